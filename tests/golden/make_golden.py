"""Writes tests/golden/models.json: per-step SHA-256 digests of the checker's
state for small runs of the three benchmark models, and counter-hash known
answers.  Run from the repository root: python tests/golden/make_golden.py"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import oracle_api as ora  # noqa: E402

vh = ora.vh
M = vh.models


def digest(m):
    st, died = m.sim._read_agents(m.T)
    st = np.ascontiguousarray(st)
    rows = st.view(np.uint8).reshape(len(st), st.dtype.itemsize).copy()
    rows[np.asarray(died, bool)] = 0      # removed agents' state is not observable
    b = rows.tobytes() + np.asarray(died, np.uint8).tobytes()
    return hashlib.sha256(b).hexdigest()


def edges_digest(m):
    to, fr, _ = m.sim._read_edges(m.E)
    return hashlib.sha256(to.tobytes() + fr.tobytes()).hexdigest()


def runs():
    out = {}
    n = 2000
    m = M.build_sir(ora, n, mean_deg=8.0, seed=31, n_init=40)
    seq = []
    for _ in range(20):
        M.step_sir(ora, m, p_inf=0.12, seed=31)
        seq.append(digest(m))
    out["sir"] = {"n": n, "mean_deg": 8.0, "seed": 31, "n_init": 40, "p_inf": 0.12, "steps": seq}
    for f32 in (False, True):
        m = M.build_opinion(ora, n, seed=32, f32=f32)
        seq = []
        for _ in range(10):
            M.step_opinion(ora, m, eps=0.2)
            seq.append(digest(m))
        out["opinion_f32" if f32 else "opinion_f64"] = {"n": n, "m_ba": 5, "seed": 32,
                                                        "eps": 0.2, "steps": seq}
    m = M.build_contact(ora, n, seed=33, n_init=20)
    seq, eseq = [], []
    for _ in range(10):
        M.step_contact(ora, m, seed=33, p_death=0.04)
        seq.append(digest(m))
        eseq.append(edges_digest(m))
    out["contact"] = {"n": n, "seed": 33, "n_init": 20, "p_death": 0.04, "steps": seq,
                      "edges": eseq}
    L = ora.lib()
    out["hash4"] = [[s, a, b, c, L.hash4(s, a, b, c)]
                    for (s, a, b, c) in [(0, 0, 0, 0), (1, 2, 3, 4), (1234567, 1 << 56 | 5, 7, 1 << 40),
                                         (2**63, 2**64 - 1, 0, 99)]]
    return out


if __name__ == "__main__":
    data = runs()
    with open(os.path.join(HERE, "models.json"), "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", os.path.join(HERE, "models.json"))
