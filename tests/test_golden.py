"""Committed golden vectors (tests/golden/models.json, written by
tests/golden/make_golden.py): the checker reproduces them on the CPU, the
device reproduces them on the GPU."""
import hashlib
import json
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "models.json")))


def _digest(m):
    st, died = m.sim._read_agents(m.T)
    st = np.ascontiguousarray(st)
    rows = st.view(np.uint8).reshape(len(st), st.dtype.itemsize).copy()
    rows[np.asarray(died, bool)] = 0      # removed agents' state is not observable
    b = rows.tobytes() + np.asarray(died, np.uint8).tobytes()
    return hashlib.sha256(b).hexdigest()


def _edges(m):
    to, fr, _ = m.sim._read_edges(m.E)
    return hashlib.sha256(to.tobytes() + fr.tobytes()).hexdigest()


def _replay(V, vh):
    M = vh.models
    g = GOLD["sir"]
    m = M.build_sir(V, g["n"], mean_deg=g["mean_deg"], seed=g["seed"], n_init=g["n_init"])
    for k, want in enumerate(g["steps"]):
        M.step_sir(V, m, p_inf=g["p_inf"], seed=g["seed"])
        assert _digest(m) == want, f"sir step {k}"
    for key, f32 in (("opinion_f64", False), ("opinion_f32", True)):
        g = GOLD[key]
        m = M.build_opinion(V, g["n"], m_ba=g["m_ba"], seed=g["seed"], f32=f32)
        for k, want in enumerate(g["steps"]):
            M.step_opinion(V, m, eps=g["eps"])
            assert _digest(m) == want, f"{key} step {k}"
    g = GOLD["contact"]
    m = M.build_contact(V, g["n"], seed=g["seed"], n_init=g["n_init"])
    for k, (want, wedge) in enumerate(zip(g["steps"], g["edges"])):
        M.step_contact(V, m, seed=g["seed"], p_death=g["p_death"])
        assert _digest(m) == want, f"contact step {k}"
        assert _edges(m) == wedge, f"contact edges step {k}"


def test_hash_known_answers(vh, ora):
    for s, a, b, c, h in GOLD["hash4"]:
        assert vh._abi.lib().hash4(s, a, b, c) == h
        assert ora.lib().hash4(s, a, b, c) == h


def test_checker_reproduces_golden(vh, ora):
    _replay(ora, vh)


@pytest.mark.gpu
def test_device_reproduces_golden(vh, gpu_ready):
    _replay(vh, vh)
