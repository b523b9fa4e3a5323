"""Ensembles of SIR simulations advanced by one segmented launch per step.

A parameter sweep of the SIR model: every member is the same simulation as
`models.build_sir(..., edges=inputs.graph_gnp(n, mean_deg, graph_seed),
init="frac")` stepped with `models.step_sir`, and must match it bit for bit.
"""
import ctypes as C

import numpy as np

from . import _abi
from .inputs import SIR_DTYPE


class Ensemble:
    def __init__(self):
        self.lib = _abi.lib()
        self.handle = C.c_void_p()
        self._check(self.lib.ensemble_create(C.byref(self.handle)))
        self.members = {}

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"[{rc}] {self.lib.last_error().decode(errors='replace')}")

    def add_sir(self, n, mean_deg, p_inf, seed, graph_seed, recovery=7, n_init=10):
        m = _abi.SirMember(n=int(n), mean_deg=float(mean_deg), p_inf=float(p_inf), seed=int(seed),
                           graph_seed=int(graph_seed), recovery=int(recovery), n_init=int(n_init))
        mid = C.c_uint32()
        self._check(self.lib.ensemble_add_sir(self.handle, C.byref(m), C.byref(mid)))
        self.members[mid.value] = m
        return mid.value

    def remove(self, mid):
        self._check(self.lib.ensemble_remove(self.handle, mid))
        self.members.pop(mid, None)

    def set_horizon(self, mid, last_step):
        """The member applies steps up to `last_step` (its step numbers start at 1)."""
        self._check(self.lib.ensemble_set_horizon(self.handle, mid, int(last_step)))

    def set_config(self, block, xcd_remap):
        self._check(self.lib.ensemble_set_config(self.handle, block, 1 if xcd_remap else 0))

    def step(self, steps=1, use_graph=True):
        self._check(self.lib.ensemble_step(self.handle, steps, 1 if use_graph else 0))

    def sync(self):
        self._check(self.lib.ensemble_sync(self.handle))

    def set_streams(self, step_stream, aux_stream):
        """Run step launches on `step_stream` and member builds / reads on
        `aux_stream` (raw hipStream_t handles; both None: own streams)."""
        self._check(self.lib.ensemble_set_streams(self.handle, step_stream, aux_stream))

    def autotune(self, trial_steps=3):
        """(block, xcd_remap, ms per step); the chosen agents per thread of
        the size classes <1e4, <1e5, >=1e5 agents are in self.class_apt."""
        b, r, ms = C.c_uint32(), C.c_uint32(), C.c_float()
        apt = (C.c_uint32 * 3)()
        self._check(self.lib.ensemble_autotune(self.handle, trial_steps, C.byref(b), C.byref(r), apt,
                                               C.byref(ms)))
        self.class_apt = [apt[i] for i in range(3)]
        return b.value, bool(r.value), ms.value

    def set_tiles(self, apt):
        """Agents per thread (1, 2, 4) for the size classes <1e4, <1e5, >=1e5."""
        arr = (C.c_uint32 * 3)(*[int(v) for v in apt])
        self._check(self.lib.ensemble_set_tiles(self.handle, arr))

    def read(self, mid):
        n = self.members[mid].n
        out = np.zeros(n, dtype=SIR_DTYPE)
        nn, step = C.c_uint64(), C.c_uint64()
        self._check(self.lib.ensemble_read(self.handle, mid, out.ctypes.data_as(C.c_void_p), n,
                                           C.byref(nn), C.byref(step)))
        return out, step.value

    def info(self):
        a, b, c, d = (C.c_uint64() for _ in range(4))
        self._check(self.lib.ensemble_info(self.handle, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return {"members": a.value, "agents": b.value, "edges": c.value, "bytes": d.value}

    def close(self):
        if self.handle:
            self.lib.ensemble_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
