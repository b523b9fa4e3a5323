"""Transition bodies loaded at run time (vh_register_body): bodies written
against include/vahana_device.h in vahana.jl_amd/examples/bodies.hip, built
into a code object, applied like built-in ones. The reference's own
test/hdf5_snapshot.jl runs here with its own transitions (test/hdf5_common.jl:
107-109, 129-134)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import scenarios as SC

pytestmark = pytest.mark.gpu


class AddF64(C.Structure):
    _fields_ = [("offset", C.c_uint32), ("pad", C.c_uint32), ("delta", C.c_double)]


class RasterP(C.Structure):
    _fields_ = [("f_offset", C.c_uint32), ("pad", C.c_uint32), ("modulus", C.c_double),
                ("threshold", C.c_double)]


def _bodies(vh):
    co = vh.example_bodies()
    sub = vh.register_body(co, "vh_ex_add_f64")
    ras = vh.register_body(co, "vh_ex_remove_some_rasternodes", adds_edges=True)
    return sub, ras


def test_user_body_field_update(vh, gpu_ready):
    """state.f += delta on every agent, the states of the other fields kept."""
    print("[gpu] run-time body: add to a Float64 field", flush=True)
    sub, _ras = _bodies(vh)
    model, t = SC.snapshot_hint_model(vh, "default")
    sim = SC.snapshot_hint_createsim(vh, model, t)
    A = t["Agent"]
    before = np.asarray(sim._read_agents(A)[0]).copy()
    p = AddF64(A.dtype.fields["f"][1], 0, -2.0)
    vh.apply_(sim, sub(params=p), [A], [A], [A])
    after = np.asarray(sim._read_agents(A)[0])
    assert np.array_equal(after["f"], before["f"] - 2.0)
    assert np.array_equal(after["foo"], before["foo"]) and np.array_equal(after["inner_f"], before["inner_f"])


def _runsim(vh, model, t, path, sub, ras):
    """runsim test/hdf5_common.jl:101-127 with its own transitions."""
    A, RA, RE = t["Agent"], t["RasterAgent"], t["RasterEdge"]
    sim = SC.snapshot_hint_createsim(vh, model, t)
    vh.snapshot.write_snapshot(sim, path, "Initial state")
    vh.apply_(sim, sub(params=AddF64(A.dtype.fields["f"][1], 0, -2.0)), [A], [A], [A])
    vh.push_global_(sim, "gb", 2.1)
    vh.push_global_(sim, "gc", 4)
    vh.snapshot.write_snapshot(sim, path)
    _raster_step(vh, sim, t, ras)
    vh.snapshot.write_snapshot(sim, path, "Final state")
    return sim


def _raster_step(vh, sim, t, ras):
    A, RA, RE = t["Agent"], t["RasterAgent"], t["RasterEdge"]
    p = RasterP(A.dtype.fields["f"][1], 0, 8.0, 4.5)
    vh.apply_(sim, ras(role_edges=[RE], params=p), [RA], [RA, RE, A], [RA, RE])


def _raster_edges(sim, t):
    to, fr, st = sim._read_edges(t["RasterEdge"])
    nr = lambda x: (x & np.uint64(0xFFFFFFFFF)).astype(np.int64)  # noqa: E731
    return sorted(zip(nr(to).tolist(), nr(fr).tolist(), np.asarray(st["a"]).tolist()))


@pytest.mark.parametrize("variant", ["default", "immortal", "neighbors"])
def test_hdf5_snapshot_with_reference_transitions(vh, gpu_ready, tmp_path, variant):
    """test/hdf5_snapshot.jl:5-27 with test/hdf5_common.jl's transitions as
    run-time bodies: the restored simulation equals the one that wrote the
    snapshot, before and after both apply remove_some_rasternodes again; the
    raster cells whose first in-neighbour has mod1(f, 8) > 4.5 (cells 5-8)
    die, the survivors keep one edge with the state doubled each time."""
    print(f"[gpu] reference snapshot test with run-time bodies ({variant})", flush=True)
    sub, ras = _bodies(vh)
    model, t = SC.snapshot_hint_model(vh, variant)
    path = str(tmp_path / variant)
    sim = _runsim(vh, model, t, path, sub, ras)
    RA = t["RasterAgent"]
    with open(os.path.join(path, "index.json")) as fh:
        idx = json.load(fh)
    assert [s[0] for s in idx["snapshots"]] == [0, 1, 2]
    assert idx["agents"]["Agent"] == [[0, 0], [1, 1]] and idx["agents"]["RasterAgent"] == [[0, 0], [2, 2]]
    assert idx["edges"]["RasterEdge"] == [[0, 0], [2, 2]] and idx["edges"]["EdgeState"] == [[0, 0]]
    st, died = sim._read_agents(RA)
    assert np.asarray(st["foo"])[~died].tolist() == [1, 2, 3, 4]
    # cell c's first in-edge comes from agent c (state a = c % 8 + 1)
    assert _raster_edges(sim, t) == [(c, c, 2 * (c % 8 + 1)) for c in (1, 2, 3, 4)]
    restored = SC.snapshot_hint_restore(vh, model, path)
    assert SC.snapshot_hint_state(vh, sim, t) == SC.snapshot_hint_state(vh, restored, t)
    _raster_step(vh, sim, t, ras)
    _raster_step(vh, restored, t, ras)
    assert SC.snapshot_hint_state(vh, sim, t) == SC.snapshot_hint_state(vh, restored, t)
    assert _raster_edges(restored, t) == [(c, c, 4 * (c % 8 + 1)) for c in (1, 2, 3, 4)]


def test_register_body_errors(vh, gpu_ready, tmp_path):
    """A missing file or kernel is an error; the same kernel again gives the
    same id; other flags for it are an error."""
    print("[gpu] run-time body registration errors", flush=True)
    co = vh.example_bodies()
    with pytest.raises(vh.VahanaError):
        vh.register_body(str(tmp_path / "none.co"), "vh_ex_add_f64")
    with pytest.raises(vh.VahanaError):
        vh.register_body(co, "no_such_kernel")
    a = vh.register_body(co, "vh_ex_add_f64")
    b = vh.register_body(co, "vh_ex_add_f64")
    assert a.id == b.id >= vh._abi.TR_USER
    with pytest.raises(vh.VahanaError):
        vh.register_body(co, "vh_ex_add_f64", adds_edges=True)
    # the failed loads leave no HIP error behind for later launches to report
    m = vh.models.build_sir(vh, 200_000, seed=3, n_init=2000)
    for _ in range(3):
        vh.models.step_sir(vh, m, p_inf=0.2, seed=3)
    m.sim._destroy()
