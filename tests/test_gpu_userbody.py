"""Transition bodies loaded at run time (vh_register_body): bodies written
against include/vahana_device.h in vahana.jl_amd/examples/bodies.hip, built
into a code object, applied like built-in ones and compared with their C
restatements in the checker (oracle/example_bodies.c). The reference's own
test/hdf5_snapshot.jl runs here with its own transitions
(test/hdf5_common.jl:107-109, 129-134)."""
import json
import os

import numpy as np
import pytest

import scenarios as SC

pytestmark = pytest.mark.gpu


def test_user_body_field_update(vh, ora, gpu_ready):
    """state.f += delta on every agent, the states of the other fields kept;
    equal to the checker."""
    print("[gpu] run-time body: add to a Float64 field", flush=True)
    out = []
    for V in (vh, ora):
        sub, _ras = SC.reference_bodies(V)
        model, t = SC.snapshot_hint_model(V, "default")
        sim = SC.snapshot_hint_createsim(V, model, t)
        A = t["Agent"]
        before = np.asarray(sim._read_agents(A)[0]).copy()
        V.apply_(sim, sub(params=SC.AddF64Params(A.dtype.fields["f"][1], 0, -2.0)), [A], [A], [A])
        after = np.asarray(sim._read_agents(A)[0])
        assert np.array_equal(after["f"], before["f"] - 2.0)
        assert np.array_equal(after["foo"], before["foo"]) and np.array_equal(after["inner_f"], before["inner_f"])
        out.append(after.tobytes())
    assert out[0] == out[1]


@pytest.mark.parametrize("variant", ["default", "immortal", "neighbors"])
def test_hdf5_snapshot_with_reference_transitions(vh, ora, gpu_ready, tmp_path, variant):
    """test/hdf5_snapshot.jl:5-27 with test/hdf5_common.jl's transitions as
    run-time bodies: device and checker write the same snapshot records; the
    restored simulation equals the one that wrote the snapshot, before and
    after both apply remove_some_rasternodes again; the raster cells whose
    first in-neighbour has mod1(f, 8) > 4.5 (cells 5-8) die, the survivors
    keep one edge with the state doubled each time."""
    print(f"[gpu] reference snapshot test with run-time bodies ({variant})", flush=True)
    res = {}
    for V, tag in ((vh, "g"), (ora, "c")):
        bodies = SC.reference_bodies(V)
        model, t = SC.snapshot_hint_model(V, variant)
        path = str(tmp_path / f"{variant}_{tag}")
        sim = SC.reference_runsim(V, model, t, path, bodies)
        with open(os.path.join(path, "index.json")) as fh:
            idx = json.load(fh)
        assert [s[0] for s in idx["snapshots"]] == [0, 1, 2]
        assert idx["agents"]["Agent"] == [[0, 0], [1, 1]] and idx["agents"]["RasterAgent"] == [[0, 0], [2, 2]]
        assert idx["edges"]["RasterEdge"] == [[0, 0], [2, 2]] and idx["edges"]["EdgeState"] == [[0, 0]]
        st, died = sim._read_agents(t["RasterAgent"])
        assert np.asarray(st["foo"])[~died].tolist() == [1, 2, 3, 4]
        # cell c's first in-edge comes from agent c (state a = c % 8 + 1)
        assert SC.raster_edges(V, sim, t) == [(c, c, 2 * (c % 8 + 1)) for c in (1, 2, 3, 4)]
        restored = SC.snapshot_hint_restore(V, model, path)
        first = SC.snapshot_hint_state(V, sim, t)
        assert first == SC.snapshot_hint_state(V, restored, t)
        SC.reference_raster_step(V, sim, t, bodies[1])
        SC.reference_raster_step(V, restored, t, bodies[1])
        second = SC.snapshot_hint_state(V, sim, t)
        assert second == SC.snapshot_hint_state(V, restored, t)
        assert SC.raster_edges(V, restored, t) == [(c, c, 4 * (c % 8 + 1)) for c in (1, 2, 3, 4)]
        res[tag] = (first, second, idx)
    assert res["g"] == res["c"]


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
