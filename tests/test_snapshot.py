"""Snapshots per edge hint with last_change skipping (src/HDF5.jl), on the
checker: the three models of test/hdf5_snapshot.jl, restored and advanced
side by side with the run that wrote them."""
import json
import os

import numpy as np
import pytest

import scenarios as SC

VARIANTS = ["default", "immortal", "neighbors"]


def _index(path):
    with open(os.path.join(path, "index.json")) as fh:
        return json.load(fh)


@pytest.mark.parametrize("variant", VARIANTS)
def test_write_restore(ora, tmp_path, variant):
    """test_write_restore test/hdf5_snapshot.jl:5-27: the restored
    simulation equals the one that wrote the snapshot (states, died flags,
    slots, edges, last_change of every type, params, globals_last_change),
    and again after both run the raster transitions."""
    model, t = SC.snapshot_hint_model(ora, variant)
    path = str(tmp_path / variant)
    sim = SC.snapshot_hint_runsim(ora, model, t, path)
    restored = SC.snapshot_hint_restore(ora, model, path)
    assert SC.snapshot_hint_state(ora, sim, t) == SC.snapshot_hint_state(ora, restored, t)
    assert sim.reuse_stack(t["RasterAgent"]) == restored.reuse_stack(t["RasterAgent"])
    SC.snapshot_hint_raster_step(ora, sim, t)
    SC.snapshot_hint_raster_step(ora, restored, t)
    assert SC.snapshot_hint_state(ora, sim, t) == SC.snapshot_hint_state(ora, restored, t)


@pytest.mark.parametrize("variant", VARIANTS)
def test_unchanged_types_skipped(ora, tmp_path, variant):
    """A type is written only when its last_change moved since its previous
    record (src/HDF5.jl:496-501, 592-597, 374-378): Agent changes in the
    first transition, the raster types in the last two (RasterEdge by the
    re-add and then by the deaths of its targets), the edge types of the
    Agents never, the globals by push_global!."""
    model, t = SC.snapshot_hint_model(ora, variant)
    path = str(tmp_path / variant)
    SC.snapshot_hint_runsim(ora, model, t, path)
    idx = _index(path)
    assert idx["snapshots"] == [[0, "Initial state", []], [1, "", []], [3, "Final state", []]]
    assert idx["agents"] == {"Agent": [[0, 0], [1, 1]], "RasterAgent": [[0, 0], [3, 3]],
                             "EmptyAgentVector": [[0, 0]]}
    assert idx["edges"] == {"EdgeState": [[0, 0]], "RasterEdge": [[0, 0], [3, 3]],
                            "StatelessEdge": [[0, 0]], "EmptyEdgeVector": [[0, 0]]}
    assert idx["globals"] == [[0, 0], [1, 1]]
    assert sorted(os.listdir(os.path.join(path, "agents", "Agent"))) == ["t_0.npz", "t_1.npz"]
    assert os.listdir(os.path.join(path, "edges", "EdgeState")) == ["t_0.npz"]


def test_death_stamps_edge_last_change(ora):
    """An edge type that loses edges to deaths gets the transition's number
    as last_change, though the transition does not write it
    (src/AgentMethods.jl:347-353)."""
    model, t = SC.snapshot_hint_model(ora, "default")
    sim = SC.snapshot_hint_createsim(ora, model, t)
    RA, RE = t["RasterAgent"], t["RasterEdge"]
    lc = ora.snapshot.last_change
    assert lc(sim, RE) == 0 and lc(sim, RA) == 0
    ora.apply_(sim, ora.transitions.die_if("never"), [RA], [RA], [RA])     # no deaths
    assert lc(sim, RA) == 1 and lc(sim, RE) == 0
    ora.apply_(sim, ora.transitions.die_if("odd"), [RA], [RA], [RA])
    assert lc(sim, RA) == 2 and lc(sim, RE) == 2
    assert lc(sim, t["EdgeState"]) == 0                  # no edge of it touched a dead agent


RECORD_KINDS = {
    "default": {"EdgeState": "complete", "StatelessEdge": "stateless", "RasterEdge": "complete",
                "EmptyEdgeVector": "stateless"},
    "immortal": {"EdgeState": "ignore_from", "StatelessEdge": "stateless", "RasterEdge": "complete",
                 "EmptyEdgeVector": "stateless"},
    "neighbors": {"EdgeState": "edge_count", "StatelessEdge": "count_vector",
                  "RasterEdge": "complete", "EmptyEdgeVector": "stateless"},
}


@pytest.mark.parametrize("variant", VARIANTS)
def test_record_kinds(ora, tmp_path, variant):
    """The record written per edge type follows its hints
    (src/HDF5.jl:543-575, 610-632): CompleteEdge, StatelessEdge (also for
    types without fields), IgnoreFromEdge, EdgeCount for neighbours-only
    types, the raw count vector for neighbours-only :SingleType types
    (Bool with :SingleEdge / :HasEdgeOnly)."""
    model, t = SC.snapshot_hint_model(ora, variant)
    path = str(tmp_path / variant)
    SC.snapshot_hint_runsim(ora, model, t, path)
    for name, kind in RECORD_KINDS[variant].items():
        with np.load(os.path.join(path, "edges", name, "t_0.npz")) as z:
            meta = json.loads(bytes(z["meta"]).decode())
            assert meta["kind"] == kind, name
            files = set(z.files)
            if kind == "edge_count":
                # 16 self edges + 16 to agent 2: agent 2 counts 17
                cnt = dict(zip(z["to"].tolist(), z["count"].tolist()))
                assert sorted(cnt.values()) == [1] * 15 + [17]
            if kind == "count_vector":
                assert z["count"].dtype == np.bool_ and z["count"].tolist() == [False, True, False, False, True]
            if kind == "ignore_from":
                assert files == {"meta", "to", "state"}
            if kind == "stateless":
                assert files == {"meta", "to", "from"}


def test_read_earlier_snapshot(ora, tmp_path):
    """read_snapshot!(; comment) / (; transition): the latest snapshot at or
    before the transition; every type from its latest record <= it."""
    model, t = SC.snapshot_hint_model(ora, "default")
    path = str(tmp_path / "d")
    SC.snapshot_hint_runsim(ora, model, t, path)
    fresh = SC.snapshot_hint_createsim(ora, model, t)
    first = SC.snapshot_hint_restore(ora, model, path, comment="Initial state")
    assert SC.snapshot_hint_state(ora, fresh, t) == SC.snapshot_hint_state(ora, first, t)
    # transition 2 -> snapshot 1: Agent from its record 1, the rest from 0
    mid = SC.snapshot_hint_restore(ora, model, path, transition=2)
    s = SC.snapshot_hint_state(ora, mid, t)
    assert s["num_transitions"] == 2 and s["globals_last_change"] == 1
    assert s["Agent"][3] == 1 and s["RasterAgent"][3] == 0 and s["RasterEdge"][3] == 0
    assert mid.globals["gb"] == [4.0, 4.1, 2.1] and mid.globals["gc"] == [5, 6, 7, 4]
    assert ora.snapshot.list_snapshots(path) == [(0, "Initial state"), (1, ""), (3, "Final state")]
    with pytest.raises(FileNotFoundError):
        SC.snapshot_hint_restore(ora, model, path, comment="no such comment")


def test_file_readers(ora, tmp_path):
    """read_agents / read_edges / read_params / read_globals on the files
    (src/HDF5.jl:1113-1160, 1292-1330, 938-1090)."""
    model, t = SC.snapshot_hint_model(ora, "default")
    path = str(tmp_path / "d")
    sim = SC.snapshot_hint_runsim(ora, model, t, path)
    ra = ora.snapshot.read_agents(path, "RasterAgent")
    assert ra["foo"].tolist() == [2, 4, 6, 8]                  # the even cells live
    a0 = ora.snapshot.read_agents(path, t["Agent"], comment="Initial state")
    assert a0["foo"].tolist() == list(range(2, 18)) and a0["f"].tolist() == [float(i + 2) for i in range(1, 17)]
    es = ora.snapshot.read_edges(path, "EdgeState")
    assert len(es) == 32 and set(es.dtype.names) == {"to", "from", "f", "t"}
    assert (es["f"] == es["from"]).all()
    with pytest.raises(AssertionError):
        ora.snapshot.read_edges(path, "StatelessEdge")
    assert ora.snapshot.read_params(path) == {"pb": 1, "pa": [2.0, 2.1], "pos": {"x": 3, "y": 4, "z": 5}}
    assert ora.snapshot.read_globals(path, transition=0)["gb"] == [4.0, 4.1]
    assert ora.snapshot.read_globals(path)["gb"] == [4.0, 4.1, 2.1]
    assert sim.params["pb"] == 1


def test_writeable_restore_appends(ora, tmp_path):
    """read_snapshot!(; writeable = true): later snapshots of the restored
    simulation go into the same files."""
    model, t = SC.snapshot_hint_model(ora, "default")
    path = str(tmp_path / "d")
    SC.snapshot_hint_runsim(ora, model, t, path)
    r = SC.snapshot_hint_restore(ora, model, path, writeable=True)
    SC.snapshot_hint_raster_step(ora, r, t)
    assert ora.snapshot.write_snapshot(r, path, "resumed") == 5
    assert [s[0] for s in ora.snapshot.list_snapshots(path)] == [0, 1, 3, 5]
    again = SC.snapshot_hint_restore(ora, model, path)
    a, b = SC.snapshot_hint_state(ora, r, t), SC.snapshot_hint_state(ora, again, t)
    # the deaths of the last transition removed nothing (the odd cells were
    # dead already): RasterEdge's last_change is 4, its record is filed under
    # 5, and a restore sets last_change to the record's number
    # (src/HDF5.jl:1434-1436)
    assert a["RasterEdge"][3] == 4 and b["RasterEdge"][3] == 5
    assert a["RasterEdge"][:3] == b["RasterEdge"][:3]
    a.pop("RasterEdge"), b.pop("RasterEdge")
    assert a == b


def test_param_values_restored(ora, tmp_path):
    """The restored scalar params/globals are the values transition bodies
    read (the library's copies), unless ignore_params."""
    model, t = SC.snapshot_hint_model(ora, "default")
    path = str(tmp_path / "d")
    sim = SC.snapshot_hint_runsim(ora, model, t, path)
    ora.set_global_(sim, "ga", 11)
    ora.snapshot.write_snapshot(sim, path)
    r = SC.snapshot_hint_restore(ora, model, path)
    assert r._value(ora._abi.VALUE_PARAM, "pb") == 1 and r._value(ora._abi.VALUE_GLOBAL, "ga") == 11
    assert r.globals_last_change == 3
    r2 = SC.snapshot_hint_restore(ora, model, path, ignore_params=True)
    assert r2.params["pb"] == 0 and r2._value(ora._abi.VALUE_PARAM, "pb") == 0


@pytest.mark.parametrize("variant", VARIANTS)
def test_reference_transitions_checker(ora, tmp_path, variant):
    """test/hdf5_snapshot.jl with test/hdf5_common.jl's own transitions (the
    checker's restatements in oracle/example_bodies.c): the known answers of
    remove_some_rasternodes (cells 5-8 die, the survivors keep one edge with
    the state doubled), the records it writes, and restore = original."""
    bodies = SC.reference_bodies(ora)
    model, t = SC.snapshot_hint_model(ora, variant)
    path = str(tmp_path / variant)
    sim = SC.reference_runsim(ora, model, t, path, bodies)
    st, died = sim._read_agents(t["RasterAgent"])
    assert np.asarray(st["foo"])[~died].tolist() == [1, 2, 3, 4]
    assert SC.raster_edges(ora, sim, t) == [(c, c, 2 * (c % 8 + 1)) for c in (1, 2, 3, 4)]
    assert np.asarray(sim._read_agents(t["Agent"])[0]["f"]).tolist() == [float(i) for i in range(1, 17)]
    restored = SC.snapshot_hint_restore(ora, model, path)
    assert SC.snapshot_hint_state(ora, sim, t) == SC.snapshot_hint_state(ora, restored, t)
    SC.reference_raster_step(ora, sim, t, bodies[1])
    SC.reference_raster_step(ora, restored, t, bodies[1])
    assert SC.snapshot_hint_state(ora, sim, t) == SC.snapshot_hint_state(ora, restored, t)
    assert SC.raster_edges(ora, restored, t) == [(c, c, 4 * (c % 8 + 1)) for c in (1, 2, 3, 4)]
