"""The checker against the reference's own known answers (CPU only).

Each expected value below is the one the reference's test file asserts
(file:line cited); the scenario code lives in scenarios.py and is shared with
the GPU parity tests.
"""
import numpy as np
import pytest

import scenarios as SC


def test_mapreduce_known_answers(ora):
    # test/core.jl:87-94: foo = 1..10 with +, *, & (Int), | (Int), max, min
    out = SC.core_aggregate(ora)
    expect = (55, 3628800, 0, 15, 10, 1)
    for T in ("AImmFixed", "AImmFixedOversize", "AMortalFixed", "ADefault"):
        assert out[T] == expect, T
    # test/core.jl:421-422: all Bool fields true
    assert out["bool"] == (True, True)
    # test/core.jl:96-107: no agent left -> the init value, no assertion
    assert out["empty_AMortalFixed"] == 0 and out["n_AMortalFixed"] == 0
    assert out["empty_ADefault"] == 0 and out["n_ADefault"] == 0


def test_synchronous_neighbor_sums(ora):
    out = SC.core_sum_neighbors(ora)
    s = sum(range(1, 11))
    assert out["AMortal"] == [s + 1, 2 * s + 1]          # test/core.jl:317-330
    assert out["AImm"] == [2, 3, 4]                       # test/core.jl:340-360
    assert out["AImmFixed"] == [3, 5, 7]                  # test/core.jl:375-392
    assert out["assert_missing_read"]                     # test/core.jl:298-301


def test_core_queries(ora):
    out = SC.core_queries(ora)
    assert out["edges_a1_ESDict"] == 4                    # test/core.jl:195
    assert out["neighbors_a1_ESLDict1"][0] == out["avids"][0]
    assert out["neighbors_a1_ESLDict1"][9] == out["avids"][9]   # test/core.jl:208-209
    assert out["edges_a2_ESDict"] is None                 # test/core.jl:198
    assert out["neighbors_avids10"] == [out["avfids"][9]]  # test/core.jl:210-211
    assert out["num_edges"] == (4, 0, 1)                  # test/core.jl:281-283
    assert out["state_a2"] == 2 and out["state_avfids10"] == 10
    assert out["assert_wrong_type"]                       # test/core.jl:167
    assert out["all_agents_fixed"] == [1, 2, 3, 4, 5]     # test/core.jl:229-240
    assert out["num_fixed"] == 5 and out["num_fixed_orig"] == 10
    assert out["with_edge"] == (99, 99)                   # test/core.jl:143-155
    assert out["with_edge_dead"]
    assert out["all_agentids_even"] == 5                  # test/core.jl:256-259


def test_edge_hint_matrix(ora):
    out = SC.edges_matrix(ora)
    a1, a2, a3 = out["ids"]
    res = out["res"]
    # test/edges.jl:85-97: a second, different edge to a SingleEdge target asserts
    assert out["second_add_asserts"] == {"EdgeSE": True, "EdgeE": True, "EdgeEI": True}
    for t in ["EdgeD", "EdgeT", "EdgeTs"]:                # test/edges.jl:126-133
        assert res[t]["edges"][2] == [(a2, 1), (a1, 2)]
        assert res[t]["edges"][1] is None
    assert res["EdgeE"]["edges"][2] == [(a2, 1)]          # test/edges.jl:134-137
    for t in ["EdgeI", "EdgeEI", "EdgeTI", "EdgeTsI", "EdgeS", "EdgeSE", "EdgeST", "EdgeSI",
              "EdgeSEI", "EdgeSTI", "EdgeSETI", "EdgeSTs", "EdgeSTsI", "EdgeSETsI"]:
        assert res[t]["edges"][0] == "assert", t          # test/edges.jl:140-142
    for t in ["EdgeD", "EdgeT", "EdgeTs", "EdgeS", "EdgeST", "EdgeSTs"]:   # :145-152
        assert res[t]["neighborids"][2] == [a2, a1]
        assert res[t]["neighborids"][1] is None
    for t in ["EdgeE", "EdgeSE"]:
        assert res[t]["neighborids"][2] == [a2]
    for t in ["EdgeI", "EdgeTI", "EdgeTsI", "EdgeSI", "EdgeSTI", "EdgeSTsI", "EdgeEI", "EdgeSEI",
              "EdgeSETI", "EdgeSETsI"]:
        assert res[t]["neighborids"][0] == "assert", t
    for t in ["EdgeD", "EdgeT", "EdgeTs", "EdgeI", "EdgeTI", "EdgeTsI"]:   # :166-181
        assert res[t]["edgestates"][2] == [1, 2]
        assert res[t]["edgestates"][1] is None
    for t in ["EdgeE", "EdgeEI"]:
        assert res[t]["edgestates"][2] == [1]
    for t in ["EdgeS", "EdgeST", "EdgeSTs", "EdgeSI", "EdgeSTI", "EdgeSTsI", "EdgeSE", "EdgeSEI",
              "EdgeSETI", "EdgeSETsI"]:
        assert res[t]["edgestates"][0] == "assert", t
    for t in ["EdgeD", "EdgeT", "EdgeTs", "EdgeI", "EdgeTI", "EdgeTsI", "EdgeS", "EdgeST",
              "EdgeSTs", "EdgeSTI", "EdgeSTsI"]:            # :193-199
        assert res[t]["num_edges"] == [1, 0, 2], t
    for t in ["EdgeE", "EdgeEI", "EdgeSE", "EdgeSETI", "EdgeSETsI"]:
        assert res[t]["num_edges"] == "assert", t
    for t in SC.EDGE_HINTS:                                 # :204-212
        assert res[t]["has_edge"] == [True, False, True], t
    for t in ["EdgeD", "EdgeT", "EdgeI", "EdgeTI", "EdgeTs", "EdgeTsI"]:   # :215-220
        assert res[t]["mapreduce"] == 6, t
    for t in ["EdgeE", "EdgeEI"]:
        assert res[t]["mapreduce"] == 4, t
    for t in ["EdgeS", "EdgeSE", "EdgeST", "EdgeSI", "EdgeSEI", "EdgeSTI", "EdgeSETI", "EdgeSTs",
              "EdgeSTsI", "EdgeSETsI"]:
        assert res[t]["mapreduce"] == "assert", t


def test_edges_rebuild_vs_add_existing(ora):
    out = SC.edges_transition(ora)
    n1 = 3                                                  # complete graph K4: n - 1 in-edges
    for t in ["EdgeD", "EdgeT", "EdgeI", "EdgeTI"]:        # test/edges.jl:336-349
        assert out[t][0] == [n1] * 4 and out[t][1] == [n1] * 4, t
    assert out["EdgeD"][2] == [n1] * 4                      # :352-356
    assert out["EdgeD"][3] == [2 * n1] * 4                  # :362-370
    assert out["EdgeD"][4] == [0] * 4                       # :373-378


def test_remove_agents(ora):
    out = SC.remove_agents(ora)
    for E in ("DEdge", "DEdgeST"):                          # test/remove_agents.jl:37-63
        assert out[(E, 1)] == [(4, 1), (4, 0), (3, 0)], E
        assert out[(E, 2)] == [(8, 2), (8, 0), (6, 0)], E   # :82-112


def test_independent_equivalence(ora):
    res = SC.independent(ora)                               # test/independent.jl:77-83,193-198
    assert res["indep"] == res["not"] == res["immortal"]
    # first step: agent foo=1 dies, 9 agents each add 2*9 edges (the 2*9 to or
    # from the dead one are removed with it)
    assert res["not"][0] == (10, 100)
    assert res["not"][1][0] == 9


def test_addexisting(ora):
    out = SC.addexisting(ora)                               # test/addexisting.jl:20-63
    for v in ("plain", "imm_comp"):
        assert out[v] == [(1, 1, 1), (1, 0, 0), (1, 1, 1), (1, 1, 1)], v
    assert out["assert"]                                    # :70-81


def test_graphs_k4(ora):
    assert SC.graphs_k4(ora) == sum(range(1, 5)) * 3        # test/graphs.jl:42-43


def test_edges_iterator_and_removal(ora):
    out = SC.edges_iter_removal(ora)
    for name, (before, after) in out["iter"].items():         # test/edgesiterator.jl:33-43
        expected = 10 if SC.hashint(name, "E") else 20
        assert before == after == expected, name
    for name, r in out["agg"].items():                        # :66-101
        assert r == [55, 0, 55, 0], name
    for name, r in out["remove"].items():                     # :131-210
        assert r[:2] == [10, 0], name
        hints = SC.EDGE_HINTS[name]
        if "IgnoreFrom" in hints:
            assert len(r) == 2
            continue
        count = 5
        if "SingleEdge" not in hints:
            count += 10
        if "SingleType" not in hints:
            count += 5
        if "SingleEdge" not in hints and "SingleType" not in hints:
            count += 5
        assert r[2] == count, (name, r)


def test_remove_edges_in_transitions(ora):
    """test/mpi/test_edgetypes.jl:295-445 at one process: remove_edges!(to)
    leaves 50 of the 100 cycle edges (own or the neighbour's); remove then
    add_edge! of the same edge keeps all 100; remove_edges!(from, to) takes
    2 x 100 of 9900 (or the single edge), 50 of the other direction (or 99 of
    the cycle via SendId); an unknown source removes nothing; without the
    edge type in `write`, or in `add_existing`, remove_edges! asserts."""
    r = SC.remove_edges(ora)
    for name in SC.REMOVE_TYPES:
        want = {"own": 50}
        if "IgnoreFrom" not in SC.MPI_EDGES[name]:
            want.update(neighbor=50, readd=100)
        assert r["remove/" + name] == want, name
    for name in SC.REMOVE_FROM_TYPES:
        single = "SingleEdge" in SC.MPI_EDGES[name]
        want = {"init": 100 if single else 9900, "from_two": 0 if single else 9700,
                "sim2": 1 if single else 9850, "zero_from": 100 if single else 9900,
                "assert_not_written": "AssertionError", "assert_not_add_existing": "AssertionError"}
        assert r["remove_from/" + name] == want, name


def test_raster_diffusion(ora):
    out = SC.raster_diffusion(ora)
    for key in ("2d_10x8", "2d_7x9", "2d_17x6", "2d_6x7"):   # test/raster.jl:54-58, 81-85, ...
        assert out[key] == [1, 9, 25], key
    assert out["rasterstate"] == (False, True, False)         # :60-63
    assert out["raster"] == (False, True, False)              # :65-68
    assert out["pos_of_3_5"] == (3, 5)
    assert out["3d"] == [1, 3 * 3 * 3, 5 * 5 * 5]              # :162-166
    assert out["3d_nonperiodic"] == [1, 2 * 2 * 2, 3 * 3 * 3]  # :182-186
    assert out["3d_dist2"] == [1, 3 * 3 * 3, 5 * 5 * 5]        # :203-207
    assert out["3d_euclidean"] == [1, 3 * 3 * 3 - 8]          # :222-224
    assert out["3d_manhatten"] == [1, 7]                      # :240-242
    assert out["3d_manhatten_d2"] == [1, 10]                  # :258-260


def test_raster_nodeid(ora):
    out = SC.raster_nodeid(ora)
    assert out["first"] == (1, 0, 5, 0)                       # test/raster.jl:301-305
    assert out["first_values"] == (1, 0, 5, 0)
    assert out["second"] == (1, 0, 0, 10)                     # :327-331
    assert out["values"] == [1, 5, 5]


def test_raster_move_to_dist(ora):
    out = SC.raster_move_to_dist(ora)
    assert out["2d"] == [25, 13, 21]                          # test/raster.jl:374-376
    assert out["4d"] == [3 * 3 * 3 * 3, 1 + 4 * 2]            # :401-402


def test_raster_random(ora):
    pos, cellpos = SC.raster_random(ora)                      # test/raster.jl:427-433
    assert pos == (7, 19, 23) and cellpos == (7, 19, 23)


def test_raster_stencils(vh):
    """Stencil sizes from their closed forms: the Chebyshev cube (2d+1)^n - 1;
    the Manhattan ball: 2n points at d = 1, 12 at d = 2 in 2D; the Euclidean
    ball of radius 1.5 in 3D: the 3x3x3 cube without the origin and the 8
    corners at sqrt(3); radius 2.5 in 2D: the 5x5 square without the origin
    and the 4 corners at sqrt(8). Offsets run first component fastest
    (src/Raster.jl:80-96)."""
    st = vh.raster.stencil
    for n in (1, 2, 3, 4):
        for d in (1, 2):
            assert len(st("chebyshev", n, d)) == (2 * d + 1) ** n - 1
        assert len(st("manhatten", n, 1)) == 2 * n
    assert len(st("manhatten", 2, 2)) == 12
    assert len(st("euclidean", 3, 1.5)) == 26 - 8
    assert len(st("euclidean", 2, 2.5)) == 20               # 5x5 minus origin minus 4 corners
    assert st("chebyshev", 2, 1)[:3] == [(-1, -1), (0, -1), (1, -1)]


def test_snapshot_resume_checker(ora, tmp_path):
    out = SC.snapshot_sir_resume(ora, str(tmp_path / "sir"))
    # filed under num_transitions - 1 = 5 (src/HDF5.jl:119); the counter
    # resumes at 1 + five apply!
    assert out["t"] == out["t_read"] == 5 and out["resumed"] == 6
    assert out["same"]
    assert out["list"] == [(5, "after five")]


def test_snapshot_contact_restore_checker(ora, tmp_path):
    out = SC.snapshot_contact_restore(ora, str(tmp_path / "contact"))
    assert out["died"] and out["states"] and out["edges"]
    b = out["b"]
    assert out["free"] and b.sim.reuse_stack(b.T) == out["free"]   # ascending
    # one step on: the free slots are exactly the died ones (a restored died
    # slot that no birth took stays died)
    ora.models.step_contact(ora, b, seed=8, p_death=0.05)
    _, died = b.sim._read_agents(b.T)
    assert sorted(b.sim.reuse_stack(b.T)) == [int(i) + 1 for i in np.nonzero(died)[0]]


def test_snapshot_raster_ids(ora, tmp_path):
    model, GridA, Grid3D, GridE, *_ = SC.raster_model(ora)
    sim = ora.create_simulation(model)
    ora.raster.add_raster_(sim, "grid", (4, 3), lambda p: GridA(p, p == (2, 2)))
    ora.raster.connect_raster_neighbors_(sim, "grid", lambda a, b: GridE())
    ora.finish_init_(sim)
    ora.snapshot.write_snapshot(sim, str(tmp_path / "r"))
    sim2 = ora.create_simulation(model)
    ora.snapshot.read_snapshot(sim2, str(tmp_path / "r"))
    assert np.array_equal(sim2.rasters["grid"].ids, sim.rasters["grid"].ids)
    assert ora.raster.cellid(sim2, "grid", (2, 2)) == ora.raster.cellid(sim, "grid", (2, 2))
    ora.apply_(sim2, ora.transitions.diffuse(GridE, GridA), [GridA], [GridA, GridE], [GridA])
    assert ora.mapreduce(sim2, "active", "+", GridA) == 9


def test_model_destroyed_before_simulation(ora):
    """A simulation keeps its own copy of the model's types: destroying the
    model handle first (as a garbage collector may, for objects in a
    reference cycle) leaves the simulation usable and destroyable."""
    model, SIR, Contact = SC.sir_model(ora)
    sim = ora.create_simulation(model)
    ids = sim._add_agents_array(SIR, np.zeros(3, dtype=SIR.dtype))
    model.lib.model_destroy(model.handle)
    model.handle = type(model.handle)()      # Model.__del__ then skips it
    sim._add_edges_array(Contact, [ids[0]], [ids[1]])
    ora.finish_init_(sim)
    assert ora.num_edges(sim, Contact) == 1
    sim._destroy()
