"""Init-phase edge checks of the C ABI on a simulation (GPU: a simulation
owns a HIP stream from vh_sim_create on). The checker runs the same calls."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _single_edge_sim(V, stateless=False):
    A = V.struct("A", [("v", "<i8")])
    E = V.struct("E") if stateless else V.struct("E", [("w", "<i8")])
    hints = ("SingleEdge", "Stateless") if stateless else ("SingleEdge",)
    model = (V.ModelTypes() | V.register_agenttype_(A) | V.register_edgetype_(E, *hints)
             | V.create_model("single_edge_bulk"))
    sim = V.create_simulation(model)
    ids = sim._add_agents_array(A, np.zeros(4, dtype=A.dtype))
    return sim, A, E, ids


def _st(E, vals):
    s = np.zeros(len(vals), dtype=E.dtype)
    s["w"] = vals
    return s


def _bulk_identical(V):
    sim, A, E, ids = _single_edge_sim(V)
    sim._add_edges_array(E, [ids[0], ids[0]], [ids[1], ids[1]], _st(E, [5, 5]))
    sim._add_edges_array(E, [ids[0]], [ids[1]], _st(E, [5]))
    sim._add_edges_array(E, [ids[2], ids[3], ids[2]], [ids[3], ids[2], ids[3]], _st(E, [1, 2, 1]))
    V.finish_init_(sim)
    return [repr(V.edges(sim, i, E)) for i in ids]


def test_single_edge_bulk_identical_accepted(vh, ora, gpu_ready):
    """_can_add (src/EdgeMethods.jl:257-294): equal edges to one :SingleEdge
    target are accepted, inside one bulk call and across calls."""
    print("[gpu] single-edge bulk, identical edges", flush=True)
    got, want = _bulk_identical(vh), _bulk_identical(ora)
    assert got == want
    assert got[0] == "None" and "0x100000000000001" in got[1]


@pytest.mark.parametrize("case", ["source", "state", "across_calls", "later_in_call"])
def test_single_edge_bulk_different_asserts(vh, ora, gpu_ready, case):
    """A second, different edge to the same :SingleEdge target asserts,
    whether its first edge came in the same bulk call or an earlier one."""
    print(f"[gpu] single-edge bulk, different edges ({case})", flush=True)
    for V in (vh, ora):
        sim, A, E, ids = _single_edge_sim(V)
        with pytest.raises(AssertionError):
            if case == "source":
                sim._add_edges_array(E, [ids[0], ids[2]], [ids[1], ids[1]], _st(E, [5, 5]))
            elif case == "state":
                sim._add_edges_array(E, [ids[0], ids[0]], [ids[1], ids[1]], _st(E, [5, 6]))
            elif case == "across_calls":
                sim._add_edges_array(E, [ids[0], ids[3]], [ids[1], ids[2]], _st(E, [5, 5]))
                sim._add_edges_array(E, [ids[0], ids[0]], [ids[2], ids[1]], _st(E, [7, 7]))
            else:   # first edge of the target in this call, after other entries
                sim._add_edges_array(E, [ids[0]], [ids[3]], _st(E, [1]))
                sim._add_edges_array(E, [ids[0], ids[1], ids[2], ids[1]],
                                     [ids[3], ids[2], ids[0], ids[2]], _st(E, [1, 4, 4, 8]))


def test_stateful_edges_need_states(vh, gpu_ready):
    print("[gpu] stateful edge type without states", flush=True)
    sim, A, E, ids = _single_edge_sim(vh)
    with pytest.raises(vh.VahanaError):
        sim._add_edges_array(E, [ids[0]], [ids[1]], None)
    # a :Stateless type needs none (equal sources: accepted)
    sim2, A2, E2, ids2 = _single_edge_sim(vh, stateless=True)
    sim2._add_edges_array(E2, [ids2[0], ids2[0]], [ids2[1], ids2[1]], None)
    vh.finish_init_(sim2)
    assert vh.neighborids(sim2, ids2[1], E2) == ids2[0]


def test_bound_values_on_device(vh, ora, gpu_ready):
    """SIR reading p_inf and seed from parameters and the recovery days from
    a global changed mid-run (vh_value_bind): equal to the literal run on the
    device, step by step, and to the checker's bound run."""
    import test_globals as TG
    print("[gpu] bound params/globals, SIR", flush=True)
    got = TG.bound_sir_matches_literal(vh)
    want = TG.bound_sir_matches_literal(ora)
    assert np.array_equal(got, want)


def test_batch_apply_opinion_matches_solo(vh, ora, gpu_ready, opts):
    """vh_batch_apply: opinion simulations of several sizes and step counts
    advanced together (one fused launch per step for the one-pass ones; one
    of them large enough for the two-pass gather, which runs its own
    kernels): every state equals the same simulation stepped alone, and one
    the checker's."""
    print("[gpu] batched opinion apply vs solo", flush=True)
    opts(opinion_plan_bytes=200_000 * 8)   # the 300k-agent simulation uses the plan
    M = vh.models
    specs = [(3000, 8.0, 5), (20000, 12.0, 9), (777, 5.0, 7), (300_000, 10.0, 6), (64, 4.0, 9)]
    lib = vh._abi.lib()
    st = C.c_void_p()
    assert lib.stream_create(C.byref(st)) == 0
    sims = []
    for k, (n, md, steps) in enumerate(specs):
        m = M.build_opinion_gnp(vh, n, mean_deg=md, seed=50 + k, graph_seed=60 + k, weight_seed=70 + k)
        m.sim._set_streams(st, st)
        sims.append(m)
    B = vh.Batch()
    B.apply([m.sim for m in sims], vh.transitions.opinion(sims[0].E, 0.2), sims[0].T, [sims[0].T, sims[0].E],
            sims[0].T, steps=[s[2] for s in specs])
    lib.sync(sims[0].sim.handle)
    for k, ((n, md, steps), m) in enumerate(zip(specs, sims)):
        solo = M.build_opinion_gnp(vh, n, mean_deg=md, seed=50 + k, graph_seed=60 + k, weight_seed=70 + k)
        for _ in range(steps):
            M.step_opinion(vh, solo)
        a = np.asarray(m.sim._read_agents(m.T)[0]).view(np.uint8)
        b = np.asarray(solo.sim._read_agents(solo.T)[0]).view(np.uint8)
        assert np.array_equal(a, b), f"member {k} n={n}"
        assert m.sim._num_transitions() == steps + 1
    n, md, steps = specs[0]
    c = M.build_opinion_gnp(ora, n, mean_deg=md, seed=50, graph_seed=60, weight_seed=70, device_graph=False)
    for _ in range(steps):
        M.step_opinion(ora, c)
    assert np.array_equal(np.asarray(sims[0].sim._read_agents(sims[0].T)[0]).view(np.uint8),
                          np.asarray(c.sim._read_agents(c.T)[0]).view(np.uint8))
    B.close()
    for m in sims:
        m.sim._destroy()
    lib.stream_destroy(st)


def test_model_destroyed_before_simulation_device(vh, gpu_ready):
    """The device simulation keeps its own copy of the model's types: the
    model handle destroyed first leaves it usable and destroyable."""
    print("[gpu] model destroyed before its simulation", flush=True)
    sim, A, E, ids = _single_edge_sim(vh)
    sim.model.lib.model_destroy(sim.model.handle)
    sim.model.handle = type(sim.model.handle)()
    sim._add_edges_array(E, [ids[0]], [ids[1]], _st(E, [3]))
    vh.finish_init_(sim)
    c = vh.copy_simulation(sim)
    assert vh.num_edges(sim, E) == 1 and vh.num_edges(c, E) == 1
    c._destroy()
    sim._destroy()
