"""Parameters and globals (CPU, on the checker): register_param!/
register_global!, param, set_param!, get_global, set_global!,
modify_global!, push_global! (src/Simulation.jl:261-339,587-630,
src/Global.jl, src/ModelTypes.jl:244-288), and transition bodies that read
them (blob fields bound to the simulation's values)."""
import types

import numpy as np
import pytest


def _model(V, name="globals", **reg):
    A = V.struct("A", [("foo", "<i8")])
    t = V.ModelTypes() | V.register_agenttype_(A)
    for k, v in reg.get("params", {}).items():
        t = t | V.register_param_(k, v)
    for k, v in reg.get("globals", {}).items():
        t = t | V.register_global_(k, v)
    return t | V.create_model(name), A


def test_globals_known_answers(ora):
    """test/globals.jl:1-17: a mutable globals struct passed to
    create_simulation; set_global!, push_global! twice, get_global."""
    model, _ = _model(ora)
    g = types.SimpleNamespace(foo=0.0, bar=[])
    sim = ora.create_simulation(model, None, g)
    ora.set_global_(sim, "foo", 1.1)
    ora.push_global_(sim, "bar", 1)
    ora.push_global_(sim, "bar", 2)
    assert ora.get_global(sim, "foo") == 1.1
    assert ora.get_global(sim, "bar") == [1, 2]
    assert sim._value(1, "foo") == 1.1          # the library's copy (what bodies read)
    ora.finish_simulation_(sim)


def test_registered_params_and_globals(ora):
    model, A = _model(ora, params={"p": 0.25, "k": 3}, globals={"g": 1.5, "series": []})
    sim = ora.create_simulation(model)
    assert ora.param(sim, "p") == 0.25 and ora.param(sim, ":k") == 3
    ora.set_param_(sim, "p", 0.5)
    ora.set_param_("k", 7)(sim)                 # the pipeable form
    assert ora.param(sim, "p") == 0.5 and sim._value(0, "p") == 0.5 and sim._value(0, "k") == 7
    with pytest.warns(UserWarning):
        ora.set_param_(sim, "nope", 1)
    ora.modify_global_(sim, "g", lambda v: v * 2)
    assert ora.get_global(sim, "g") == 3.0 and sim._value(1, "g") == 3.0
    ora.push_global_(sim, "series", 4)
    assert ora.get_global(sim, "series") == [4]
    assert sim.globals_last_change == -1        # num_transitions - 1 before finish_init!
    ora.add_agents_(sim, A(1), A(2))
    ora.finish_init_(sim)
    with pytest.raises(AssertionError):         # set_param! only before finish_init!
        ora.set_param_(sim, "p", 0.1)
    ora.set_global_(sim, "g", 2.0)              # globals stay settable
    assert sim.globals_last_change == 0
    # a second simulation of the model starts from the registered defaults
    sim2 = ora.create_simulation(model)
    assert ora.param(sim2, "p") == 0.25 and ora.get_global(sim2, "series") == []


def test_register_and_argument_exclusive(ora):
    model, _ = _model(ora, params={"p": 0.25})
    with pytest.raises(AssertionError):
        ora.create_simulation(model, types.SimpleNamespace(p=1.0))
    model2, _ = _model(ora, globals={"g": 1})
    with pytest.raises(AssertionError):
        ora.create_simulation(model2, None, {"g": 2})


def _sir_sim(V, n, params=None, globals_=None):
    M = V.models
    SIR, Contact, *_ = M._types(V)
    t = (V.ModelTypes() | V.register_agenttype_(SIR, "Immortal")
         | V.register_edgetype_(Contact, "Stateless", target=SIR))
    for k, v in (params or {}).items():
        t = t | V.register_param_(k, v)
    for k, v in (globals_ or {}).items():
        t = t | V.register_global_(k, v)
    model = t | V.create_model("sir_values")
    sim = V.create_simulation(model)
    sim._add_agents_array(SIR, V.inputs.sir_init(n, 5, 20))
    fr, to = V.inputs.graph_er(n, 8.0 / (n - 1), 5)
    tid = np.uint64(1) << np.uint64(56)
    sim._add_edges_array(Contact, tid | fr.astype(np.uint64), tid | to.astype(np.uint64))
    V.finish_init_(sim)
    return types.SimpleNamespace(sim=sim, T=SIR, E=Contact)


def _states(m):
    return np.ascontiguousarray(m.sim._read_agents(m.T)[0]).view(np.uint8)


def bound_sir_matches_literal(V, n=600):
    """SIR with p_inf from a parameter and the recovery days from a global
    (changed after 4 steps) equals the literal runs step by step."""
    a = _sir_sim(V, n, params={"p_inf": 0.3, "seed": 5}, globals_={"recovery": 7})
    b = _sir_sim(V, n)
    tr = V.transitions.sir(a.E, V.param_ref("p_inf"), V.param_ref("seed"), V.global_ref("recovery"))
    for step in range(9):
        if step == 4:
            V.set_global_(a.sim, "recovery", 3)
        V.apply_(a.sim, tr, a.T, [a.T, a.E], a.T)
        V.apply_(b.sim, V.transitions.sir(b.E, 0.3, 5, 7 if step < 4 else 3), b.T, [b.T, b.E], b.T)
        assert np.array_equal(_states(a), _states(b)), step
    return _states(a)


def test_bound_sir_matches_literal(ora):
    bound_sir_matches_literal(ora)


def test_binding_errors(ora):
    m = _sir_sim(ora, 50, params={"p_inf": 0.3, "frac": 2.5})
    with pytest.raises(AssertionError):         # no such parameter
        ora.apply_(m.sim, ora.transitions.sir(m.E, ora.param_ref("nope"), 1, 7), m.T, [m.T, m.E], m.T)
    with pytest.raises(AssertionError):         # InexactError: 2.5 into a UInt32 field
        ora.apply_(m.sim, ora.transitions.sir(m.E, 0.3, 1, ora.param_ref("frac")), m.T, [m.T, m.E], m.T)
    # an integral float converts (convert(UInt32, 7.0) == 7)
    m2 = _sir_sim(ora, 50, params={"days": 7.0})
    ora.apply_(m2.sim, ora.transitions.sir(m2.E, 0.3, 1, ora.param_ref("days")), m2.T, [m2.T, m2.E], m2.T)
