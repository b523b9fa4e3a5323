"""Benchmark of apply! on one MI355X.

Workload (default): weighted bounded-confidence opinion dynamics, 10^7 agents
on a symmetrised Barabasi-Albert graph (m = 5, ~10^8 directed edges), Float64
opinions, Float32 edge weights, eps = 0.2.  A step is one apply! over every
agent.  Inputs are synthetic (seeded counter hash), resident in HBM before the
timed region starts.

Prints one JSON line: agent-updates/s (value), edges/s, the roofline of the
transition kernel (algorithmic bytes / HIP-event time of the launch on the
simulation's stream) and the CPU baseline (the checker's apply! restatement on
one host core over a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def opinion_bytes(n, e, f32):
    """Compulsory traffic of one step: read + write the state, the offsets,
    one source index and one Float32 weight per edge; every neighbour state
    read once (as part of the state read)."""
    sb = 4 if f32 else 8
    return n * (sb + sb) + (n + 1) * 4 + e * (4 + 4)


def sir_bytes(n, e):
    return n * (4 + 4) + (n + 1) * 4 + e * 4


def _cpu_sample(workload, f32, seed=1, seconds_cap=25.0, relabel=True, n=None, edges=None, steps=5,
                warmup=1):
    """The checker's apply! (a restatement of the reference's data structures)
    over the workload: `n` agents (default: 10^6; 2*10^5 for the contact
    model), median of up to `steps` steps after `warmup` warm-up steps, at
    most `seconds_cap` seconds of timed steps. `edges`: the opinion graph to
    use (otherwise generated from `seed`)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as ora
    from __graft_entry__ import load_package
    vh = load_package()
    M = vh.models
    n = n or 1_000_000
    if workload.startswith("opinion"):
        if edges is None:
            a, b = vh.inputs.graph_ba(n, 5, seed)
            edges = vh.inputs.symmetric_weighted(a, b, seed + 1)
        m = M.build_opinion(ora, n, seed=seed, f32=f32, edges=edges, relabel=relabel)
        step = lambda: M.step_opinion(ora, m, eps=0.2)  # noqa: E731
        ne = len(edges[0])
    elif workload == "contact":
        n = min(n, 200_000)
        m = M.build_contact(ora, n, seed=seed)
        step = lambda: M.step_contact(ora, m, seed=seed)  # noqa: E731
        ne = 20 * n
    elif workload == "raster":
        m = M.build_raster_diffusion(ora, 1000, seed=seed)
        n = m.n
        step = lambda: M.step_raster_diffusion(ora, m)  # noqa: E731
        ne = 8 * n
    else:
        edges = vh.inputs.graph_er(n, 10.0 / (n - 1), seed)
        m = M.build_sir(ora, n, seed=seed, edges=edges)
        step = lambda: M.step_sir(ora, m)  # noqa: E731
        ne = len(edges[0])
    for _ in range(warmup):
        step()
    times = []
    t_all = time.perf_counter()
    while len(times) < steps and time.perf_counter() - t_all < seconds_cap:
        t0 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t0)
    times.sort()
    return n, ne, times


def cpu_baseline(workload, f32, relabel=True, n=None, edges=None):
    """Setting (i): one host core (the same agent numbering as the GPU run).
    At 10^7 opinion agents (the headline workload itself, same graph): the
    median of 2 steps, no warm-up (a step takes seconds; the checker has no
    JIT)."""
    big = n is not None and n >= 5_000_000
    n_, ne, times = _cpu_sample(workload, f32, relabel=relabel, n=n, edges=edges,
                                steps=2 if big else 5, warmup=0 if big else 1,
                                seconds_cap=40.0 if big else 25.0)
    med = times[len(times) // 2]
    return {"value": n_ / med, "unit": "agent-updates/s", "cores": 1, "kind": "port",
            "edges_per_s": ne / med,
            "sample": f"{workload} N={n_} E={ne}, median of {len(times)} apply! steps"
                      f"{'' if big else ' after 1 warm-up'}, 1 host core"}


def cpu_baseline_all_cores(workload, f32, relabel=True, n=None):
    """Setting (ii): the host cores as P independent processes, each owning a
    contiguous share of the agents (N / P of them: its own rank-seeded graph
    of that size; the exchange between ranks is not modelled), run at the
    same time. The aggregate rate is all ranks' agents over the slowest
    rank's median step, as ranks advance in lockstep. P = the CPUs this
    process may use, at most 15. The workers make no GPU calls."""
    import subprocess
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    P = max(1, min(ncpu, 15))
    env = dict(os.environ, ROCR_VISIBLE_DEVICES="", HIP_VISIBLE_DEVICES="", OMP_NUM_THREADS="1")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-worker", "--workload", workload]
    if n:
        cmd += ["--n", str(max(1000, n // P))]
    if relabel:
        cmd.append("--relabel")
    procs = [subprocess.Popen(cmd + ["--rank", str(r)], stdout=subprocess.PIPE, env=env)
             for r in range(P)]
    res = []
    for pr in procs:
        out, _ = pr.communicate(timeout=600)
        if pr.returncode != 0:
            raise RuntimeError(f"cpu worker failed ({pr.returncode})")
        res.append(json.loads(out.decode().strip().splitlines()[-1]))
    slow = max(r["median_s"] for r in res)
    n_all = sum(r["n"] for r in res)
    ne = sum(r["edges"] for r in res)
    return {"value": n_all / slow, "unit": "agent-updates/s", "cores": P, "kind": "port",
            "edges_per_s": ne / slow,
            "sample": f"{workload}: {P} processes x N={res[0]['n']} (N={n_all} in all; rank-seeded "
                      f"graphs, no halo exchange), slowest rank's median apply! step {slow:.3f} s"}


def cpu_worker(args):
    f32 = args.workload == "opinion32"
    n, ne, times = _cpu_sample(args.workload, f32, seed=1 + args.rank, relabel=args.relabel,
                               n=args.n if args.n != 10_000_000 else None)
    print(json.dumps({"n": n, "edges": ne, "median_s": times[len(times) // 2]}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--workload", default="opinion",
                    choices=["opinion", "opinion32", "sir", "contact", "raster", "ensemble",
                             "serve"])
    ap.add_argument("--replicas", type=int, default=1024)
    ap.add_argument("--requests", type=int, default=60)
    ap.add_argument("--rate", type=float, default=0.0,
                    help="serve: Poisson arrivals/s (0: 0.8 x the measured one-at-a-time capacity)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-classes", action="store_true",
                    help="serve: also measure capacity per size class (30 requests per decade)")
    ap.add_argument("--relabel", action="store_true",
                    help="opinion: renumber the agents breadth-first at finish_init! (default: "
                         "generation order)")
    ap.add_argument("--no-compare", action="store_true",
                    help="opinion: skip the second timing in the other agent order")
    ap.add_argument("--one-pass", action="store_true",
                    help="opinion: gather in one pass (option opinion_plan_bytes = -1) instead of "
                         "through the two-pass plan")
    ap.add_argument("--cpu-worker", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--rank", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--calib", action="store_true",
                    help="run a 256 MiB device copy first (HBM counter calibration)")
    args = ap.parse_args()
    if args.cpu_worker:   # one rank of the all-cores CPU baseline (no GPU use)
        return cpu_worker(args)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: every simulation here lives on a single GPU "
              "(no multi-GPU path is built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from __graft_entry__ import load_package

    vh = load_package()
    M = vh.models
    torch.cuda.init()
    if args.one_pass:
        vh.set_options(opinion_plan_bytes=-1)
    if args.workload == "ensemble":
        return bench_ensemble(vh, args)
    if args.workload == "serve":
        return bench_serve(vh, args)
    dev = torch.device("cuda:0")
    n = args.n
    f32 = args.workload == "opinion32"
    t0 = time.perf_counter()
    if args.workload.startswith("opinion"):
        # the graph and the breadth-first numbering built on the device at
        # finish_init! (the same graph the host generator gives; the CPU
        # baseline generates its copy on the host)
        edges = None
        relabel = args.relabel
        m = M.build_opinion(vh, n, seed=1, f32=f32, relabel=relabel, device_graph=True)
        step = lambda: M.step_opinion(vh, m, eps=0.2)  # noqa: E731
        nbytes = opinion_bytes(n, m.nedges, f32)
        dtype = "f32" if f32 else "f64"
        workload = (f"opinion dynamics (weighted bounded confidence), Barabasi-Albert m=5 symmetrised, "
                    f"N={n}, E={m.nedges}, eps=0.2, {'Float32' if f32 else 'Float64'} opinions"
                    + (", agents renumbered breadth-first at finish_init! (idmapping)" if relabel
                       else ""))
    elif args.workload == "raster":
        import math
        side = int(round(math.sqrt(n)))
        n = side * side
        edges = None
        m = M.build_raster_diffusion(vh, side, seed=1)
        step = lambda: M.step_raster_diffusion(vh, m)  # noqa: E731
        m.nedges = 8 * n
        # cell state read + written (1 B each); neighbour states count once as
        # part of the state read. The stencil's in-edges are computed from the
        # stencil (the device-built CSR is not read by the diffuse step).
        nbytes = n * (1 + 1)
        dtype = "u8"
        workload = (f"raster diffusion (test/raster.jl diffuse), {side}x{side} periodic Moore "
                    f"neighbourhood, N={n}, stencil edges built on the device")
    elif args.workload == "contact":
        edges = None
        m = M.build_contact(vh, n, seed=1)
        step = lambda: M.step_contact(vh, m, seed=1)  # noqa: E731
        m.nedges = 20 * n
        # survivors' state r/w, died flags, previous contacts read, new contacts
        # written (offsets + sources)
        nbytes = n * (8 + 8 + 1) + 2 * (n + 1) * 4 + 4 * 20 * n + 4 * 20 * n
        dtype = "u8"
        workload = (f"dynamic contact network, N={n}, 20 new contacts per agent and step, "
                    f"2% deaths with 1:1 births")
    else:
        edges = vh.inputs.graph_er(n, 10.0 / (n - 1), 1)
        m = M.build_sir(vh, n, seed=1, edges=edges)
        step = lambda: M.step_sir(vh, m)  # noqa: E731
        nbytes = sir_bytes(n, m.nedges)
        dtype = "u8"
        workload = f"SIR contagion, directed G(n,p), N={n}, E={m.nedges}, p_inf=0.05"
    if not args.workload.startswith("opinion"):
        del edges
    init_s = time.perf_counter() - t0
    log(f"init {init_s:.2f}s: {workload}")

    stream = torch.cuda.Stream(device=dev)
    L = vh._abi.lib()

    if args.calib:
        a = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    def timed(sim, step):
        """W warm-up steps, then K steps between two device synchronisations,
        with HIP events on the simulation's stream around every step: (wall
        ms per step, mean event ms, sorted event ms)."""
        L.set_stream(sim.handle, stream.cuda_stream)
        vh.enable_asserts(False)   # production mode (enable_asserts(false) in the reference)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            starts[k].record(stream)
            step()
            ends[k].record(stream)
        L.sync(sim.handle)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ev = sorted(s.elapsed_time(e) for s, e in zip(starts, ends))
        vh.enable_asserts(True)
        assert sim._num_transitions() == 1 + args.warmup + args.steps
        return wall * 1e3 / args.steps, sum(ev) / len(ev), ev

    ms, ev_mean, ev = timed(m.sim, step)
    value = n / (ms * 1e-3)
    edges_s = m.nedges / (ms * 1e-3)
    achieved = nbytes / (ev_mean * 1e-3) / 1e9
    log(f"step wall {ms:.3f} ms, kernel events mean {ev_mean:.3f} ms (min {ev[0]:.3f} max {ev[-1]:.3f})")

    other_order = None
    if args.workload.startswith("opinion") and not args.no_compare:
        # the same graph and states in the other agent order (breadth-first
        # numbering at finish_init! when the headline keeps the generation
        # order, and the reverse), timed the same way
        m2 = M.build_opinion(vh, n, seed=1, f32=f32, relabel=not args.relabel, device_graph=True)
        ms2, ev2, _ = timed(m2.sim, lambda: M.step_opinion(vh, m2, eps=0.2))
        other_order = {"ms_per_step": ms2, "kernel_ms": ev2, "value": n / (ms2 * 1e-3),
                       "roofline_frac": nbytes / (ev2 * 1e-3) / 1e9 / HBM_PEAK_GBS}
        log(f"{'generation' if args.relabel else 'breadth-first'} order: step wall {ms2:.3f} ms, "
            f"kernel events mean {ev2:.3f} ms")
        m2.sim._destroy()
        del m2

    traffic = None
    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    # runs of this exact workload (scaled by the copy-kernel calibration)
    # (the Float64 state of 10^7 agents is gathered in two passes unless
    # the option opinion_plan_bytes turns that off)
    pb = vh.get_options()["opinion_plan_bytes"]
    two_pass = 0 <= pb <= n * 8
    pmc = {("opinion", 10_000_000, False, False): "k_opinion_f64_ba10M_traffic.json",
           ("opinion", 10_000_000, True, False): "k_opinion_f64_ba10M_bfs_traffic.json",
           ("opinion", 10_000_000, False, True): "k_opinion_f64_ba10M_plan_traffic.json"}.get(
        (args.workload, n, args.relabel, two_pass))
    if pmc and os.path.exists(os.path.join(ROOT, "profiles", pmc)):
        traffic = json.load(open(os.path.join(ROOT, "profiles", pmc))).get("bytes_per_launch")

    drift = None
    if f32:
        # the Float32 variant against the Float64 run of the same inputs
        m64 = M.build_opinion(vh, n, seed=1, f32=False, relabel=args.relabel, device_graph=True)
        for _ in range(1 + args.warmup + args.steps - 1):
            M.step_opinion(vh, m64, eps=0.2)
        x64 = np.asarray(m64.sim._read_agents(m64.T)[0]["opinion"])
        x32 = np.asarray(m.sim._read_agents(m.T)[0]["opinion"]).astype(np.float64)
        d = np.abs(x32 - x64)
        drift = {"steps": args.warmup + args.steps, "max_abs": float(d.max()),
                 "mean_abs": float(d.mean()), "frac_above_1e-3": float((d > 1e-3).mean())}
        log(f"f32 drift vs f64: {drift}")
        m64.sim._destroy()

    cpu = None
    if not args.no_cpu_baseline:
        # opinion: the checker on the headline workload itself (same N, same
        # graph); other workloads: their bounded samples
        opin = args.workload.startswith("opinion")
        m.sim._destroy()
        t0 = time.perf_counter()
        cpu = cpu_baseline(args.workload, f32, relabel=args.relabel, n=n if opin else None,
                           edges=edges if opin else None)
        log(f"cpu baseline {time.perf_counter() - t0:.1f}s: {cpu}")
        t0 = time.perf_counter()
        cpu["all_cores"] = cpu_baseline_all_cores(args.workload, f32, relabel=args.relabel,
                                                  n=n if opin else None)
        log(f"cpu baseline, all cores {time.perf_counter() - t0:.1f}s: {cpu['all_cores']}")

    # the metric's second half (ensemble p50/p99 against one at a time):
    # config c3, 1024 SIR replicas x 200 steps in one launch per step
    ens = None
    if args.workload == "opinion" and not args.no_compare:
        t0 = time.perf_counter()
        e = ensemble_run(vh)
        ens = {k: e[k] for k in ("replicas", "steps", "ms_per_step", "replicas_per_s",
                                 "one_at_a_time_replicas_per_s", "speedup_vs_one_at_a_time", "value",
                                 "edges_per_s", "autotune")}
        ens["agent_updates_per_s"] = ens.pop("value")
        ens["latency_s"] = {k: e["latency_s"][k] for k in ("ensemble", "one_at_a_time")}
        log(f"ensemble {time.perf_counter() - t0:.1f}s: {ens['speedup_vs_one_at_a_time']:.2f}x one at a time, "
            f"p99 {ens['latency_s']['ensemble']['p99']:.3f} s vs {ens['latency_s']['one_at_a_time']['p99']:.3f} s")

    out = {
        "metric": "agent-updates/s per apply! (edges/s alongside)",
        "value": value,
        "unit": "agent-updates/s",
        "edges_per_s": edges_s,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded counter-hash graph and states)",
        "config": {"workload": workload, "agents": n, "edges": m.nedges, "parallelism": "single GPU",
                   "init_s": round(init_s, 3),
                   **({"gather": "two-pass" if two_pass else "one-pass"} if args.workload.startswith("opinion")
                      else {})},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": nbytes, "kernel_ms": ev_mean},
        "cpu_baseline": cpu,
    }
    if other_order is not None:
        out["generation_order" if args.relabel else "breadth_first_order"] = other_order
    if ens is not None:
        out["ensemble"] = ens
    if drift is not None:
        out["f32_drift_vs_f64"] = drift
    print(json.dumps(out), flush=True)


def bench_ensemble(vh, args):
    out = ensemble_run(vh, args.replicas, args.steps if args.steps != 50 else 200)
    print(json.dumps(out), flush=True)


def ensemble_run(vh, R=1024, steps=200):
    """1024 SIR replicas (sizes log-uniform in [1e3, 1e6], mean in-degree
    U[4, 20], p_inf U[0.01, 0.1], distinct seeds), 200 steps in one
    graph-captured launch sequence, against the same replicas one at a time."""
    import numpy as np
    rng = np.random.default_rng(7)
    specs = [(int(10 ** rng.uniform(3, 6)), float(rng.uniform(4, 20)), float(rng.uniform(0.01, 0.1)),
              1 + i) for i in range(R)]
    E = vh.ensemble.Ensemble()
    t0 = time.perf_counter()
    ids = [E.add_sir(n, md, p, seed, 10_000 + seed, 7, max(1, n // 1000)) for n, md, p, seed in specs]
    info = E.info()
    log(f"ensemble built in {time.perf_counter() - t0:.1f}s: {info}")
    block, remap, tms = E.autotune(2)
    log(f"autotune: block {block}, xcd_remap {remap}, agents per thread by size class {E.class_apt}, "
        f"{tms:.3f} ms/step")
    E.step(1, use_graph=False)
    E.sync()
    t0 = time.perf_counter()
    E.step(steps, use_graph=True)
    E.sync()
    wall = time.perf_counter() - t0
    agent_steps = info["agents"] * steps
    # one at a time: a sample of the replicas, each alone (same kernels, same
    # graph-captured step sequence), scaled to all replicas by agent count.
    # Each replica runs twice with and twice without the exposure pass
    # (option sir_expose) and its fastest run counts: alone, a small replica pays
    # the pass's two extra launches per step without gaining from it.
    sample = list(range(0, R, max(1, R // 64)))
    solo_time, solo_agents = 0.0, 0
    solo_by_mode = {"1": 0.0, "0": 0.0}
    service = []   # each sampled replica's run time alone (its faster mode)
    saved_opts = vh.get_options()
    for i in sample:
        n, md, p, seed = specs[i]
        best = None
        for mode in ("1", "0"):
            vh.set_options(sir_expose=int(mode))
            mode_best = None
            for _rep in range(2):   # the same steps from the same start each time
                S = vh.ensemble.Ensemble()
                S.add_sir(n, md, p, seed, 10_000 + seed, 7, max(1, n // 1000))
                S.step(1, use_graph=False)
                S.sync()
                t1 = time.perf_counter()
                S.step(steps, use_graph=True)
                S.sync()
                dt = time.perf_counter() - t1
                S.close()
                mode_best = dt if mode_best is None else min(mode_best, dt)
            solo_by_mode[mode] += mode_best
            best = mode_best if best is None else min(best, mode_best)
        solo_time += best
        solo_agents += n
        service.append(best)
    vh.set_options(**saved_opts)
    solo_rate = solo_agents * steps / solo_time
    solo_total = len(specs) * (solo_time / len(sample))
    # latency of every replica when all R arrive at t = 0: in the ensemble all
    # finish with the last launch; one at a time, replica k finishes after the
    # run times of those before it (each sampled replica stands for R / sample
    # replicas in spec order)
    rep = max(1, R // len(sample))
    seq_done = np.cumsum(np.repeat(np.asarray(service), rep))
    pct = lambda a, q: float(np.percentile(a, q))  # noqa: E731
    out = {"metric": "ensemble agent-updates/s (1024 SIR replicas, one launch per step)",
           "value": agent_steps / wall, "unit": "agent-updates/s",
           "edges_per_s": info["edges"] * steps / wall, "steps": steps,
           "ms_per_step": wall * 1e3 / steps, "replicas": R, "agents": info["agents"],
           "edges": info["edges"],
           "autotune": {"block": block, "xcd_remap": remap,
                        "agents_per_thread_by_size_class": {"<1e4": E.class_apt[0], "1e4-1e5": E.class_apt[1],
                                                            ">=1e5": E.class_apt[2]}},
           "latency_s": {"ensemble": {"p50": wall, "p99": wall},
                         "one_at_a_time": {"p50": pct(seq_done, 50), "p99": pct(seq_done, 99)},
                         "one_at_a_time_alone": {"p50": pct(service, 50), "p99": pct(service, 99)},
                         "note": "all replicas submitted at t=0; 'alone' is a replica's own run time"},
           "one_at_a_time": {"agent_updates_per_s": solo_rate, "sample_replicas": len(sample),
                             "est_total_s": solo_total,
                             "est_total_s_exposure_on": len(specs) * solo_by_mode["1"] / len(sample),
                             "est_total_s_exposure_off": len(specs) * solo_by_mode["0"] / len(sample),
                             "note": "each sampled replica's faster run (exposure pass on or off)"},
           "replicas_per_s": R / wall, "one_at_a_time_replicas_per_s": R / solo_total,
           "speedup_vs_one_at_a_time": solo_total / wall, "dtype": "u8", "n_gpus": 1,
           "higher_is_better": True}
    E.close()
    return out


def bench_serve(vh, args):
    """Mixed SIR / opinion requests (sizes log-uniform in [1e4, 1e7], horizons
    U[50, 500], mean degree U[4, 20]) served with continuous batching under an
    HBM budget, against one-at-a-time execution of the same requests.
    1) capacity: all requests present at t = 0, requests/s of each mode;
    2) latency: Poisson arrivals at 0.8 x the one-at-a-time capacity, p50/p99
       of each mode on the same arrival sequence."""
    import torch
    S = vh.serve
    free, total = torch.cuda.mem_get_info()
    budget = int(0.9 * free)

    def served(reqs):
        srv = S.Server(vh, hbm_budget=budget, quantum=10)
        done, wall = srv.run(reqs)
        srv.close()
        log(f"served phases (host s): {({k: round(v, 3) for k, v in srv.stats.items()})}")
        return done, S.summarize(done, wall)

    burst = S.make_requests(args.requests, rate=0, seed=11)
    d_cap, cap_srv = served(burst)
    log(f"capacity, served: {cap_srv}")
    ph = {}
    s_cap, wall = S.run_sequential(vh, burst, stats=ph)
    log(f"one-at-a-time phases (host s): {({k: round(v, 3) for k, v in ph.items()})}")
    cap_seq = S.summarize(s_cap, wall)
    # GPU-busy accounting of the one-at-a-time run: the time a request's steps
    # hold the GPU (enqueue to completion, nothing else running), per size
    # class; no schedule serves the mix faster than the sum of the classes'
    # step time that batching cannot remove (the large requests fill the GPU
    # alone)
    by_class = {}
    for lo, hi, name in ((1e4, 1e5, "1e4-1e5"), (1e5, 1e6, "1e5-1e6"), (1e6, 1e7 + 1, "1e6-1e7")):
        rs = [r for r in s_cap if lo <= r.n < hi]
        by_class[name] = {"requests": len(rs), "agent_steps": int(sum(r.n * r.horizon for r in rs)),
                          "run_s": sum(r.run_s for r in rs), "build_s": sum(r.build_s for r in rs),
                          "read_s": sum(r.read_s for r in rs)}
    run_sum = sum(r.run_s for r in s_cap)
    log(f"capacity, one at a time: {cap_seq}")
    same = {r.id: r.digest for r in d_cap} == {r.id: r.digest for r in s_cap}
    rate = args.rate if args.rate > 0 else 0.8 * cap_seq["throughput_rps"]
    reqs = S.make_requests(args.requests, rate=rate, seed=12)
    d_lat, lat_srv = served(reqs)
    log(f"Poisson {rate:.2f}/s, served: {lat_srv}")
    s_lat, wall = S.run_sequential(vh, reqs)
    lat_seq = S.summarize(s_lat, wall)
    log(f"Poisson {rate:.2f}/s, one at a time: {lat_seq}")
    same = same and {r.id: r.digest for r in d_lat} == {r.id: r.digest for r in s_lat}
    classes = None
    if args.serve_classes:
        # capacity by size class: 30 requests of one decade each, all at t = 0
        classes = {}
        for lo, hi, name in ((1e4, 1e5, "1e4-1e5"), (1e5, 1e6, "1e5-1e6"), (1e6, 1e7, "1e6-1e7")):
            reqs_c = S.make_requests(30, rate=0, seed=21, n_min=lo, n_max=hi)
            dc, srv_c = served(reqs_c)
            sc, wall_c = S.run_sequential(vh, reqs_c)
            seq_c = S.summarize(sc, wall_c)
            classes[name] = {"served_rps": srv_c["throughput_rps"], "one_at_a_time_rps": seq_c["throughput_rps"],
                             "ratio": srv_c["throughput_rps"] / seq_c["throughput_rps"],
                             "served_p99_s": srv_c["p99_s"], "one_at_a_time_p99_s": seq_c["p99_s"],
                             "identical": {r.id: r.digest for r in dc} == {r.id: r.digest for r in sc}}
            log(f"class {name}: {classes[name]}")
    out = {"metric": "served requests/s with p50/p99 latency (SIR + opinion mix)",
           "value": lat_srv["throughput_rps"], "unit": "requests/s",
           "capacity_served_rps": cap_srv["throughput_rps"],
           "capacity_one_at_a_time_rps": cap_seq["throughput_rps"],
           "capacity_ratio": cap_srv["throughput_rps"] / cap_seq["throughput_rps"],
           "poisson_rate_per_s": rate, "served": lat_srv, "one_at_a_time": lat_seq,
           "p50_ratio": lat_srv["p50_s"] / lat_seq["p50_s"],
           "p99_ratio": lat_srv["p99_s"] / lat_seq["p99_s"],
           "results_identical": same, "requests": args.requests, "hbm_budget_bytes": budget,
           "gpu_busy": {"one_at_a_time_wall_s": cap_seq["wall_s"], "one_at_a_time_step_s": run_sum,
                        "served_wall_s": cap_srv["wall_s"], "by_size_class": by_class,
                        "capacity_ratio_bound_large_classes": cap_seq["wall_s"] / max(
                            1e-9, by_class["1e6-1e7"]["run_s"] + by_class["1e5-1e6"]["run_s"]),
                        "capacity_ratio_bound_steps_and_builds": cap_seq["wall_s"] / max(
                            1e-9, sum(c["run_s"] + c["build_s"] for c in by_class.values())),
                        "note": "step_s: enqueue-to-completion time of each request's steps when it runs "
                                "alone; the 1e5-1e7 classes' step time is GPU-bound work that batching "
                                "cannot remove; graph builds are device kernels as well"},
           "n_gpus": 1, "higher_is_better": True,
           "config": {"workload": "serving: Poisson SIR/opinion requests, N ~ logU[1e4, 1e7], "
                                  "horizon ~ U[50, 500], mean degree ~ U[4, 20], device-built G(n, p)"}}
    if classes is not None:
        out["capacity_by_size_class"] = classes
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
