"""Serving loop: a request stream of SIR and opinion-dynamics simulations.

Requests arrive over time (Poisson), are admitted first-come-first-served
while their estimated device footprint fits the HBM budget, keep their graph
and state on the device between steps, and leave when their horizon is done.
Both models build their graphs on the device (directed G(n, p), mean degree
from the request): SIR requests as ensemble members, opinion requests as
simulations whose edges come from vh_add_edges_gnp.
Continuous batching: every scheduler iteration advances all SIR requests by
one segmented, graph-captured ensemble launch per step, all small opinion
requests (state below the two-pass threshold) by one fused launch per step
(vh_batch_apply) and every large opinion request by one vh_apply_n call, all
on one compute stream; requests that arrive meanwhile join at the next
iteration as budget frees up.

`run_sequential` is the one-at-a-time baseline: the same requests, each built,
run to its horizon and read back before the next one starts.
"""
import ctypes as C
import time
import types

import numpy as np
import xxhash

from . import VahanaError, _abi, models
from .ensemble import Ensemble


def make_requests(count, rate, seed=1, n_min=10_000, n_max=10_000_000, h_min=50, h_max=500,
                  sir_fraction=0.5):
    """Poisson arrivals (rate per second), sizes log-uniform in [n_min, n_max],
    horizons uniform in [h_min, h_max], a model mix."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count)) if rate > 0 else np.zeros(count)
    out = []
    for i in range(count):
        n = int(round(10 ** rng.uniform(np.log10(n_min), np.log10(n_max))))
        out.append(types.SimpleNamespace(
            id=i, kind="sir" if rng.random() < sir_fraction else "opinion", n=n,
            horizon=int(rng.integers(h_min, h_max + 1)), seed=1000 + i, arrival=float(t[i]),
            mean_deg=float(rng.uniform(4, 20)), p_inf=float(rng.uniform(0.01, 0.1))))
    return out


# state vectors from this size on are gathered in two passes by the opinion
# kernels (the default of the option opinion_plan_bytes)
OPINION_PLAN_MIN_BYTES = 16 << 20


def footprint(req):
    """Device bytes a request holds while active (state, CSR, build scratch)."""
    e = req.n * req.mean_deg
    if req.kind == "sir":
        # states, offsets, sources, generator scratch; out-edges and the
        # exposure bitmaps
        return int(req.n * 12 + e * 4 + req.n * 8 + e * 4 + req.n * 5)
    # two Float64 state buffers, offsets, sources + weights, generator scratch
    fp = req.n * 16 + req.n * 4 + e * 8 + req.n * 8
    if req.n * 8 >= OPINION_PLAN_MIN_BYTES:
        # the two-pass gather: local source indices, gathered states, lists
        # with their weights, run tables; the scratch of the one-time plan
        # build (its sorts)
        fp += e * 17 + req.n * 8 + e * 53
    return int(fp)


def _build_opinion(vh, req, stream=None, keep=None):
    return models.build_opinion_gnp(vh, req.n, mean_deg=req.mean_deg, seed=req.seed,
                                    graph_seed=req.seed + 1, weight_seed=req.seed + 7, stream=stream,
                                    keep=keep)


def _stream(lib):
    h = C.c_void_p()
    rc = lib.stream_create(C.byref(h))
    if rc != _abi.VH_OK:
        raise RuntimeError(f"vh_stream_create failed ({rc})")
    return h


def _digest(arr):
    """128-bit digest of the result bytes, to compare results across runs."""
    return xxhash.xxh3_128_hexdigest(np.ascontiguousarray(arr).view(np.uint8))


class Server:
    """`solo_agents`: SIR requests with at least this many agents run in an
    ensemble of their own (their `quantum` steps back to back, so that their
    graph stays in the 256 MB Infinity Cache from step to step); smaller ones
    share one ensemble whose launches advance all of them at once.
    `large_concurrency`: how many large requests (those with kernels of
    their own: solo SIR, opinion through the two-pass plan) advance in one
    quantum, oldest admitted first (None: all). The small ones always all
    advance. Large requests take nearly all GPU time, so sharing it among
    them only delays each; in admission order they finish as early as one at
    a time, while the small ones no longer wait behind them."""

    def __init__(self, vh, hbm_budget, quantum=10, keep_results=False, solo_agents=1_000_000,
                 large_concurrency=1):
        self.vh = vh
        self.budget = int(hbm_budget)
        self.quantum = quantum
        self.keep = keep_results
        self.solo_agents = solo_agents
        self.large_concurrency = large_concurrency
        # Two streams in all: every step launch goes to `compute`, every
        # graph build and result read to `io`. Simulations do not get streams
        # of their own: with a handful of hardware queues per process, many
        # streams share queues and a read would wait behind unrelated steps.
        lib = _abi.lib()
        self.compute = _stream(lib)
        self.io = _stream(lib)
        self.ens = Ensemble()
        self.ens.set_streams(self.compute, self.io)
        # small opinion requests (one-pass kernel) advance together, one
        # fused launch per step
        from . import Batch
        self.batch = Batch()
        # host seconds per scheduler phase of the last run (observability)
        self.stats = {}
        lib.enable_asserts(0)

    def run(self, requests, time_scale=1.0, cancel=None, before_admit=None):
        """Serve `requests` (arrival times in seconds, scaled by time_scale);
        returns per-request records with latency and the final-state digest.

        Every record ends with a `status`: "done" (digest, latency), "rejected"
        (its footprint exceeds the whole budget: never admitted), "failed" (its
        graph or state could not be built, e.g. out of device memory; what it
        had allocated is released) or "cancelled" (`cancel` maps a request id
        to the number of quanta it runs before it is dropped). Other requests
        are not affected by any of these. `before_admit(r)` is called just
        before a request is built (a test hook).

        One scheduler iteration: (1) enqueue `quantum` steps of every active
        request (asynchronous: opinion requests on their own streams, SIR
        requests in one graph-captured ensemble launch sequence); while those
        run, (2) read back and digest the requests that finished in the
        previous iteration and (3) admit arrivals that fit the budget, their
        graphs built on streams of their own; (4) wait for the quantum;
        (5) free what was read back (frees wait for an idle GPU) and set the
        requests that just finished aside for the next iteration's (2)."""
        vh = self.vh
        pending = sorted(requests, key=lambda r: r.arrival)
        active = {}
        finished = []
        used = 0
        done = []
        t0 = time.perf_counter()
        qi = 0
        ph = dict(enqueue_opinion=0.0, enqueue_sir=0.0, read=0.0, admit=0.0, wait=0.0, release=0.0,
                  idle=0.0, iterations=0)
        self.stats = ph
        q = self.quantum
        while qi < len(pending) or active or finished:
            # (1) one quantum for every active small request and the oldest
            # large ones
            ta = time.perf_counter()
            stepping = list(active.values())
            if self.large_concurrency is not None:
                large = sorted((r for r in stepping if self._large(r)), key=lambda r: (r.admitted, r.id))
                if len(large) > self.large_concurrency:
                    held = set(id(r) for r in large[self.large_concurrency:])
                    stepping = [r for r in stepping if id(r) not in held]
            if any(r.kind == "sir" and r.ens is self.ens for r in stepping):
                self.ens.step(q, use_graph=True)
            for r in stepping:
                if r.kind == "sir" and r.ens is not self.ens:
                    r.ens.step(min(q, r.left), use_graph=True)
            ta2 = time.perf_counter()
            small = [r for r in stepping if r.kind == "opinion" and r.n * 8 < OPINION_PLAN_MIN_BYTES]
            if small:
                m0 = small[0].model
                self.batch.apply([r.model.sim for r in small], vh.transitions.opinion(m0.E, 0.2), m0.T,
                                 [m0.T, m0.E], m0.T, steps=[min(q, r.left) for r in small])
            for r in stepping:
                if r.kind == "opinion" and r.n * 8 >= OPINION_PLAN_MIN_BYTES:
                    models.step_opinion_n(vh, r.model, min(q, r.left), eps=0.2)
            for r in stepping:
                r.left -= min(q, r.left)
            if stepping:
                ph["iterations"] += 1
            tb = time.perf_counter()
            # (2) results of the previous iteration, beside the running quantum
            for r in finished:
                if r.kind == "sir":
                    st, step = r.ens.read(r.mid)
                    assert step == r.horizon + 1
                else:
                    st, _ = r.model.sim._read_agents(r.model.T)
                r.digest = _digest(st)
                if self.keep:
                    r.result = np.array(st)
                self._end(r, "done", t0, time_scale, done)
            tc = time.perf_counter()
            # (3) admission: FIFO while the head fits the budget
            now = tc - t0
            while qi < len(pending) and pending[qi].arrival * time_scale <= now:
                r = pending[qi]
                need = footprint(r)
                if need > self.budget:
                    qi += 1
                    self._end(r, "rejected", t0, time_scale, done)
                    continue
                if used + need > self.budget and (active or finished):
                    break
                qi += 1
                r.admitted = now
                r.quanta = 0
                try:
                    if before_admit is not None:
                        before_admit(r)
                    self._build(vh, r)
                except (VahanaError, AssertionError) as ex:
                    r.error = str(ex)
                    self._end(r, "failed", t0, time_scale, done)
                    continue
                used += need
                r.left = r.horizon
                active[r.id] = r
            td = time.perf_counter()
            # (4) the quantum's end: everything stepped on the compute stream
            self.ens.sync()
            te = time.perf_counter()
            # (5) release what was read back; set aside what just finished
            for r in finished:
                self._release(r)
                used -= footprint(r)
            for r in stepping:
                r.quanta += 1
                if cancel and r.left > 0 and r.id in cancel and r.quanta >= cancel[r.id]:
                    del active[r.id]
                    self._release(r)
                    used -= footprint(r)
                    self._end(r, "cancelled", t0, time_scale, done)
            finished = [r for r in stepping if r.left == 0 and r.id in active]
            for r in finished:
                del active[r.id]
            tf = time.perf_counter()
            ph["enqueue_sir"] += ta2 - ta
            ph["enqueue_opinion"] += tb - ta2
            ph["read"] += tc - tb
            ph["admit"] += td - tc
            ph["wait"] += te - td
            ph["release"] += tf - te
            if not active and not finished and qi < len(pending):
                wait = pending[qi].arrival * time_scale - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.05))
                    ph["idle"] += min(wait, 0.05)
        wall = time.perf_counter() - t0
        return done, wall

    def _large(self, r):
        return ((r.kind == "sir" and r.ens is not self.ens) or
                (r.kind == "opinion" and r.n * 8 >= OPINION_PLAN_MIN_BYTES))

    def _build(self, vh, r):
        """Graph and state of a request on the device; on failure nothing of
        it stays allocated."""
        if r.kind == "sir":
            solo = r.n >= self.solo_agents
            ens = Ensemble() if solo else self.ens
            try:
                if solo:
                    ens.set_streams(self.compute, self.io)
                r.mid = ens.add_sir(r.n, r.mean_deg, r.p_inf, r.seed, r.seed + 1,
                                    recovery=7, n_init=max(1, r.n // 1000))
                ens.set_horizon(r.mid, r.horizon)
            except Exception:
                if solo:
                    ens.close()
                raise
            r.ens = ens
        else:
            m = None
            try:
                m = _build_opinion(vh, r, stream=self.io, keep=lambda mm: setattr(r, "_partial", mm))
                # the gather plan's host round trips on the build stream,
                # not on the step stream the other requests share
                m.sim._opinion_prepare(m.T, m.E, m.f32)
                m.sim._set_streams(self.compute, self.io)
            except Exception:
                part = getattr(r, "_partial", None)
                if part is not None:
                    part.sim._destroy()
                raise
            finally:
                if hasattr(r, "_partial"):
                    del r._partial
            r.model = m

    def _release(self, r):
        if r.kind == "sir":
            if r.ens is self.ens:
                self.ens.remove(r.mid)
            else:
                r.ens.close()
            r.ens = None
        else:
            r.model.sim._destroy()
            r.model = None

    @staticmethod
    def _end(r, status, t0, time_scale, done):
        r.status = status
        r.done = time.perf_counter() - t0
        r.latency = r.done - r.arrival * time_scale
        done.append(r)

    def close(self):
        self.batch.close()
        self.ens.close()
        lib = _abi.lib()
        lib.stream_destroy(self.io)
        lib.stream_destroy(self.compute)


def run_sequential(vh, requests, time_scale=1.0, keep_results=False, stats=None):
    """One request at a time on the GPU, in arrival order. With a `stats`
    dict, the host seconds spent building, running and reading back are
    added to it."""
    _abi.lib().enable_asserts(0)
    done = []
    ph = stats if stats is not None else {}
    for k in ("build", "run", "read", "idle"):
        ph.setdefault(k, 0.0)
    t0 = time.perf_counter()
    for r in sorted(requests, key=lambda x: x.arrival):
        now = time.perf_counter() - t0
        if r.arrival * time_scale > now:
            time.sleep(r.arrival * time_scale - now)
            ph["idle"] += r.arrival * time_scale - now
        ta = time.perf_counter()
        if r.kind == "sir":
            # the same device-built graph and launch as in the mix, alone
            e = Ensemble()
            mid = e.add_sir(r.n, r.mean_deg, r.p_inf, r.seed, r.seed + 1, recovery=7,
                            n_init=max(1, r.n // 1000))
            tb = time.perf_counter()
            e.step(r.horizon, use_graph=True)
            e.sync()
            tc = time.perf_counter()
            st, _ = e.read(mid)
            e.close()
        else:
            m = _build_opinion(vh, r)
            tb = time.perf_counter()
            models.step_opinion_n(vh, m, r.horizon, eps=0.2)
            _abi.lib().sync(m.sim.handle)
            tc = time.perf_counter()
            st, _ = m.sim._read_agents(m.T)
            m.sim._destroy()
        td = time.perf_counter()
        ph["build"] += tb - ta
        ph["run"] += tc - tb
        ph["read"] += td - tc
        r2 = types.SimpleNamespace(**{k: v for k, v in vars(r).items()})
        r2.build_s, r2.run_s, r2.read_s = tb - ta, tc - tb, td - tc
        r2.status = "done"
        r2.digest = _digest(st)
        if keep_results:
            r2.result = np.array(st)
        r2.done = time.perf_counter() - t0
        r2.latency = r2.done - r.arrival * time_scale
        done.append(r2)
    return done, time.perf_counter() - t0


def summarize(done, wall):
    """Throughput and latency percentiles of the completed requests; the
    counts of the others by status."""
    ok = [r for r in done if getattr(r, "status", "done") == "done"]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
    out = {"requests": len(ok), "wall_s": wall, "throughput_rps": len(ok) / wall,
           "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
           "agent_steps_per_s": sum(r.n * r.horizon for r in ok) / wall}
    for st in ("rejected", "failed", "cancelled"):
        k = sum(1 for r in done if getattr(r, "status", "done") == st)
        if k:
            out[st] = k
    return out
