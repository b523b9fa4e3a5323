"""Serving loop: a request stream of SIR and opinion-dynamics simulations.

Requests arrive over time (Poisson), are admitted first-come-first-served
while their estimated device footprint fits the HBM budget, keep their graph
and state on the device between steps, and leave when their horizon is done.
Continuous batching: every scheduler iteration advances all SIR requests by
one segmented, graph-captured ensemble launch per step and every opinion
request by apply! calls on its own stream, so their work overlaps on the GPU;
requests that arrive meanwhile join at the next iteration as budget frees up.

`run_sequential` is the one-at-a-time baseline: the same requests, each built,
run to its horizon and read back before the next one starts.
"""
import hashlib
import time
import types

import numpy as np

from . import _abi, inputs, models
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


def footprint(req):
    """Device bytes a request holds while active (state, CSR, build scratch)."""
    if req.kind == "sir":
        e = req.n * req.mean_deg
        return int(req.n * 12 + e * 4 + req.n * 8)
    e = 2 * req.n * 5
    return int(req.n * 20 + e * 8 + e * 48)


def _opinion_edges(req):
    fr, to = inputs.graph_gnp(req.n, 5.0, req.seed)
    return inputs.symmetric_weighted(fr, to, req.seed + 7)


def _digest(arr):
    return hashlib.sha256(np.ascontiguousarray(arr).view(np.uint8)).hexdigest()


class Server:
    def __init__(self, vh, hbm_budget, quantum=10, keep_results=False):
        self.vh = vh
        self.budget = int(hbm_budget)
        self.quantum = quantum
        self.keep = keep_results
        self.ens = Ensemble()
        _abi.lib().enable_asserts(0)

    def run(self, requests, time_scale=1.0):
        """Serve `requests` (arrival times in seconds, scaled by time_scale);
        returns per-request records with latency and the final-state digest."""
        vh = self.vh
        pending = sorted(requests, key=lambda r: r.arrival)
        active = {}
        used = 0
        done = []
        t0 = time.perf_counter()
        qi = 0
        while qi < len(pending) or active:
            now = time.perf_counter() - t0
            # admission: FIFO while the head fits the budget
            while qi < len(pending) and pending[qi].arrival * time_scale <= now:
                r = pending[qi]
                need = footprint(r)
                if used + need > self.budget and active:
                    break
                qi += 1
                used += need
                r.admitted = now
                if r.kind == "sir":
                    r.mid = self.ens.add_sir(r.n, r.mean_deg, r.p_inf, r.seed, r.seed + 1,
                                             recovery=7, n_init=max(1, r.n // 1000))
                    self.ens.set_horizon(r.mid, r.horizon)
                    r.left = r.horizon
                else:
                    r.model = models.build_opinion(vh, r.n, seed=r.seed, edges=_opinion_edges(r))
                    r.left = r.horizon
                active[r.id] = r
            if not active:
                if qi < len(pending):
                    wait = pending[qi].arrival * time_scale - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.05))
                continue
            # one scheduler iteration: `q` steps for everybody (ensemble members
            # stop exactly at their horizon inside the launch)
            q = self.quantum
            # opinion steps are enqueued first (asynchronous, one stream per
            # simulation), the ensemble launch then overlaps with them
            for r in active.values():
                if r.kind == "opinion":
                    for _ in range(min(q, r.left)):
                        models.step_opinion(vh, r.model, eps=0.2)
            if any(r.kind == "sir" for r in active.values()):
                self.ens.step(q, use_graph=True)
            for r in active.values():
                r.left -= min(q, r.left)
            self.ens.sync()
            for r in list(active.values()):
                if r.kind == "opinion":
                    _abi.lib().sync(r.model.sim.handle)
            now = time.perf_counter() - t0
            for r in list(active.values()):
                if r.left > 0:
                    continue
                if r.kind == "sir":
                    st, step = self.ens.read(r.mid)
                    assert step == r.horizon + 1
                    self.ens.remove(r.mid)
                else:
                    st, _ = r.model.sim._read_agents(r.model.T)
                    r.model.sim._destroy()
                    r.model = None
                r.digest = _digest(st)
                if self.keep:
                    r.result = np.array(st)
                r.done = now
                r.latency = now - r.arrival * time_scale
                used -= footprint(r)
                done.append(r)
                del active[r.id]
        wall = time.perf_counter() - t0
        return done, wall

    def close(self):
        self.ens.close()


def run_sequential(vh, requests, time_scale=1.0, keep_results=False):
    """One request at a time on the GPU, in arrival order."""
    _abi.lib().enable_asserts(0)
    done = []
    t0 = time.perf_counter()
    for r in sorted(requests, key=lambda x: x.arrival):
        now = time.perf_counter() - t0
        if r.arrival * time_scale > now:
            time.sleep(r.arrival * time_scale - now)
        if r.kind == "sir":
            # the same device-built graph and launch as in the mix, alone
            e = Ensemble()
            mid = e.add_sir(r.n, r.mean_deg, r.p_inf, r.seed, r.seed + 1, recovery=7,
                            n_init=max(1, r.n // 1000))
            e.step(r.horizon, use_graph=True)
            st, _ = e.read(mid)
            e.close()
        else:
            m = models.build_opinion(vh, r.n, seed=r.seed, edges=_opinion_edges(r))
            for _ in range(r.horizon):
                models.step_opinion(vh, m, eps=0.2)
            st, _ = m.sim._read_agents(m.T)
            m.sim._destroy()
        r2 = types.SimpleNamespace(**{k: v for k, v in vars(r).items()})
        r2.digest = _digest(st)
        if keep_results:
            r2.result = np.array(st)
        r2.done = time.perf_counter() - t0
        r2.latency = r2.done - r.arrival * time_scale
        done.append(r2)
    return done, time.perf_counter() - t0


def summarize(done, wall):
    lat = np.array([r.latency for r in done])
    return {"requests": len(done), "wall_s": wall, "throughput_rps": len(done) / wall,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "agent_steps_per_s": sum(r.n * r.horizon for r in done) / wall}
