"""Serving loop over batches of scenes (continuous batching).

Scene requests arrive over time (e.g. a Poisson process); each asks for a number of time
steps of one scene.  Requests of one kind (a parameter set: sphere scenes NoFriction,
superellipsoid scenes Friction, ...) share one resident batch (scopi_hip_batch): an
arrived request joins its kind's batch as soon as the device memory budget allows
(admission by each scene's estimated HBM footprint, first come first served), advances by
one time step per step of its batch, and leaves the batch when its steps are done, which
frees its share of the budget for the requests still waiting.  Particle state stays on the
device between steps.  The batches of different kinds step concurrently (one host thread
per kind, one HIP stream per batch), so their kernels overlap on the GPU.

The reference has no serving layer: every scene follows ScopiSolver::run
(include/scopi/solver.hpp:266-303) step by step, and a scene's results do not depend on
the other scenes of its batch (bitwise), so a served scene ends exactly where it ends when
it runs alone.

Reported per request: the latency of each of its steps (the duration of the batch step
that advanced it, packing of arrivals and departures included), its wait before admission
and its completion time; over the run: p50 / p99 step latency and scene-steps/s.
"""
from __future__ import annotations

import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    rid: int
    kind: str
    n: int              # particles asked for (the scene factory's size argument)
    seed: int
    steps: int
    arrival: float      # seconds after the start of the run
    # filled by the server
    admitted: float | None = None
    finished: float | None = None
    done_steps: int = 0
    step_latencies: list = field(default_factory=list)
    result: dict | None = None


class KindGroup:
    """The resident batch of one kind of request and its step loop."""

    def __init__(self, kind, batch):
        self.kind = kind
        self.batch = batch
        self.active = {}       # scene id -> (request, footprint)
        self.inbox = []        # admitted (request, scene, footprint), not in the batch yet
        self.steps = 0


class Server:
    """Continuous-batching request loop.

    kinds: {kind: make_batch()} -- a factory of the batch of each kind (HipBatch-like:
    add(scene) -> id, remove(id), step(n), state(id));
    make_scene(request) -> scene; footprint(request, scene) -> bytes;
    hbm_budget: bytes the resident scenes may hold together.

    The calling thread admits arrived requests (first come, first served, while the budget
    allows); every kind's batch steps in a thread of its own, taking in its admitted
    requests between two steps and releasing the budget of the finished ones.
    """

    def __init__(self, kinds, make_scene, footprint, hbm_budget, keep_results=False, clock=time.perf_counter,
                 sleep=time.sleep, on_finish=None):
        self.groups = {k: KindGroup(k, f()) for k, f in kinds.items()}
        self.make_scene = make_scene
        self.footprint = footprint
        self.budget = int(hbm_budget)
        self.keep_results = keep_results
        self.clock = clock
        self.sleep = sleep
        self.on_finish = on_finish  # called with each request when its steps are done
        self.used = 0
        self.peak_used = 0
        self.peak_active = 0
        self.ticks = 0
        self._cv = threading.Condition()
        self._closed = False
        self._error = None
        self._t0 = 0.0

    def _now(self):
        return self.clock() - self._t0

    def _group_loop(self, g):
        try:
            while True:
                with self._cv:
                    while not g.inbox and not g.active and not self._closed:
                        self._cv.wait(0.05)
                    if self._closed and not g.inbox and not g.active:
                        return
                    incoming, g.inbox = g.inbox, []
                for r, scene, fp in incoming:
                    g.active[g.batch.add(scene)] = (r, fp)
                if not g.active:
                    continue
                t = self.clock()
                g.batch.step(1)
                dt = self.clock() - t
                g.steps += 1
                now = self._now()
                freed = 0
                for sid, (r, fp) in list(g.active.items()):
                    r.step_latencies.append(dt)
                    r.done_steps += 1
                    if r.done_steps >= r.steps:
                        if self.keep_results:
                            r.result = g.batch.state(sid)
                        g.batch.remove(sid)
                        del g.active[sid]
                        freed += fp
                        r.finished = now
                        if self.on_finish:
                            self.on_finish(r)
                with self._cv:
                    self.used -= freed
                    self.ticks += 1
                    self._cv.notify_all()
        except BaseException as e:  # surfaced by run()
            with self._cv:
                self._error = e
                self._closed = True
                self._cv.notify_all()

    def run(self, requests):
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        self._t0 = self.clock()
        threads = [threading.Thread(target=self._group_loop, args=(g,), daemon=True) for g in self.groups.values()]
        for t in threads:
            t.start()
        try:
            while pending:
                r = pending[0]
                wait = r.arrival - self._now()
                if wait > 0:
                    self.sleep(min(wait, 0.05))
                    continue
                scene = self.make_scene(r)
                fp = int(self.footprint(r, scene))
                if fp > self.budget:
                    raise RuntimeError(f"request {r.rid}: footprint {fp} B exceeds the budget {self.budget} B")
                with self._cv:
                    while self.used + fp > self.budget and self._error is None:
                        self._cv.wait(0.05)  # head of the queue waits for memory
                    if self._error is not None:
                        break
                    self.used += fp
                    self.peak_used = max(self.peak_used, self.used)
                    r.admitted = self._now()
                    self.groups[r.kind].inbox.append((r, scene, fp))
                    resident = sum(len(g.active) + len(g.inbox) for g in self.groups.values())
                    self.peak_active = max(self.peak_active, resident)
                    self._cv.notify_all()
                pending.popleft()
        finally:
            with self._cv:
                self._closed = True
                self._cv.notify_all()
            for t in threads:
                t.join()
        if self._error is not None:
            raise self._error
        return stats(requests, self)


def stats(requests, server=None):
    lat = np.concatenate([np.asarray(r.step_latencies, float) for r in requests]) if requests else np.zeros(0)
    first = min(r.arrival for r in requests)
    last = max(r.finished for r in requests)
    steps = sum(r.done_steps for r in requests)
    out = {
        "requests": len(requests),
        "scene_steps": steps,
        "wall_s": last - first,
        "scene_steps_per_s": steps / (last - first) if last > first else float("nan"),
        "step_latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99)),
                            "max": 1e3 * float(lat.max())},
        "wait_ms": {"p50": 1e3 * float(np.median([r.admitted - r.arrival for r in requests])),
                    "max": 1e3 * float(max(r.admitted - r.arrival for r in requests))},
        "completion_s": {"p50": float(np.median([r.finished - r.arrival for r in requests])),
                         "max": float(max(r.finished - r.arrival for r in requests))},
    }
    if server is not None:
        out.update(ticks=server.ticks, peak_resident_scenes=server.peak_active, peak_budget_bytes=server.peak_used,
                   budget_bytes=server.budget)
    return out


def poisson_requests(n, rate, kinds, sizes, steps, seed=0):
    """n requests with exponential inter-arrival times (mean 1/rate s); kinds = [(kind, probability)],
    sizes = (lo, hi): log-uniform particle counts."""
    rng = np.random.default_rng(seed)
    t = 0.0
    names = [k for k, _ in kinds]
    prob = np.array([p for _, p in kinds], float)
    prob /= prob.sum()
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        kind = names[int(rng.choice(len(names), p=prob))]
        size = int(round(float(np.exp(rng.uniform(np.log(sizes[0]), np.log(sizes[1]))))))
        out.append(Request(rid=i, kind=kind, n=size, seed=int(rng.integers(1 << 30)), steps=steps, arrival=t))
    return out


def served_scene(r: Request):
    """The scene of a request: the 131k-sphere recipe scaled to r.n spheres (kind "spheres"), or the
    32k-superellipsoid recipe on a lattice of about r.n sites (kind "superellipsoids")."""
    from scopi_amd import scenes
    if r.kind == "spheres":
        return scenes.spheres_3d_count(r.n, seed=r.seed, exact=True)
    side = max(2, int(round(r.n ** (1.0 / 3.0))))
    return scenes.superellipsoids_3d(nx=side, ny=side, nz=max(1, int(round(r.n / (side * side)))), seed=r.seed)


def hip_kinds(dt=1e-3, sphere_precision=None, warm_start=True, schedules=None):
    """Batch factories of the two request kinds of the served workload: sphere scenes
    (NoFriction, vap_fpd, dmax 0.02, kd_tree_radius 0.042^2) and superellipsoid scenes
    (Friction mu 0.3, dmax 0.02, kd_tree_radius (2 sqrt(3) 0.015 + 0.02)^2), APGD tol 1e-6.
    With warm_start, each scene's solves start from its own last step's multipliers, kept on the
    device with its state.  schedules: optional {kind: batch schedule} (scopi_amd.tune's keys) set
    on every batch of that kind."""
    from scopi_amd import hip, scenes
    c3 = scenes.CONFIG3
    ws = 1 if warm_start else 0
    sp = hip.params(vap=hip.VAP_FPD, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-6, warm_start=ws,
                    precision=hip.FP64 if sphere_precision is None else sphere_precision)
    se = hip.params(problem=hip.FRICTION, mu=c3["mu"], vap=hip.VAP_FPD, dmax=c3["dmax"], warm_start=ws,
                    kd_tree_radius=c3["kd_tree_radius"], tolerance=c3["tolerance"])
    prm = {"spheres": sp, "superellipsoids": se}
    def factory(kind, p):
        def make():
            b = hip.HipBatch(p, dt)
            if schedules and kind in schedules:
                from scopi_amd import tune
                tune.apply_batch_schedule(b, schedules[kind])
            return b
        return make
    kinds = {k: factory(k, p) for k, p in prm.items()}
    footprint = lambda r, sc: hip.scene_footprint(sc.n, sc.n_inactive, prm[r.kind])
    return kinds, footprint, prm
