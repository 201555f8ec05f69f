"""Host logic of the serving loop (scopi_amd.serve) on the CPU: admission by memory budget,
first come first served, every request stepped exactly its number of steps, latencies
recorded per step.  A stand-in batch and clock replace the GPU (no compute call)."""
import pytest

from scopi_amd import serve


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-6)


class FakeBatch:
    def __init__(self, clock, step_s, log):
        self.clock, self.step_s, self.log = clock, step_s, log
        self.scenes, self.next = {}, 1

    def add(self, scene):
        sid = self.next
        self.next += 1
        self.scenes[sid] = dict(scene=scene, steps=0)
        self.log.append(("add", scene["rid"]))
        return sid

    def remove(self, sid):
        self.log.append(("remove", self.scenes[sid]["scene"]["rid"]))
        del self.scenes[sid]

    def step(self, n):
        for v in self.scenes.values():
            v["steps"] += n
        self.clock.t += self.step_s * max(1, len(self.scenes)) * n

    def state(self, sid):
        return dict(steps=self.scenes[sid]["steps"])


def make_server(budget, log, clock, step_s=0.01):
    kinds = {"a": lambda: FakeBatch(clock, step_s, log), "b": lambda: FakeBatch(clock, step_s, log)}
    return serve.Server(kinds, make_scene=lambda r: dict(rid=r.rid, n=r.n), footprint=lambda r, sc: r.n,
                        hbm_budget=budget, keep_results=True, clock=clock, sleep=clock.sleep)


def test_kinds_step_independently():
    """A slow kind does not hold back a fast one: each kind's batch steps in its own loop."""
    import threading
    import time

    class SlowBatch(FakeBatch):
        def step(self, n):
            time.sleep(0.05)
            super().step(n)

    clock, log = FakeClock(), []
    kinds = {"fast": lambda: FakeBatch(clock, 0.0, log), "slow": lambda: SlowBatch(clock, 0.0, log)}
    srv = serve.Server(kinds, make_scene=lambda r: dict(rid=r.rid, n=r.n), footprint=lambda r, sc: r.n,
                       hbm_budget=1000, clock=time.perf_counter)
    reqs = [serve.Request(rid=0, kind="slow", n=1, seed=0, steps=10, arrival=0.0),
            serve.Request(rid=1, kind="fast", n=1, seed=0, steps=10, arrival=0.0)]
    srv.run(reqs)
    assert reqs[1].finished < 0.5 * reqs[0].finished


def test_every_request_gets_its_steps_and_results():
    clock, log = FakeClock(), []
    reqs = serve.poisson_requests(20, rate=50.0, kinds=[("a", 0.8), ("b", 0.2)], sizes=(10, 100), steps=5, seed=3)
    srv = make_server(10_000, log, clock)
    st = srv.run(reqs)
    for r in reqs:
        assert r.done_steps == 5 and len(r.step_latencies) == 5
        assert r.result == dict(steps=5)
        assert r.arrival <= r.admitted <= r.finished
    assert st["scene_steps"] == 100 and st["requests"] == 20
    assert st["step_latency_ms"]["p99"] >= st["step_latency_ms"]["p50"] > 0
    assert sum(1 for e in log if e[0] == "add") == 20 and sum(1 for e in log if e[0] == "remove") == 20


def test_budget_bounds_resident_scenes_fifo():
    clock, log = FakeClock(), []
    # every request needs 40 of a budget of 100: at most two resident at once
    reqs = [serve.Request(rid=i, kind="a" if i % 2 else "b", n=40, seed=i, steps=3, arrival=0.001 * i)
            for i in range(8)]
    srv = make_server(100, log, clock)
    st = srv.run(reqs)
    assert st["peak_resident_scenes"] <= 2 and srv.peak_used <= 100
    adm = [r.admitted for r in reqs]
    assert adm == sorted(adm)  # admitted in arrival order
    assert reqs[-1].admitted > reqs[-1].arrival  # the last ones waited for memory


def test_oversized_request_is_refused():
    clock, log = FakeClock(), []
    srv = make_server(50, log, clock)
    with pytest.raises(RuntimeError):
        srv.run([serve.Request(rid=0, kind="a", n=60, seed=0, steps=1, arrival=0.0)])


def test_poisson_requests_shape():
    reqs = serve.poisson_requests(2000, rate=10.0, kinds=[("s", 0.8), ("e", 0.2)], sizes=(512, 65536), steps=100,
                                  seed=1)
    arr = [r.arrival for r in reqs]
    assert arr == sorted(arr)
    gaps = [b - a for a, b in zip([0.0] + arr[:-1], arr)]
    assert abs(sum(gaps) / len(gaps) - 0.1) < 0.01
    frac = sum(r.kind == "e" for r in reqs) / len(reqs)
    assert 0.17 < frac < 0.23
    assert min(r.n for r in reqs) >= 512 and max(r.n for r in reqs) <= 65536
