"""The serving loop on the GPU: requests of two kinds (sphere scenes NoFriction, superellipsoid
scenes Friction mu 0.3) arrive while others are stepping, share resident batches, and leave
when done.  Bar: every served scene's final state is bitwise the state of the same scene
stepped alone through the one-scene context for the same number of steps."""
import sys

import numpy as np
import pytest

from scopi_amd import hip, serve

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone():
    kinds, footprint, prm = serve.hip_kinds()
    reqs = [serve.Request(rid=0, kind="spheres", n=700, seed=1, steps=4, arrival=0.0),
            serve.Request(rid=1, kind="superellipsoids", n=125, seed=2, steps=3, arrival=0.0),
            serve.Request(rid=2, kind="spheres", n=1500, seed=3, steps=2, arrival=0.05),
            serve.Request(rid=3, kind="spheres", n=400, seed=4, steps=5, arrival=0.1),
            serve.Request(rid=4, kind="superellipsoids", n=216, seed=5, steps=2, arrival=0.15)]
    print(f"[serve] {len(reqs)} requests", flush=True)
    srv = serve.Server(kinds, serve.served_scene, footprint, hbm_budget=8 << 30, keep_results=True)
    st = srv.run(reqs)
    assert st["scene_steps"] == sum(r.steps for r in reqs)
    assert st["peak_resident_scenes"] >= 2
    for r in reqs:
        sc = serve.served_scene(r)
        g = hip.HipScene(sc)
        g.run(prm[r.kind], 1e-3, r.steps)
        alone = g.state()
        sys.stdout.flush()
        for k in ("pos", "q", "v", "omega"):
            assert np.array_equal(alone[k], r.result[k]), f"request {r.rid} ({r.kind}): {k} differs"


def test_superellipsoid_batch_bitwise_one_scene():
    kinds, footprint, prm = serve.hip_kinds()
    reqs = [serve.Request(rid=k, kind="superellipsoids", n=n, seed=10 + k, steps=2, arrival=0.0)
            for k, n in enumerate([125, 216, 343])]
    b = kinds["superellipsoids"]()
    ids = [b.add(serve.served_scene(r)) for r in reqs]
    b.step(2)
    for sid, r in zip(ids, reqs):
        g = hip.HipScene(serve.served_scene(r))
        g.run(prm["superellipsoids"], 1e-3, 2)
        gc, bc = g.contacts(), b.contacts(sid)
        assert len(gc["ij"]) > 0
        for k in gc:
            assert np.array_equal(gc[k], bc[k]), f"n={r.n}: contacts {k}"
        gs, bs = g.solution(), b.solution(sid)
        for k in ("u", "w", "lam"):
            assert np.array_equal(gs[k], bs[k]), f"n={r.n}: solution {k}"
