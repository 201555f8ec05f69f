"""Batches of independent scenes (scopi_hip_batch_*) on the GPU, through the C ABI.

Bars:
  * a scene's contacts in a batch are bitwise those of the same scene stepped alone
    through the one-scene context (which is bit-exact against the oracle,
    test_gpu_parity.py), and bit-exact against the oracle directly on a small scene;
  * a scene's trajectory in a batch is bitwise the one-scene context's, FP64 and
    FP32 + FP64 refinement, reference defaults (adaptive line search): a scene's share
    of the batch's grids follows the one-scene rule, so it runs the same arithmetic;
  * a scene's trajectory is bitwise the same alone and in any mix, across scenes added
    and removed between steps.
"""
import sys

import numpy as np
import pytest

from scopi_amd import hip, scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def prm_c2(**over):
    kw = dict(vap=hip.VAP_FPD, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-6)
    kw.update(over)
    return hip.params(**kw)


SIZES = [300, 1234, 2500, 777, 4096]


def mix(sizes=SIZES, seed=5):
    return [scenes.spheres_3d_count(n, seed=seed + k, exact=True) for k, n in enumerate(sizes)]


def same_state(a, b, tag):
    for k in ("pos", "q", "v", "omega"):
        assert np.array_equal(a[k], b[k]), f"{tag}: {k} differs (max {np.abs(a[k] - b[k]).max():.3e})"


@pytest.mark.parametrize("method", [hip.KDTREE, hip.BRUTE_FORCE], ids=["kdtree", "brute_force"])
def test_batch_contacts_match_one_scene(method):
    prm = prm_c2(contact_method=method)
    scs = mix()
    say(f"[batch] contacts method={method} sizes={SIZES}")
    b = hip.HipBatch(prm, 1e-3)
    ids = [b.add(sc) for sc in scs]
    b.step(1)
    for sid, sc in zip(ids, scs):
        g = hip.HipScene(sc)
        g.run(prm, 1e-3, 1)
        gc, bc = g.contacts(), b.contacts(sid)
        assert len(gc["ij"]) > 0
        for k in gc:
            assert np.array_equal(gc[k], bc[k]), f"scene n={sc.n}: {k} differs"


def test_batch_contacts_against_oracle(oracle_lib):
    prm = prm_c2()
    scs = mix([500, 900])
    b = hip.HipBatch(prm, 1e-3)
    ids = [b.add(sc) for sc in scs]
    b.step(1)
    for sid, sc in zip(ids, scs):
        o = oracle_lib.Oracle(sc, problem=prm.problem, contact_method=prm.contact_method, vap=prm.vap,
                              method=prm.method, dmax=prm.dmax, kd_tree_radius=prm.kd_tree_radius, alpha=prm.alpha,
                              max_ite=prm.max_ite, tolerance=prm.tolerance, dynamic_descent=prm.dynamic_descent,
                              mu=prm.mu)
        o.compute_contacts()
        oc, bc = o.contacts(), b.contacts(sid)
        assert np.array_equal(oc["ij"], bc["ij"])
        for k in ("dij", "nij", "pi", "pj"):
            np.testing.assert_allclose(bc[k], oc[k], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("precision", [hip.FP64, hip.FP32_REFINE], ids=["fp64", "fp32_refine"])
def test_batch_trajectory_bitwise_one_scene(precision):
    prm = prm_c2(precision=precision)
    scs = mix()
    say(f"[batch] trajectory precision={precision}")
    b = hip.HipBatch(prm, 1e-3)
    ids = [b.add(sc) for sc in scs]
    b.step(3)
    for sid, sc in zip(ids, scs):
        g = hip.HipScene(sc)
        info = g.run(prm, 1e-3, 3)
        bi = b.info(sid)
        assert bi["steps"] == 3 and bi["converged"] == 1
        assert bi["n_contacts"] == g.nc
        same_state(g.state(), b.state(sid), f"n={sc.n}")
        gs, bs = g.solution(), b.solution(sid)
        for k in ("u", "w", "lam"):
            assert np.array_equal(gs[k], bs[k]), f"n={sc.n}: solution {k} differs"
        assert info["converged"] == 1


def test_batch_scene_independent_of_mix():
    prm = prm_c2(precision=hip.FP32_REFINE)
    scs = mix()
    target = scs[2]
    say("[batch] alone vs mix, with additions and removals")
    alone = hip.HipBatch(prm, 1e-3)
    a = alone.add(target)
    alone.step(4)
    m = hip.HipBatch(prm, 1e-3)
    ids = [m.add(sc) for sc in scs[:2]]
    t = m.add(target)
    m.step(2)
    m.remove(ids[0])
    extra = [m.add(sc) for sc in mix([640, 2000], seed=11)]
    m.step(1)
    m.remove(extra[1])
    m.step(1)
    same_state(alone.state(a), m.state(t), "alone vs mix")
    ia, im = alone.info(a), m.info(t)
    for k in ("iterations", "backtracks", "n_contacts", "steps"):
        assert ia[k] == im[k], k
    assert m.size()["scenes"] == 3


def test_batch_scene_without_contacts():
    """A scene without contacts (spheres far apart, no obstacles, no force) shares a batch."""
    prm = prm_c2()
    sc = scenes.spheres_3d_count(64, seed=3, exact=True)
    ni = sc.n_inactive
    far = scenes.Scene(3, 0, *(getattr(sc, k)[ni:].copy() for k in (
        "shape", "pos", "q", "size", "squareness", "v", "vd", "omega", "domega", "f", "m", "J")))
    far.pos *= 40.0
    far.f[:] = 0.0
    b = hip.HipBatch(prm, 1e-3)
    i1 = b.add(sc)
    i2 = b.add(far)
    b.step(2)
    assert b.info(i2)["n_contacts"] == 0 and b.info(i2)["converged"] == 1
    st = b.state(i2)
    np.testing.assert_array_equal(st["v"], far.v)
    g = hip.HipScene(sc)
    g.run(prm, 1e-3, 2)
    same_state(g.state(), b.state(i1), "small scene")
