"""Periodic boxes on the GPU (scopi_hip_set_box) against the oracle's restatement of the
reference's periodic rules (tests/test_oracle_periodic.py pins it).

Bars: the contact set (pairs, order) bit-exact, distances / normals / points within 1e-12
relative, as for every other contact set (test_gpu_parity.py); trajectories with the fixed-step
APGD (no line search) within 1e-9 in positions over 30 steps, wraps across the box included.
"""
import math
import sys

import numpy as np
import pytest

from scopi_amd import hip
from scopi_amd.scenes import Container, Property, quaternion

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def periodic_scene(dim, n, seed, planes=False, speed=0.5):
    rng = np.random.default_rng(seed)
    c = Container(dim)
    if planes:
        if dim == 2:
            c.plane((0.5, 0.01), math.pi / 2)  # near y = 0: copied when y is periodic
            c.plane((0.3, 0.5), 0.0)           # x-normal plane through the box
        else:
            c.plane((0.5, 0.5, 0.01), q=quaternion(-math.pi / 2, (0.0, 1.0, 0.0)))
            c.plane((0.3, 0.5, 0.5), q=quaternion(0.0))
    for _ in range(n):
        p = rng.uniform(0.0, 1.0, size=dim)
        c.sphere(p, rng.uniform(0.015, 0.03),
                 Property(mass=1.0, moment_inertia=0.1 if dim == 2 else (0.1, 0.1, 0.1),
                          velocity=rng.uniform(-speed, speed, size=dim),
                          desired_velocity=rng.uniform(-speed, speed, size=dim)))
    return c.arrays()


CASES = {
    "2d_x": (2, 300, [True, False], False),
    "2d_xy": (2, 300, [True, True], False),
    "2d_xy_planes": (2, 300, [True, True], True),
    "3d_x": (3, 800, [True, False, False], False),
    "3d_xyz": (3, 800, [True, True, True], False),
    "3d_yz_planes": (3, 800, [False, True, True], True),
}


def oracle_kw(prm):
    return dict(problem=prm.problem, contact_method=prm.contact_method, vap=prm.vap, method=prm.method,
                dmax=prm.dmax, kd_tree_radius=prm.kd_tree_radius, alpha=prm.alpha, max_ite=prm.max_ite,
                tolerance=prm.tolerance, dynamic_descent=prm.dynamic_descent, mu=prm.mu)


@pytest.mark.parametrize("method", [hip.KDTREE, hip.BRUTE_FORCE], ids=["kdtree", "brute_force"])
@pytest.mark.parametrize("case", list(CASES))
def test_periodic_contacts(oracle_lib, case, method):
    dim, n, per, planes = CASES[case]
    sc = periodic_scene(dim, n, 7, planes)
    lo, hi = [0.0] * dim, [1.0] * dim
    prm = hip.params(contact_method=method, dmax=0.04, kd_tree_radius=0.12 ** 2)
    say(f"[periodic-contacts] {case} method={method}")
    g = hip.HipScene(sc)
    g.set_box(lo, hi, per)
    nc = g.compute_contacts(prm)
    o = oracle_lib.Oracle(sc, **oracle_kw(prm))
    o.set_box(lo, hi, per)
    o.compute_contacts()
    gc, oc = g.contacts(), o.contacts()
    assert nc > 0
    assert gc["ij"].shape == oc["ij"].shape, f"{len(gc['ij'])} vs {len(oc['ij'])}"
    assert np.array_equal(gc["ij"], oc["ij"])
    assert np.any(gc["ij"][:, 0] > gc["ij"][:, 1]) or not any(per[:dim])  # pairs through a copy
    for k in ("dij", "nij", "pi", "pj"):
        np.testing.assert_allclose(gc[k], oc[k], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("case", ["2d_xy", "3d_xyz"])
def test_periodic_trajectory_fixed_step(oracle_lib, case):
    dim, n, per, planes = CASES[case]
    sc = periodic_scene(dim, n // 2, 11, planes, speed=2.0)
    lo, hi = [0.0] * dim, [1.0] * dim
    prm = hip.params(vap=hip.VAP_FIXED, dmax=0.04, kd_tree_radius=0.12 ** 2, dynamic_descent=0, alpha=50.0,
                     tolerance=1e-10, max_ite=200000)
    say(f"[periodic-trajectory] {case}")
    g = hip.HipScene(sc)
    g.set_box(lo, hi, per)
    g.run(prm, 0.01, 30)
    o = oracle_lib.Oracle(sc, **oracle_kw(prm))
    o.set_box(lo, hi, per)
    o.run(0.01, 30)
    sg, so = g.state(), o.state()
    # positions modulo the box along periodic axes: an object within rounding of a bound may be
    # wrapped by one code and not by the other, which moves it by exactly L
    d = sg["pos"] - so["pos"]
    for a in range(dim):
        if per[a]:
            d[:, a] -= (hi[a] - lo[a]) * np.round(d[:, a] / (hi[a] - lo[a]))
    err = float(np.abs(d).max())
    say(f"  max|dpos| modulo the box={err:.3e}")
    assert err <= 1e-9
    act = sc.n_inactive
    assert np.all(sg["pos"][act:, :dim] >= -0.1) and np.all(sg["pos"][act:, :dim] <= 1.1)


def test_periodic_box_rejects_superellipsoids():
    c = Container(3)
    c.superellipsoid((0.5, 0.5, 0.5), quaternion(0.0), (0.02, 0.02, 0.02), (1.0, 1.0),
                     Property(mass=1.0, moment_inertia=(0.1, 0.1, 0.1)))
    g = hip.HipScene(c.arrays())
    with pytest.raises(hip.ScopiHipError):
        g.set_box([0, 0, 0], [1, 1, 1], [True, False, False])
