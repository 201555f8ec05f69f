"""2D superellipsoids on the GPU (closest_points.hpp:64-249, 1316-1464, 1782-1968): every pair kind
and whole contact sets bit for bit against the oracle with its portable elementary functions (the
device evaluates the same operations; the hybrd1 solve is chaotic at the rounding level, so the
bar is equality), and a step's solve on those contacts against the oracle."""
import math
import sys

import numpy as np
import pytest

from scopi_amd import hip
from scopi_amd.scenes import Container, Property

pytestmark = pytest.mark.gpu
SPH, PLN, SE = hip.SHAPE_SPHERE, hip.SHAPE_PLANE, hip.SHAPE_SUPERELLIPSOID


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def q(a):
    return np.array([math.cos(a / 2), 0.0, 0.0, math.sin(a / 2)])


@pytest.fixture
def se_oracle(oracle_lib):
    oracle_lib.set_se_math(True)
    yield oracle_lib
    oracle_lib.set_se_math(False)


def se2_scene(n=14, seed=2, with_spheres=True):
    """Superellipsoids (and spheres) in two rows over a floor plane, a plane wall at x = -0.05."""
    rng = np.random.default_rng(seed)
    c = Container(2)
    c.plane((0.0, 0.0), math.pi / 2)
    c.plane((-0.05, 0.0), 0.0)
    for k in range(n):
        pos = (0.2 * (k % 7) + 0.1 + rng.uniform(-0.02, 0.02), 0.1 + 0.18 * (k // 7) + rng.uniform(0, 0.02))
        prop = Property(mass=1.0, moment_inertia=0.01, force=(0, -10), velocity=(rng.uniform(-1, 1), 0))
        if with_spheres and k % 3 == 2:
            c.sphere(pos, 0.07, prop)
        else:
            c.superellipsoid(pos, q(rng.uniform(0, 3)), (0.1, 0.06, 0.0), (rng.uniform(0.5, 1.5), 1.0), prop)
    return c.arrays()


@pytest.fixture(scope="module")
def ctx2():
    g = hip.HipScene(se2_scene(3))
    yield g
    g.close()


KINDS = [(SE, SE), (SPH, SE), (SE, SPH), (SE, PLN), (PLN, SE)]


@pytest.mark.parametrize("si,sj", KINDS, ids=["se-se", "sphere-se", "se-sphere", "se-plane", "plane-se"])
def test_pairs_bit_exact(se_oracle, ctx2, si, sj):
    rng = np.random.default_rng(si * 7 + sj)
    n = 40
    pos_i = np.zeros((n, 3)); pos_j = np.zeros((n, 3))
    pos_i[:, :2] = rng.uniform(-1, 1, (n, 2))
    ang = rng.uniform(-math.pi, math.pi, n)
    pos_j[:, :2] = pos_i[:, :2] + rng.uniform(0.2, 0.5, (n, 1)) * np.stack([np.cos(ang), np.sin(ang)], 1)
    qi = np.array([q(a) for a in rng.uniform(0, 6, n)]); qj = np.array([q(a) for a in rng.uniform(0, 6, n)])
    size_i = np.zeros((n, 3)); size_j = np.zeros((n, 3))
    size_i[:, :2] = rng.uniform(0.05, 0.15, (n, 2)); size_j[:, :2] = rng.uniform(0.05, 0.15, (n, 2))
    sq_i = np.ones((n, 2)); sq_j = np.ones((n, 2))
    sq_i[:, 0] = rng.uniform(0.5, 1.5, n); sq_j[:, 0] = rng.uniform(0.5, 1.5, n)
    say(f"[se2-pairs] {si}-{sj}")
    g = ctx2.closest_points(np.full(n, si), pos_i, qi, size_i, sq_i, np.full(n, sj), pos_j, qj, size_j, sq_j)
    for k in range(n):
        o = se_oracle.closest_points(2, si, pos_i[k, :2], qi[k], size_i[k, :2], sj, pos_j[k, :2], qj[k], size_j[k, :2],
                                     sq_i=sq_i[k], sq_j=sq_j[k])
        assert g["dij"][k] == o["dij"], k
        for key in ("nij", "pi", "pj"):
            assert np.array_equal(g[key][k, :2], o[key][:2]), (k, key)


@pytest.mark.parametrize("method", [hip.KDTREE, hip.BRUTE_FORCE], ids=["kdtree", "brute_force"])
def test_contact_set_bit_exact(se_oracle, method):
    sc = se2_scene()
    prm = hip.params(dmax=0.05, kd_tree_radius=0.3 ** 2, contact_method=method)
    say(f"[se2-contacts] method={method} n={sc.n}")
    g = hip.HipScene(sc)
    nc = g.compute_contacts(prm)
    gc = g.contacts()
    o = se_oracle.Oracle(sc, dmax=0.05, kd_tree_radius=0.3 ** 2, contact_method=method)
    assert o.compute_contacts() == nc and nc > 0
    oc = o.contacts()
    assert np.array_equal(gc["ij"], oc["ij"])
    for k in ("dij", "nij", "pi", "pj"):
        assert np.array_equal(gc[k], oc[k]), k
    kinds = {(int(sc.shape[i]), int(sc.shape[j])) for i, j in gc["ij"]}
    say(f"  nc={nc} pair kinds={sorted(kinds)}")
    assert (SE, SE) in kinds or (SE, SPH) in kinds


def test_step_against_oracle(se_oracle):
    """One step with the fixed-step projected gradient (alpha below 1/L, tol 1e-12) on the contact
    set that the tests above find bit-exact: positions within 1e-9, rotation velocities within 1e-7.
    Longer runs are not compared: the superellipsoid / plane Newton of the reference does not
    converge for these scenes (its residual treats (c_s - c_p) . n_p as a sign), so its end point,
    and the trajectory after it, changes with the last bit of the input."""
    sc = se2_scene()
    kw = dict(vap=hip.VAP_FPD, dmax=0.05, kd_tree_radius=0.3 ** 2, dynamic_descent=0, alpha=20.0, tolerance=1e-12,
              max_ite=200000)
    g = hip.HipScene(sc)
    info = g.run(hip.params(**kw), 0.005, 1)
    o = se_oracle.Oracle(sc, **kw)
    o.run(0.005, 1)
    d = float(np.abs(g.state()["pos"] - o.state()["pos"]).max())
    say(f"[se2-step] contacts={info['n_contacts']} max|dpos|={d:.3e}")
    assert info["n_contacts"] > 0
    np.testing.assert_allclose(g.state()["pos"], o.state()["pos"], atol=1e-9)
    np.testing.assert_allclose(g.state()["omega"], o.state()["omega"], atol=1e-7)
