"""2D superellipsoids in the CPU restatement (closest_points.hpp:64-249, 1316-1464, 1782-1968).

Pins:
* the reference's own 2D superellipsoid pair known answers (test_closest_points.cpp:1502-1562,
  commented out there): dij = 0.2 axis-aligned and dij = 0.5 rotated by 30 degrees;
* e = 1 with equal radii is a circle: the pair equals the circle-circle closed form, and a
  sphere against it likewise;
* e = 1 ellipses against each other: the distance equals a direct minimisation of |p1 - p2| over
  both curves;
* invariants at converged ends for random squareness: both points on their implicit curves, unit
  normal, pi - pj parallel to nij;
* sphere / superellipsoid: pi lies on the superellipsoid and nij is the sphere's normal whatever the
  argument order asks (closest_points.hpp:1443-1446), the reason the reference's own test of
  this pair is marked as failing;
* superellipsoid / plane: the reference's residual uses (c_s - c_p) . n_p as a sign; with that value
  equal to -1 the Newton root is the true closest point;
* the SE2-SE2 Jacobian, written term by term as the reference's (closest_points.hpp:127-189),
  against forward-mode differentiation of F and against finite differences.
"""
import math

import numpy as np
import pytest
from scipy.optimize import minimize

SPH, PLN, SE = 0, 1, 3


def q(a):
    return np.array([math.cos(a / 2), 0.0, 0.0, math.sin(a / 2)])


@pytest.fixture(params=[0, 1], ids=["libm", "portable"])
def cp(oracle_lib, request):
    oracle_lib.lib().oracle_set_se_math(request.param)
    yield lambda *a, **k: oracle_lib.closest_points(2, *a, **k)
    oracle_lib.lib().oracle_set_se_math(0)


def rot(a):
    return np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])


def test_reference_known_answers(cp):
    r = cp(SE, (-0.2, 0.0), q(0), (0.1, 0.2), SE, (0.2, 0.0), q(0), (0.1, 0.3), sq_i=(1, 1), sq_j=(1, 1))
    np.testing.assert_allclose(r["pi"], [-0.1, 0.0], atol=1e-7)
    np.testing.assert_allclose(r["pj"], [0.1, 0.0], atol=1e-7)
    np.testing.assert_allclose(r["nij"], [-1.0, 0.0], atol=1e-7)
    assert abs(r["dij"] - 0.2) < 1e-7
    c, s, d = math.sqrt(3) / 2, 0.5, 0.4
    r = cp(SE, (-d * c, d * s), q(-math.pi / 6), (0.1, 0.4), SE, (d * c, -d * s), q(math.pi - math.pi / 6), (0.2, 0.3),
           sq_i=(1, 1), sq_j=(1, 1))
    np.testing.assert_allclose(r["pi"], [-0.3 * c, 0.3 * s], atol=1e-7)
    np.testing.assert_allclose(r["pj"], [0.2 * c, -0.2 * s], atol=1e-7)
    np.testing.assert_allclose(r["nij"], [-c, s], atol=1e-7)
    assert abs(r["dij"] - 0.5) < 1e-7


def test_circles_closed_form(cp):
    rng = np.random.default_rng(1)
    for _ in range(20):
        r1, r2 = rng.uniform(0.05, 0.3, 2)
        c1 = rng.uniform(-1, 1, 2)
        ang = rng.uniform(-math.pi, math.pi)
        dist = r1 + r2 + rng.uniform(0.01, 0.5)
        c2 = c1 + dist * np.array([math.cos(ang), math.sin(ang)])
        r = cp(SE, c1, q(rng.uniform(0, 6)), (r1, r1), SE, c2, q(rng.uniform(0, 6)), (r2, r2), sq_i=(1, 1), sq_j=(1, 1))
        u = (c2 - c1) / dist
        assert abs(r["dij"] - (dist - r1 - r2)) < 1e-7
        np.testing.assert_allclose(r["pi"], c1 + r1 * u, atol=1e-7)
        np.testing.assert_allclose(r["pj"], c2 - r2 * u, atol=1e-7)
        np.testing.assert_allclose(r["nij"], -u, atol=1e-7)
        # a sphere against the circle: pi on the circle (the superellipsoid), nij the sphere's normal.
        # newton_method stops on |F| < 1e-10 (utils.hpp:81-118), and F[1] = n1.n2 + |n1||n2| =
        # |n1||n2| (1 - cos t) is quadratic in the normals' angle t (|n1| = r1, |n2| = r2 here), so
        # t is only held to sqrt(2e-10 / (r1 r2)); the distance is second order in t
        r = cp(SPH, c2, q(0), (r2,), SE, c1, q(rng.uniform(0, 6)), (r1, r1), sq_j=(1, 1))
        t = 2 * math.sqrt(2e-10 / (r1 * r2)) + 1e-7
        assert abs(r["dij"] - (dist - r1 - r2)) < 1e-6
        np.testing.assert_allclose(r["pi"], c1 + r1 * u, atol=r1 * t)
        np.testing.assert_allclose(r["nij"], -u, atol=t)


def ellipse_point(c, rx, ry, a, t):
    return np.asarray(c) + rot(a) @ np.array([rx * math.cos(t), ry * math.sin(t)])


def test_ellipses_direct_minimisation(cp):
    rng = np.random.default_rng(2)
    checked = 0
    for _ in range(30):
        rx1, ry1, rx2, ry2 = rng.uniform(0.05, 0.25, 4)
        a1, a2 = rng.uniform(0, math.pi, 2)
        c1 = np.zeros(2)
        ang = rng.uniform(-math.pi, math.pi)
        c2 = (max(rx1, ry1) + max(rx2, ry2) + rng.uniform(0.02, 0.3)) * np.array([math.cos(ang), math.sin(ang)])
        r = cp(SE, c1, q(a1), (rx1, ry1), SE, c2, q(a2), (rx2, ry2), sq_i=(1, 1), sq_j=(1, 1))
        f = lambda x: np.linalg.norm(ellipse_point(c1, rx1, ry1, a1, x[0]) - ellipse_point(c2, rx2, ry2, a2, x[1]))
        best = min((minimize(f, [t1, t2], method="Nelder-Mead", options=dict(xatol=1e-12, fatol=1e-14, maxiter=4000))
                    for t1 in np.linspace(-3, 3, 5) for t2 in np.linspace(-3, 3, 5)), key=lambda m: m.fun)
        if abs(r["dij"] - best.fun) < 1e-6:
            checked += 1
    assert checked >= 27  # hybrd1 may end at a local solution from its grid start


def test_invariants_random_squareness(oracle_lib, cp):
    """At the ends where hybrd1 converged (info 1); most random pairs do."""
    rng = np.random.default_rng(3)
    converged = 0
    for _ in range(30):
        rx1, ry1, rx2, ry2 = rng.uniform(0.05, 0.25, 4)
        e1, e2 = rng.uniform(0.4, 1.6, 2)
        a1, a2 = rng.uniform(0, math.pi, 2)
        ang = rng.uniform(-math.pi, math.pi)
        c2 = (max(rx1, ry1) + max(rx2, ry2) + rng.uniform(0.02, 0.3)) * np.array([math.cos(ang), math.sin(ang)])
        r = cp(SE, (0, 0), q(a1), (rx1, ry1), SE, c2, q(a2), (rx2, ry2), sq_i=(e1, 1), sq_j=(e2, 1))
        full = oracle_lib.se2_closest_points((0, 0), q(a1), (rx1, ry1), e1, c2, q(a2), (rx2, ry2), e2)
        assert full["dij"] == r["dij"]
        if full["info"] != 1:
            continue
        converged += 1
        for (c, a, rx, ry, e, p) in ((np.zeros(2), a1, rx1, ry1, e1, r["pi"]), (c2, a2, rx2, ry2, e2, r["pj"])):
            b = rot(a).T @ (p - c)  # body frame
            assert abs(abs(b[0] / rx) ** (2 / e) + abs(b[1] / ry) ** (2 / e) - 1) < 1e-6
        assert abs(np.linalg.norm(r["nij"]) - 1) < 1e-12
        d = r["pi"] - r["pj"]
        assert abs(d[0] * r["nij"][1] - d[1] * r["nij"][0]) < 1e-6 * (1 + np.linalg.norm(d))
        assert abs(abs(r["dij"]) - np.linalg.norm(d)) < 1e-12
    assert converged >= 15  # hybrd1 often ends on "no progress" (info 4) near a root, as in 3D


def test_sphere_superellipsoid_point_order(cp):
    """closest_points(sphere, superellipsoid) of test_closest_points.cpp:972-997: the reference's test
    expects pi on the sphere; the function puts pi on the superellipsoid (the test is marked as failing)."""
    r = cp(SPH, (-0.2, 0.0), q(0), (0.1,), SE, (0.2, 0.0), q(0), (0.1, 0.2), sq_j=(1, 1))
    np.testing.assert_allclose(r["pi"], [0.1, 0.0], atol=1e-4)
    np.testing.assert_allclose(r["pj"], [-0.1, 0.0], atol=1e-4)
    np.testing.assert_allclose(r["nij"], [1.0, 0.0], atol=1e-4)
    assert abs(r["dij"] - 0.2) < 1e-6
    r2 = cp(SE, (0.2, 0.0), q(0), (0.1, 0.2), SPH, (-0.2, 0.0), q(0), (0.1,), sq_i=(1, 1))
    np.testing.assert_array_equal(r2["pi"], r["pj"])
    np.testing.assert_array_equal(r2["pj"], r["pi"])
    np.testing.assert_array_equal(r2["nij"], -r["nij"])


def test_superellipsoid_plane_at_unit_offset(cp):
    """With (c_s - c_p) . n_p = -1 the residual's root is the closest point: pj is pi projected on
    the plane, and pi is the superellipsoid's extreme point along n_p."""
    a, e = 0.3, 0.7
    r = cp(SE, (0, 0), q(a), (0.1, 0.2), PLN, (1.0, 0.5), q(0), (0,), sq_i=(e, 1))
    np.testing.assert_allclose(r["nij"], [1.0, 0.0])
    # F[1] is quadratic in the angle between n1 and n_p, so |F| < 1e-10 holds that angle to about
    # 1e-5 and pj's offset along the plane to |pi - pj| times it
    assert abs(r["pi"][1] - r["pj"][1]) < 1e-4 and r["pj"][0] == 1.0
    ts = np.linspace(-math.pi, math.pi, 200001)
    pts = rot(a) @ np.array([0.1 * np.sign(np.cos(ts)) * np.abs(np.cos(ts)) ** e, 0.2 * np.sign(np.sin(ts)) * np.abs(np.sin(ts)) ** e])
    assert abs(r["pi"][0] - pts[0].max()) < 1e-6
    assert abs(r["dij"] - (pts[0].max() - 1.0)) < 1e-6


def test_se2_reference_jacobian(oracle_lib):
    rng = np.random.default_rng(31)
    for portable in (0, 1):
        oracle_lib.lib().oracle_set_se_math(portable)
        try:
            for _ in range(30):
                a = ((0.0, 0.0), q(rng.uniform(-3, 3)), rng.uniform(0.05, 0.2, 2), rng.uniform(0.4, 1.6))
                b = ((rng.uniform(0.3, 0.5), rng.uniform(-0.1, 0.1)), q(rng.uniform(-3, 3)), rng.uniform(0.05, 0.2, 2),
                     rng.uniform(0.4, 1.6))
                u = rng.uniform(-3, 3, 2)
                F, J = oracle_lib.se2_residual(*a, *b, u, jacobian=True)
                oracle_lib.set_jacobian_ad(True)
                try:
                    Fa, Ja = oracle_lib.se2_residual(*a, *b, u, jacobian=True)
                finally:
                    oracle_lib.set_jacobian_ad(False)
                assert np.array_equal(F, Fa)
                assert np.abs(J - Ja).max() <= 1e-12 * np.abs(Ja).max()
                h = 1e-7
                for k in range(2):
                    e = np.zeros(2); e[k] = h
                    fd = (oracle_lib.se2_residual(*a, *b, u + e) - oracle_lib.se2_residual(*a, *b, u - e)) / (2 * h)
                    assert np.allclose(J[:, k], fd, rtol=1e-5, atol=1e-9 * np.abs(J).max())
        finally:
            oracle_lib.lib().oracle_set_se_math(0)
