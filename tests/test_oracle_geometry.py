"""Pins the CPU restatement's closest points, contact set and A / A^T.

Known answers come from test/test_closest_points.cpp, test/test_contacts_kdtree.cpp,
test/test_contacts_brute_force.cpp and test/test_matrices.cpp; the rest are
invariants the mathematics fixes (brute-force set equality, adjointness, and
A u = relative velocity of the contact points under rigid motion).
"""
import math

import numpy as np
import pytest

from scopi_amd.scenes import Container, Property, quaternion

S, P = 0, 1
c30, s30 = math.sqrt(3.0) / 2.0, 0.5
PI = math.pi

# (name, dim, (shape, pos, angle, radius) i, same for j, pi, pj, nij, dij)
KNOWN = [
    ("sphere_sphere_2d", 2, (S, (-0.2, 0.0), 0.0, 0.1), (S, (0.2, 0.0), 0.0, 0.1), (-0.1, 0.0), (0.1, 0.0), (-1.0, 0.0), 0.2),
    ("sphere_sphere_2d_rotation_30_deg", 2, (S, (-0.3 * c30, 0.3 * s30), 0.0, 0.1), (S, (0.3 * c30, -0.3 * s30), 0.0, 0.1),
     (-0.2 * c30, 0.2 * s30), (0.2 * c30, -0.2 * s30), (-c30, s30), 0.4),
    ("sphere_sphere_3d", 3, (S, (-0.2, 0.0, 0.0), 0.0, 0.1), (S, (0.2, 0.0, 0.0), 0.0, 0.1), (-0.1, 0, 0), (0.1, 0, 0),
     (-1.0, 0, 0), 0.2),
    ("sphere_plane_2d", 2, (S, (0.0, 0.0), 0.0, 0.1), (P, (0.3, 0.0), 0.0, 0.0), (0.1, 0.0), (0.3, 0.0), (-1.0, 0.0), 0.2),
    ("sphere_plane_2d_rotation_30_deg", 2, (S, (0.0, 0.0), 0.0, 0.1), (P, (0.3 * c30, 0.3 * s30), PI / 6, 0.0),
     (0.1 * c30, 0.1 * s30), (0.3 * c30, 0.3 * s30), (-c30, -s30), 0.2),
    ("sphere_plane_2d_rotation_90_deg", 2, (S, (0.0, 0.0), 0.0, 0.1), (P, (0.0, -0.2), PI / 2, 0.0), (0.0, -0.1),
     (0.0, -0.2), (0.0, 1.0), 0.1),
    ("sphere_plane_3d", 3, (S, (0, 0, 0), 0.0, 0.1), (P, (0.3, 0, 0), 0.0, 0.0), (0.1, 0, 0), (0.3, 0, 0), (-1.0, 0, 0), 0.2),
    ("sphere_plane_3d_rotation_30_deg", 3, (S, (0, 0, 0), 0.0, 0.1), (P, (0.3 * c30, 0.3 * s30, 0.0), PI / 6, 0.0),
     (0.1 * c30, 0.1 * s30, 0), (0.3 * c30, 0.3 * s30, 0), (-c30, -s30, 0), 0.2),
    ("sphere_plane_3d_rotation_90_deg", 3, (S, (0, 0, 0), 0.0, 0.1), (P, (0, -0.2, 0), PI / 2, 0.0), (0, -0.1, 0),
     (0, -0.2, 0), (0, 1.0, 0), 0.1),
    ("plane_sphere_3d_rotation_30_deg", 3, (P, (0.3 * c30, 0.3 * s30, 0.0), PI / 6, 0.0), (S, (0, 0, 0), 0.0, 0.1),
     (0.3 * c30, 0.3 * s30, 0), (0.1 * c30, 0.1 * s30, 0), (c30, s30, 0), 0.2),
]


def approx(x, ref):
    # doctest::Approx default: |x - ref| < 100 eps_f * (1 + max(|x|, |ref|))
    return abs(x - ref) < np.finfo(np.float32).eps * 100 * (1.0 + max(abs(x), abs(ref)))


@pytest.mark.parametrize("case", KNOWN, ids=[k[0] for k in KNOWN])
def test_closest_points_known_answers(oracle_lib, case):
    _, dim, a, b, pi, pj, nij, dij = case
    out = oracle_lib.closest_points(dim, a[0], a[1], quaternion(a[2]), (a[3], 0, 0), b[0], b[1], quaternion(b[2]),
                                    (b[3], 0, 0))
    assert approx(out["dij"], dij)
    for got, ref in ((out["pi"], pi), (out["pj"], pj), (out["nij"], nij)):
        for g, r in zip(got, ref):
            assert approx(g, r)


def three_spheres():
    c = Container(2)
    c.sphere((0.0, 0.0), 0.1)
    c.sphere((1.0, 1.0), 0.2)
    c.sphere((5.0, 10.0), 0.1)
    return c.arrays()


@pytest.mark.parametrize("method", [0, 1], ids=["kdtree", "brute_force"])
def test_contacts_known_answer(oracle_lib, method):
    """test_contacts_kdtree.cpp:14-72 / test_contacts_brute_force.cpp:12-70."""
    o = oracle_lib.Oracle(three_spheres(), contact_method=method, kd_tree_radius=100.0)
    assert o.compute_contacts() == 1
    c = o.contacts()
    assert tuple(c["ij"][0]) == (0, 1)
    assert approx(c["dij"][0], math.sqrt(2.0) - 0.1 - 0.2)
    r2 = math.sqrt(2.0) / 2.0
    assert approx(c["nij"][0, 0], -r2) and approx(c["nij"][0, 1], -r2)
    assert approx(c["pi"][0, 0], 0.1 * r2) and approx(c["pi"][0, 1], 0.1 * r2)
    assert approx(c["pj"][0, 0], 1.0 - 0.2 * r2) and approx(c["pj"][0, 1], 1.0 - 0.2 * r2)


def unit(x):
    return x / np.linalg.norm(x)


def random_scene(dim, n, seed, box=1.0, rmin=0.02, rmax=0.06, planes=True):
    rng = np.random.default_rng(seed)
    c = Container(dim)
    if planes:
        if dim == 2:
            c.plane((0.0, 0.0), PI / 2)
            c.plane((0.0, 0.0), 0.0)
        else:
            c.plane((0.0, 0.0, 0.0), q=quaternion(-PI / 2, (0.0, 1.0, 0.0)))
            c.plane((0.0, 0.0, 0.0), q=quaternion(0.0))
    for _ in range(n):
        pos = rng.uniform(0.0, box, size=dim)
        c.sphere(pos, rng.uniform(rmin, rmax), Property(mass=rng.uniform(0.5, 2.0),
                                                        moment_inertia=rng.uniform(0.01, 0.1, size=3) if dim == 3
                                                        else rng.uniform(0.01, 0.1),
                                                        velocity=rng.uniform(-1, 1, size=dim)),
                 q=quaternion(rng.uniform(0, 2 * PI), unit(rng.normal(size=3)) if dim == 3 else (0.0, 0.0, 1.0)))
    return c.arrays()


@pytest.mark.parametrize("dim", [2, 3])
def test_kdtree_set_equals_filtered_brute_force(oracle_lib, dim):
    """The kd-tree set is the brute-force set restricted to |ci-cj|^2 < kd_tree_radius."""
    sc = random_scene(dim, 300, seed=dim)
    kd_r = 0.12 ** 2
    kd = oracle_lib.Oracle(sc, contact_method=0, kd_tree_radius=kd_r, dmax=0.05)
    bf = oracle_lib.Oracle(sc, contact_method=1, dmax=0.05)
    kd.compute_contacts()
    bf.compute_contacts()
    a, b = kd.contacts(), bf.contacts()
    keep = []
    for k, (i, j) in enumerate(b["ij"]):
        d2 = 0.0
        for d in range(dim):
            diff = sc.pos[i, d] - sc.pos[j, d]
            d2 += diff * diff
        if i < sc.n_inactive or d2 < kd_r:
            keep.append(k)
    assert np.array_equal(a["ij"], b["ij"][keep])
    assert len(a["ij"]) > 50
    # every emitted pair meets dij < dmax; sorted, unique, i < j
    assert np.all(a["dij"] < 0.05)
    assert np.all(a["ij"][:, 0] < a["ij"][:, 1])
    key = a["ij"][:, 0] * sc.n + a["ij"][:, 1]
    assert np.all(np.diff(key) > 0)


def test_matrix_a_adjoint(oracle_lib):
    """test_matrices.cpp:17-47: <A u, f> = <u, A^T f> (3D), plus a larger random scene."""
    c = Container(3)
    c.sphere((0.0, 0.0, 0.0), 0.1)
    c.sphere((1.0, 1.0, 0.0), 0.2)
    for sc in (c.arrays(), random_scene(3, 200, seed=7)):
        o = oracle_lib.Oracle(sc, contact_method=1, dmax=0.1 if sc.n > 10 else 2.0)
        nc = o.compute_contacts()
        assert nc >= 1
        rng = np.random.default_rng(0)
        u = rng.random(6 * sc.n_active)
        f = rng.random(3 * nc)
        lhs = float(np.dot(o.amat(u), f))
        rhs = float(np.dot(u, o.atmat(f)))
        assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


def test_matrix_a_case_2(oracle_lib):
    """test_matrices.cpp:49-77: known values of A u for a 2D sphere on a 45-degree plane."""
    c = Container(2)
    c.plane((0.0, 0.0), PI / 2 - PI / 4)
    c.sphere((2.0, 1.0), 0.5, Property(mass=1, moment_inertia=0.1))
    o = oracle_lib.Oracle(c.arrays(), contact_method=1)
    assert o.compute_contacts() == 1
    u = np.array([0.187562190766376, -1.184390935327111, 0.0, 0.0, 0.0, -0.00453709103433])
    sol = o.amat(u)
    assert approx(sol[0], -0.18595809)
    assert approx(sol[1], 1.18278683)


def _rot(q):
    """Textbook rotation matrix of a unit quaternion (w, x, y, z)."""
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _qmul(a, b):
    w1, v1 = a[0], a[1:]
    w2, v2 = b[0], b[1:]
    return np.concatenate([[w1 * w2 - v1 @ v2], w1 * v2 + w2 * v1 + np.cross(v1, v2)])


def test_a_is_material_point_velocity(oracle_lib):
    """(A u)_c = d/dt (p_i - p_j) for material points p_k = x_k + R(q_k) r_k moving rigidly
    with x' = v, q' from q <- q * exp(h w / 2) (body-frame w, solver.hpp:463-479).

    A central finite difference with the textbook rotation matrix pins both the
    translational and the rotational blocks of A in 3D.
    """
    sc = random_scene(3, 120, seed=11, rmin=0.04, rmax=0.08)
    o = oracle_lib.Oracle(sc, contact_method=1, dmax=0.2)
    nc = o.compute_contacts()
    ct = o.contacts()
    rng = np.random.default_rng(3)
    na, off = sc.n_active, sc.n_inactive
    v = rng.normal(size=(na, 3))
    w = rng.normal(size=(na, 3))
    au = o.amat(np.concatenate([v.ravel(), w.ravel()])).reshape(nc, 3)
    h = 1e-6

    def point(k, p, t):
        if k < off:
            return p  # obstacles do not move
        x, q = sc.pos[k], sc.q[k]
        body = _rot(q).T @ (p - x)
        wk = w[k - off]
        nw = np.linalg.norm(wk)
        e = np.concatenate([[math.cos(0.5 * nw * t)], math.sin(0.5 * nw * t) / nw * wk])
        return x + t * v[k - off] + _rot(_qmul(q, e)) @ body

    for c in range(nc):
        i, j = ct["ij"][c]
        fd = ((point(i, ct["pi"][c], h) - point(j, ct["pj"][c], h))
              - (point(i, ct["pi"][c], -h) - point(j, ct["pj"][c], -h))) / (2 * h)
        np.testing.assert_allclose(au[c], fd, rtol=1e-6, atol=1e-6)


def test_threads_do_not_change_results(oracle_lib):
    """The restatement's OpenMP loops (the reference's pragma sites: search, vap, update and move)
    give the same contact list and the same step with 1 and 4 threads."""
    from scopi_amd import scenes
    sc = scenes.spheres_3d(16, 16, 8)
    kw = dict(vap=1, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-6)
    out = []
    for t in (1, 4):
        oracle_lib.set_threads(t)
        try:
            o = oracle_lib.Oracle(sc, **kw)
            o.compute_contacts()
            c = o.contacts()
            o.run(1e-3, 2)
            out.append((c, o.state()))
        finally:
            oracle_lib.set_threads(1)
    (c1, s1), (c4, s4) = out
    for k in c1:
        assert np.array_equal(c1[k], c4[k])
    for k in s1:
        assert np.array_equal(s1[k], s4[k])
