"""The demos' random draws (oracle.gen_segment_demo, oracle.gen_box_spheres_demo) against the published
definitions of the generator they use.

demos/segment.cpp and demos/box_spheres_3d.cpp draw from a default-constructed std::default_random_engine:
in libstdc++ that is minstd_rand0, x_{k+1} = 16807 x_k mod (2^31 - 1) from x_0 = 1.  A
uniform_real_distribution<double>(a, b) takes generate_canonical<double, 53>: with range R = 2^31 - 2 it
needs ceil(53 / log2(R + 1)) = 2 draws, u = ((g1 - 1) + (g2 - 1) R) / R^2 evaluated in double (the
product, the sum and R^2 each rounded), clamped below 1, and returns u (b - a) + a.  Restated here from
those definitions, this pins the first draws of both generators and the order in which the demos
consume them.
"""
import numpy as np

import oracle

M = 2 ** 31 - 1
R = M - 1


def draws(n):
    x, out = 1, []
    for _ in range(n):
        x = 16807 * x % M
        out.append(x)
    return out


def uniforms(ranges, count):
    g = draws(2 * len(ranges) * count)
    vals, k = [], 0
    for _ in range(count):
        row = []
        for a, b in ranges:
            total = float(g[k] - 1) * 1.0
            scale = float(R)
            total += float(g[k + 1] - 1) * scale
            scale *= float(R)
            u = min(total / scale, float(np.nextafter(1.0, 0.0)))
            row.append(u * (b - a) + a)
            k += 2
        vals.append(row)
    return np.array(vals)


def test_segment_demo_draws():
    x, y, r = oracle.gen_segment_demo(5, 0.06)
    ref = uniforms([(0.1, 0.9), (0.0, 0.7), (0.01, 0.06)], 5)
    np.testing.assert_array_equal(np.stack([x, y, r], axis=1), ref)


def test_box_spheres_demo_draws():
    r0 = 0.5
    r, m, dx, dy, dz = oracle.gen_box_spheres_demo(2, r0)
    ref = uniforms([(0.8 * r0, 0.9 * r0), (1.0, 2.0), (-0.05, 0.05), (-0.05, 0.05), (-0.05, 0.05)], 8)
    np.testing.assert_array_equal(np.stack([r, m, dx, dy, dz], axis=1), ref)


def test_box_spheres_demo_scene_layout():
    """The builder's geometry: five planes first (normals x, x, y, -z, -z), then spheres in the demo's
    i, j, k order at (r0 + 0.1 + 2 i r0 + dx, 1.1 r0 + 2.1 r0 j + dy, r0 + 0.1 + 2 k r0 + dz)."""
    from scopi_amd import scenes
    r0 = 0.5
    r, m, dx, dy, dz = oracle.gen_box_spheres_demo(2, r0)
    sc = scenes.box_spheres_demo(2, r0, 1.0, r, m, dx, dy, dz)
    assert sc.n == 13 and sc.n_inactive == 5
    # sphere (i, j, k) = (1, 0, 1) is the 6th drawn (index 5)
    k = 1 * 4 + 0 * 2 + 1
    np.testing.assert_allclose(sc.pos[5 + k], [r0 + 0.1 + 2 * r0 + dx[k], 1.1 * r0 + dy[k], r0 + 0.1 + 2 * r0 + dz[k]],
                               rtol=0, atol=1e-15)
    assert sc.size[5 + k, 0] == r[k] and sc.m[5 + k] == m[k]
    np.testing.assert_allclose(sc.J[5 + k], [m[k] * r[k] ** 2 / 2] * 3, rtol=1e-15)
    np.testing.assert_allclose(sc.f[5 + k], [0.0, -1.0, 0.0])
    # plane normals: the first column of each plane's rotation
    def normal(q):
        w, x, y, z = q
        return np.array([1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)])
    for p, n in zip(range(5), ([1, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, -1], [0, 0, -1])):
        np.testing.assert_allclose(normal(sc.q[p]), n, atol=1e-15)
    np.testing.assert_allclose(sc.pos[1], [2 * 2 * r0 + 2 * r0, 0, 0])
    np.testing.assert_allclose(sc.pos[4], [0, 0, 2 * 2 * r0 + 2 * r0])


def test_uniform_draws():
    """oracle.gen_uniform (the pile-of-sand demo's masses) is the same sequence."""
    np.testing.assert_array_equal(oracle.gen_uniform(9, 1.0, 2.0), uniforms([(1.0, 2.0)], 9)[:, 0])


def test_demo_scene_builders():
    """sphere_plane: the sphere touches the 30-degree incline (its centre 2 r from the plane's origin
    along the plane's normal); pile_of_sand: 27 spheres of radius 1.25, the layer j = 0 shifted by r / 2;
    two_ellipsoids: semi-axes (0.1, 0.05), turned by +-pi/4."""
    import math
    from scopi_amd import scenes
    sp = scenes.sphere_plane_demo()
    assert sp.n == 2 and sp.n_inactive == 1
    w, z = sp.q[0, 0], sp.q[0, 3]
    normal = np.array([1 - 2 * z * z, 2 * w * z])  # first column of the plane's 2D rotation
    assert abs(float(normal @ sp.pos[1, :2]) - 2.0) < 1e-12 and abs(float(normal[1]) - math.cos(math.pi / 6)) < 1e-12
    pile = scenes.pile_of_sand_demo(oracle.gen_uniform(27, 1.0, 2.0))
    assert pile.n == 28 and np.all(pile.size[1:, 0] == 1.25)
    assert np.allclose(pile.pos[19:, 1], 1.25) and np.allclose(pile.pos[19, 0], 0.625)
    te = scenes.two_ellipsoids_demo()
    assert np.allclose(te.size[:, :2], [[0.1, 0.05], [0.1, 0.05]])
    assert np.allclose(te.q[:, 3], [math.sin(math.pi / 8), -math.sin(math.pi / 8)])
