"""vap_projection in the oracle (vap/vap_projection.hpp:34-49): the slot writes the desired
velocity (u) and the rotation velocity (w(i, 2), the scalar assigned to all three components of
a 3D rotation velocity), and leaves the velocity itself as it was; the solve starts from that
velocity.  No reference test uses the slot; the pins are its own assignments."""
import numpy as np

from scopi_amd.scenes import Container, Property


def scene(dim):
    c = Container(dim)
    rng = np.random.default_rng(4)
    for k in range(6):
        c.sphere(rng.uniform(0, 1, dim), 0.05, Property(mass=1.0, velocity=rng.uniform(-1, 1, dim),
                                                         moment_inertia=0.1 if dim == 2 else (0.1, 0.1, 0.1)))
    return c.arrays()


def test_projection_assignments(oracle_lib):
    for dim in (2, 3):
        sc = scene(dim)
        o = oracle_lib.Oracle(sc, vap=oracle_lib.VAP_PROJECTION, dmax=0.01, kd_tree_radius=0.2 ** 2)
        u = np.arange(18, dtype=float).reshape(6, 3) * 0.1
        w = np.arange(18, dtype=float).reshape(6, 3) * -0.01
        o.set_projection_velocity(u, w)
        o.set_a_priori_velocity(0.01)
        st = o.state()
        np.testing.assert_array_equal(st["v"], sc.v)  # the velocity is not touched
        om = st["omega"]
        if dim == 2:
            np.testing.assert_array_equal(om[:, 2], w[:, 2])
        else:
            np.testing.assert_array_equal(om, np.repeat(w[:, 2:3], 3, axis=1))
