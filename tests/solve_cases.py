"""Seeded first-step contact problems shared by the solver tests (CPU and GPU)."""
import math

import numpy as np

from scopi_amd import scenes
from scopi_amd.scenes import Container, Property, quaternion


def random_scene(dim, n, seed, box=1.0, rmin=0.02, rmax=0.06):
    """n random spheres (masses, inertias, velocities, spins, orientations from `seed`) above two planes."""
    rng = np.random.default_rng(seed)
    c = Container(dim)
    if dim == 2:
        c.plane((0.0, 0.0), math.pi / 2)
        c.plane((0.0, 0.0), 0.0)
    else:
        c.plane((0.0, 0.0, 0.0), q=quaternion(-math.pi / 2, (0.0, 1.0, 0.0)))
        c.plane((0.0, 0.0, 0.0), q=quaternion(0.0))
    box = box * (n / 300.0) ** (1.0 / dim)
    for _ in range(n):
        ax = rng.normal(size=3)
        c.sphere(rng.uniform(0.0, box, size=dim), rng.uniform(rmin, rmax),
                 Property(mass=rng.uniform(0.5, 2.0), velocity=rng.uniform(-1, 1, size=dim),
                          moment_inertia=rng.uniform(0.01, 0.1, size=3) if dim == 3 else rng.uniform(0.01, 0.1),
                          omega=rng.uniform(-1, 1, size=3) if dim == 3 else rng.uniform(-1, 1),
                          force=(0.0, -9.81) if dim == 2 else (0.0, 0.0, -9.81)),
                 q=quaternion(rng.uniform(0, 2 * math.pi), ax / np.linalg.norm(ax) if dim == 3 else (0, 0, 1)))
    return c.arrays()


CASES = {
    "random_2d_nofriction": dict(scene=lambda: random_scene(2, 120, 11), problem=0, dt=0.01, dmax=0.03,
                                 kd=0.15 ** 2, mu=0.3),
    "random_3d_nofriction": dict(scene=lambda: random_scene(3, 400, 12), problem=0, dt=0.01, dmax=0.03,
                                 kd=0.15 ** 2, mu=0.3),
    "random_2d_friction": dict(scene=lambda: random_scene(2, 120, 13), problem=1, dt=0.01, dmax=0.03,
                               kd=0.15 ** 2, mu=0.3),
    "random_3d_friction": dict(scene=lambda: random_scene(3, 400, 14), problem=1, dt=0.01, dmax=0.03,
                               kd=0.15 ** 2, mu=0.3),
    "lattice_3d_nofriction": dict(scene=lambda: scenes.spheres_3d(8, 8, 4, velocity_scale=2.0), problem=0, dt=1e-3,
                                  dmax=0.02, kd=0.042 ** 2, mu=0.3),
}


def oracle_kw(case, **over):
    kw = dict(problem=case["problem"], vap=1, dmax=case["dmax"], kd_tree_radius=case["kd"], tolerance=1e-9,
              max_ite=200000, mu=case["mu"])
    kw.update(over)
    return kw
