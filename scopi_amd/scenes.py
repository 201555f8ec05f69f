"""Particle containers and the synthetic workloads of the contact-step benchmarks.

`Container` mirrors `scopi_container<dim>::push_back(object, property)`
(include/scopi/container.hpp:338-384): obstacles (inactive objects) must be
pushed before active particles, and every per-particle field is stored as a
structure of arrays.  Conventions shared with the C-ABI (include/scopi_hip.h):

* vectors use 3 slots even in 2D; the 2D rotation velocity and moment of
  inertia sit in slot 2, the slot the reference's 6-dof velocity vector uses
  (include/scopi/solvers/minimization_problem.hpp:15-29);
* quaternions are scalar-first (include/scopi/quaternion.hpp:86-124);
* `size[:, 0]` is a sphere radius or a segment length; superellipsoids use
  `size[:, 0:3]` (semi-axes) and `squareness[:, 0:2]` = (e, n).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

SPHERE, PLANE, SEGMENT, SUPERELLIPSOID = 0, 1, 2, 3


def quaternion(angle: float, axis=(0.0, 0.0, 1.0)) -> np.ndarray:
    """quaternion(angle, axis) of src/quaternion.cpp:12-26 (normalised, scalar first)."""
    h = angle / 2.0
    q = np.array([math.cos(h), math.sin(h) * axis[0], math.sin(h) * axis[1], math.sin(h) * axis[2]])
    return q / math.sqrt(float(np.dot(q, q)))


@dataclass
class Property:
    """property<dim> (include/scopi/property.hpp:160-199); all fields default to zero, mass to 1."""

    velocity: tuple = (0.0, 0.0, 0.0)
    desired_velocity: tuple = (0.0, 0.0, 0.0)
    omega: tuple = (0.0, 0.0, 0.0)
    desired_omega: tuple = (0.0, 0.0, 0.0)
    force: tuple = (0.0, 0.0, 0.0)
    mass: float = 1.0
    moment_inertia: tuple = (0.0, 0.0, 0.0)
    active: bool = True


def _vec3(x, dim, rot=False):
    """Place a dim-vector (or a 2D scalar rotation quantity) into 3 slots."""
    if np.isscalar(x):
        out = np.zeros(3)
        out[2 if rot else 0] = float(x)
        return out
    x = np.asarray(x, dtype=np.float64).ravel()
    out = np.zeros(3)
    out[: x.size] = x
    return out


@dataclass
class Container:
    dim: int
    shape: list = field(default_factory=list)
    pos: list = field(default_factory=list)
    q: list = field(default_factory=list)
    size: list = field(default_factory=list)
    squareness: list = field(default_factory=list)
    v: list = field(default_factory=list)
    vd: list = field(default_factory=list)
    omega: list = field(default_factory=list)
    domega: list = field(default_factory=list)
    f: list = field(default_factory=list)
    m: list = field(default_factory=list)
    J: list = field(default_factory=list)
    n_inactive: int = 0

    def _push(self, shape, pos, q, size, sq, prop: Property):
        if not prop.active and self.n_inactive != len(self.shape):
            raise RuntimeError("All the obstacles must be pushed before the active particles.")
        self.shape.append(shape)
        self.pos.append(_vec3(pos, self.dim))
        self.q.append(np.asarray(q, dtype=np.float64))
        self.size.append(_vec3(size, 3))
        self.squareness.append(np.asarray(sq, dtype=np.float64))
        self.v.append(_vec3(prop.velocity, self.dim))
        self.vd.append(_vec3(prop.desired_velocity, self.dim))
        self.omega.append(_vec3(prop.omega, self.dim, rot=True))
        self.domega.append(_vec3(prop.desired_omega, self.dim, rot=True))
        self.f.append(_vec3(prop.force, self.dim))
        self.m.append(float(prop.mass))
        self.J.append(_vec3(prop.moment_inertia, self.dim, rot=True))
        if not prop.active:
            self.n_inactive += 1

    def sphere(self, pos, radius, prop=None, q=None):
        """sphere<dim>(pos, radius) (include/scopi/objects/types/sphere.hpp:139-151)."""
        self._push(SPHERE, pos, quaternion(0.0) if q is None else q, (radius, 0, 0), (1.0, 1.0), prop or Property())

    def plane(self, pos, angle=0.0, prop=None, q=None):
        """plane<dim>(pos, angle) / plane<dim>(pos, q) (include/scopi/objects/types/plane.hpp:122-140)."""
        self._push(PLANE, pos, quaternion(angle) if q is None else q, (0, 0, 0), (1.0, 1.0),
                   prop or Property(active=False))

    def segment(self, pos, length, angle=0.0, prop=None, q=None):
        """segment<2>(pos, q, length) (include/scopi/objects/types/segment.hpp:147-152)."""
        self._push(SEGMENT, pos, quaternion(angle) if q is None else q, (length, 0, 0), (1.0, 1.0),
                   prop or Property(active=False))

    def segment_points(self, pos1, pos2, prop=None):
        """segment<2>(pos1, pos2) (include/scopi/objects/types/segment.hpp:133-138): centre at the
        midpoint, angle atan2(x1 - x2, y2 - y1), length |pos2 - pos1|."""
        p1, p2 = np.asarray(pos1, np.float64), np.asarray(pos2, np.float64)
        d = p2 - p1
        self.segment(0.5 * (p1 + p2), math.sqrt(d[0] * d[0] + d[1] * d[1]), math.atan2(p1[0] - p2[0], p2[1] - p1[1]),
                     prop=prop)

    def superellipsoid(self, pos, q, radius, squareness, prop=None):
        """superellipsoid<3>(pos, q, radius(3), squareness(e, n)) (superellipsoid.hpp)."""
        self._push(SUPERELLIPSOID, pos, q, radius, squareness, prop or Property())

    def arrays(self) -> "Scene":
        n = len(self.shape)
        def a(x, w):
            return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(n, w))
        return Scene(
            dim=self.dim, n_inactive=self.n_inactive,
            shape=np.asarray(self.shape, dtype=np.int32).reshape(n),
            pos=a(self.pos, 3), q=a(self.q, 4), size=a(self.size, 3), squareness=a(self.squareness, 2),
            v=a(self.v, 3), vd=a(self.vd, 3), omega=a(self.omega, 3), domega=a(self.domega, 3),
            f=a(self.f, 3), m=np.asarray(self.m, dtype=np.float64).reshape(n), J=a(self.J, 3))


@dataclass
class Scene:
    """Structure-of-arrays particle state (obstacles first)."""

    dim: int
    n_inactive: int
    shape: np.ndarray
    pos: np.ndarray
    q: np.ndarray
    size: np.ndarray
    squareness: np.ndarray
    v: np.ndarray
    vd: np.ndarray
    omega: np.ndarray
    domega: np.ndarray
    f: np.ndarray
    m: np.ndarray
    J: np.ndarray

    @property
    def n(self) -> int:
        return int(self.shape.shape[0])

    @property
    def n_active(self) -> int:
        return self.n - self.n_inactive

    def copy(self) -> "Scene":
        return Scene(self.dim, self.n_inactive, *(getattr(self, k).copy() for k in (
            "shape", "pos", "q", "size", "squareness", "v", "vd", "omega", "domega", "f", "m", "J")))


# ----------------------------------------------------------------------------- workloads

def box_planes_3d(lo, hi):
    """Floor z=lo[2] and four walls x=lo/hi, y=lo/hi as planes (normal = first column of R(q)).

    Plane orientations follow demos/box_spheres_3d.cpp:34-63 (quaternion(angle, axis)).
    """
    ez, ey = (0.0, 0.0, 1.0), (0.0, 1.0, 0.0)
    return [
        ((0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), lo[2]), quaternion(-math.pi / 2, ey)),  # floor, n=(0,0,1)
        ((lo[0], 0.0, 0.0), quaternion(0.0, ez)),                                               # x = lo
        ((hi[0], 0.0, 0.0), quaternion(0.0, ez)),                                               # x = hi
        ((0.0, lo[1], 0.0), quaternion(math.pi / 2, ez)),                                       # y = lo
        ((0.0, hi[1], 0.0), quaternion(math.pi / 2, ez)),                                       # y = hi
    ]


def disks_2d(n_side=32, spacing=0.025, jitter=0.002, rmin=0.008, rmax=0.012, seed=1234):
    """1,024 disks on a jittered 32x32 lattice above the plane y=0 (gravity (0,-9.81) via vap_fpd).

    The first lattice row's centre height is `spacing` (0.025) above the plane.
    """
    rng = np.random.default_rng(seed)
    c = Container(2)
    c.plane((0.0, 0.0), math.pi / 2)  # normal (0, 1)
    for i in range(n_side):
        for j in range(n_side):
            r = rng.uniform(rmin, rmax)
            x = (i + 0.5) * spacing + rng.uniform(-jitter, jitter)
            y = (j + 1.0) * spacing + rng.uniform(-jitter, jitter)
            c.sphere((x, y), r, Property(mass=1.0, moment_inertia=0.1, force=(0.0, -9.81)))
    return c.arrays()


def spheres_3d(nx=64, ny=64, nz=32, spacing=0.025, jitter=0.003, rmin=0.009, rmax=0.011, seed=4321,
               velocity_scale=0.0, count=None):
    """Jittered nx*ny*nz lattice of spheres in a box of 5 planes (floor + 4 walls).

    Lattice site (i, j, k) sits at ((i+0.5)s, (j+0.5)s, (k+0.5)s) + jitter with the floor at z=0
    and the walls half a spacing outside the lattice.  Gravity (0,0,-9.81)*m via vap_fpd, m=1,
    J=(0.1,0.1,0.1).  `velocity_scale` > 0 adds a seeded uniform initial velocity in
    [-velocity_scale, velocity_scale]^3 per sphere (0 for the configuration as specified).
    `count` keeps the first `count` lattice sites, layer by layer from the floor up.
    """
    rng = np.random.default_rng(seed)
    n = nx * ny * nz if count is None else int(count)
    lo = (0.0, 0.0, 0.0)
    hi = (nx * spacing, ny * spacing, nz * spacing + 1.0)
    planes = box_planes_3d(lo, hi)
    ni = len(planes)
    N = ni + n
    shape = np.empty(N, np.int32)
    pos = np.zeros((N, 3))
    q = np.zeros((N, 4))
    size = np.zeros((N, 3))
    shape[:ni] = PLANE
    for k, (p, qq) in enumerate(planes):
        pos[k] = p
        q[k] = qq
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    base = np.stack([(ii.ravel() + 0.5), (jj.ravel() + 0.5), (kk.ravel() + 0.5)], axis=1) * spacing
    if count is not None:
        if n > len(base):
            raise ValueError("spheres_3d: count exceeds the lattice")
        base = base[np.lexsort((jj.ravel(), ii.ravel(), kk.ravel()))][:n]
    pos[ni:] = base + rng.uniform(-jitter, jitter, size=(n, 3))
    size[ni:, 0] = rng.uniform(rmin, rmax, size=n)
    shape[ni:] = SPHERE
    q[ni:, 0] = 1.0
    v = np.zeros((N, 3))
    if velocity_scale > 0:
        v[ni:] = rng.uniform(-velocity_scale, velocity_scale, size=(n, 3))
    f = np.zeros((N, 3))
    f[ni:, 2] = -9.81
    m = np.ones(N)
    J = np.zeros((N, 3))
    J[ni:] = 0.1
    return Scene(dim=3, n_inactive=ni, shape=shape, pos=pos, q=q, size=size, squareness=np.ones((N, 2)),
                 v=v, vd=np.zeros((N, 3)), omega=np.zeros((N, 3)), domega=np.zeros((N, 3)), f=f, m=m, J=J)


def spheres_3d_count(n_particles, seed=0, exact=False, **kw):
    """A scene of the 131k-sphere recipe scaled to `n_particles` spheres (lattice nx = ny = 2 nz).

    exact=True keeps exactly `n_particles` sites (the top layer partly filled).
    """
    nz = max(1, int(round((n_particles / 4.0) ** (1.0 / 3.0))))
    nx = max(1, int(round(math.sqrt(n_particles / nz))))
    ny = max(1, int(math.ceil(n_particles / (nx * nz))))
    if exact:
        return spheres_3d(nx=nx, ny=ny, nz=nz, seed=seed, count=n_particles, **kw)
    return spheres_3d(nx=nx, ny=ny, nz=nz, seed=seed, **kw)


def batched_sphere_scenes(n_scenes=512, nmin=256, nmax=32768, seed=77):
    """Independent sphere scenes with log-uniform sizes in [nmin, nmax] (seeded), each built by
    the 131k-sphere recipe scaled to its count (box of 5 planes, gravity, jittered lattice)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(nmin), math.log(nmax), size=n_scenes)).round().astype(np.int64)
    return [spheres_3d_count(int(n), seed=seed * 100003 + k, exact=True) for k, n in enumerate(sizes)]


def superellipsoids_3d(nx=32, ny=32, nz=32, spacing=0.035, jitter=0.003, amin=0.008, amax=0.015, emin=0.4,
                       emax=1.6, velocity_scale=0.5, seed=2024):
    """Jittered nx*ny*nz lattice of superellipsoids, no obstacles, no gravity.

    Semi-axes U[amin, amax]^3, squareness (e, n) U[emin, emax]^2, orientations uniform
    random unit quaternions (normalised 4-Gaussians), m = 1, J = (0.1, 0.1, 0.1).  The
    lattice spacing leaves most neighbours separated, so the scene is driven by seeded
    uniform initial velocities in [-velocity_scale, velocity_scale]^3 (vap_fpd with zero
    force keeps them): approaching pairs come into contact within a few steps.
    """
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    base = np.stack([(ii.ravel() + 0.5), (jj.ravel() + 0.5), (kk.ravel() + 0.5)], axis=1) * spacing
    pos = base + rng.uniform(-jitter, jitter, size=(n, 3))
    size = rng.uniform(amin, amax, size=(n, 3))
    sq = rng.uniform(emin, emax, size=(n, 2))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    v = rng.uniform(-velocity_scale, velocity_scale, size=(n, 3)) if velocity_scale > 0 else np.zeros((n, 3))
    J = np.full((n, 3), 0.1)
    return Scene(dim=3, n_inactive=0, shape=np.full(n, SUPERELLIPSOID, np.int32), pos=pos, q=q, size=size,
                 squareness=sq, v=v, vd=np.zeros((n, 3)), omega=np.zeros((n, 3)), domega=np.zeros((n, 3)),
                 f=np.zeros((n, 3)), m=np.ones(n), J=J)


def segment_demo(x, y, r):
    """The scene of demos/segment.cpp:15-60: two segment obstacles, (0, 1)-(0.4, 1) and (0.6, 1)-(1, 1),
    and one sphere per draw (x, y, r) with desired velocity (0.5 - x, 2 - y), mass 1, inertia 0.1
    (the demo takes its draws from std::default_random_engine)."""
    c = Container(2)
    c.segment_points((0.0, 1.0), (0.4, 1.0))
    c.segment_points((0.6, 1.0), (1.0, 1.0))
    for xi, yi, ri in zip(x, y, r):
        c.sphere((xi, yi), ri, Property(desired_velocity=(0.5 - xi, 2 - yi), mass=1.0, moment_inertia=0.1))
    return c.arrays()


def box_spheres_demo(n_side, r0, g, r, m, dx, dy, dz):
    """The scene of demos/box_spheres_3d.cpp:12-95: five obstacle planes (x = 0 and x = W + 2 r0 with
    normal x, y = 0 with normal y, z = 0 and z = W + 2 r0 with the y-rotated normal), then n_side^3
    spheres at (r0 + 0.1 + 2 i r0 + dx, 1.1 r0 + 2.1 r0 j + dy, r0 + 0.1 + 2 k r0 + dz), loops i, j, k
    outermost first, with force (0, -g, 0), mass m and moment of inertia m r^2 / 2 about each axis;
    W = 2 n_side r0.  (r, m, dx, dy, dz) are the demo's draws in its loop order."""
    w = 2.0 * n_side * r0
    c = Container(3)
    c.plane((0.0, 0.0, 0.0), q=quaternion(0.0, (0.0, 0.0, 1.0)))
    c.plane((w + 2 * r0, 0.0, 0.0), q=quaternion(0.0, (0.0, 0.0, 1.0)))
    c.plane((0.0, 0.0, 0.0), q=quaternion(math.pi / 2.0, (0.0, 0.0, 1.0)))
    c.plane((0.0, 0.0, 0.0), q=quaternion(math.pi / 2.0, (0.0, 1.0, 0.0)))
    c.plane((0.0, 0.0, w + 2 * r0), q=quaternion(math.pi / 2.0, (0.0, 1.0, 0.0)))
    k = 0
    for i in range(n_side):
        for j in range(n_side):
            for l in range(n_side):
                ri, mi = r[k], m[k]
                c.sphere((r0 + 0.1 + i * 2.0 * r0 + dx[k], 1.1 * r0 + 2.1 * r0 * j + dy[k], r0 + 0.1 + l * 2.0 * r0 + dz[k]),
                         ri, Property(force=(0.0, -g, 0.0), mass=mi,
                                      moment_inertia=(mi * ri * ri / 2.0, mi * ri * ri / 2.0, mi * ri * ri / 2.0)))
                k += 1
    return c.arrays()


def sphere_plane_demo():
    """The scene of demos/sphere_plane.cpp:14-40 (2D): a plane through the origin at angle pi/2 - pi/6
    (an incline of 30 degrees), a sphere of radius 1 at 2 (sin(pi/6), cos(pi/6)) with mass 1, moment of
    inertia 1/2 and force (0, -1)."""
    alpha = math.pi / 6.0
    c = Container(2)
    c.plane((0.0, 0.0), math.pi / 2.0 - alpha)
    c.sphere((2.0 * math.sin(alpha), 2.0 * math.cos(alpha)), 1.0,
             Property(mass=1.0, moment_inertia=0.5, force=(0.0, -1.0)))
    return c.arrays()


def two_ellipsoids_demo():
    """The scene of demos/two_ellipsoids.cpp:13-46 (2D): two ellipses of semi-axes (0.1, 0.05) at
    (-0.2, 0) and (0.2, 0), turned by pi/4 and -pi/4, desired velocities (0.25, 0) and (-0.25, 0),
    mass 1, moment of inertia 0.1."""
    c = Container(2)
    for x, ang, vx in ((-0.2, math.pi / 4.0, 0.25), (0.2, -math.pi / 4.0, -0.25)):
        c.superellipsoid((x, 0.0), quaternion(ang), (0.1, 0.05, 0.0), (1.0, 1.0),
                         Property(desired_velocity=(vx, 0.0), mass=1.0, moment_inertia=0.1))
    return c.arrays()


def pile_of_sand_demo(masses, n=3, width_box=10.0, g=1.0):
    """The scene of demos/pile_of_sand_spheres.cpp:14-71 (3D): the plane y = 0 (angle pi/2), then spheres
    of radius r = width_box / 2 / (n + 1) at (2 i r, r + 2 j r, 2 k r) for j >= 1 (loops i, j, k), then the
    layer j = 0 shifted by r / 2 in x, force (0, -g, 0), mass from `masses` in that order and moment of
    inertia m r^2 / 2 about each axis."""
    r = width_box / 2.0 / (n + 1)
    c = Container(3)
    c.plane((0.0, 0.0, 0.0), math.pi / 2.0)
    k = 0

    def add(pos):
        nonlocal k
        m = masses[k]
        k += 1
        c.sphere(pos, r, Property(force=(0.0, -g, 0.0), mass=m,
                                  moment_inertia=(m * r * r / 2.0, m * r * r / 2.0, m * r * r / 2.0)))
    for i in range(n):
        for j in range(1, n):
            for l in range(n):
                add((i * 2.0 * r, r + j * 2.0 * r, l * 2.0 * r))
    dec_x = 0.5 * r
    for i in range(n):
        for l in range(n):
            add((i * 2.0 * r + dec_x, r, l * 2.0 * r))
    return c.arrays()


# Default solver parameters of the workloads (BASELINE configurations).
CONFIG1 = dict(dim=2, dt=0.005, dmax=0.025, kd_tree_radius=(0.024 + 0.025) ** 2, tolerance=1e-6, steps=200)
CONFIG2 = dict(dim=3, dt=1e-3, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-6)
# kd_tree_radius = (2 |a|_max + dmax)^2 with |a|_max = sqrt(3) 0.015: never binding for these shapes
CONFIG3 = dict(dim=3, dt=1e-3, dmax=0.02, kd_tree_radius=(2 * math.sqrt(3.0) * 0.015 + 0.02) ** 2, mu=0.3,
               tolerance=1e-6, steps=100)
# 512 independent sphere scenes, log-uniform [256, 32768] particles, FP32 + FP64 refinement
CONFIG4 = dict(dim=3, dt=1e-3, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-6, steps=50, n_scenes=512,
               nmin=256, nmax=32768)
