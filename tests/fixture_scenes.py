"""The reference's fixture scenes, rebuilt with the project's container.

Each builder restates the set-up of one reference test whose output JSON is
committed under tests/golden/references/ (BSD-3, see LICENSE.scopi there).
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from scopi_amd.scenes import Container, Property

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "references")
PI = math.pi

# vap slot per reference test: vap_fixed unless the test names vap_fpd
FIXED, FPD = 0, 1


def load_fixture(name):
    with open(os.path.join(GOLDEN, name + ".json")) as fh:
        return json.load(fh)


def two_spheres(asymmetrical: bool):
    """test/test_sphere.cpp:341-421 (dt .005, 1000 steps, vap_fixed, default apgd / kd params)."""
    c = Container(2)
    dy = 0.05 if asymmetrical else 0.0
    p1 = Property(mass=1.0, moment_inertia=0.1, desired_velocity=(0.25, 0.0))
    p2 = Property(mass=1.0, moment_inertia=0.1, desired_velocity=(-0.25, 0.0))
    c.sphere((-0.2, -dy), 0.1, p1)
    c.sphere((0.2, dy), 0.1, p2)
    name = "two_spheres_asymmetrical" if asymmetrical else "two_spheres_symmetrical"
    return dict(scene=c.arrays(), dt=0.005, total_it=1000, vap=FIXED, params={}, fixture=name)


def case_2d_spheres(gen):
    """test/test_sphere.cpp:423-484: 18 spheres from std::minstd_rand0(123), dt .01, 20 steps."""
    x, y, r, vel = gen(3)
    c = Container(2)
    for k in range(x.size):
        c.sphere((x[k], y[k]), r[k], Property(mass=1.0, moment_inertia=0.1, desired_velocity=(vel[k], 0.0)))
    return dict(scene=c.arrays(), dt=0.01, total_it=20, vap=FIXED, params={}, fixture="2d_case_spheres")


def obstacles_sphere_sphere_moving():
    """test/test_obstacles.cpp:216-254: inactive sphere + sphere under force (0,-10), vap_fpd."""
    c = Container(2)
    c.sphere((0.0, 0.0), 1.0, Property(active=False))
    c.sphere((1.0, math.sqrt(3.0)), 1.0, Property(mass=1.0, moment_inertia=0.1, force=(0.0, -10.0)))
    return dict(scene=c.arrays(), dt=0.005, total_it=100, vap=FPD, params={},
                fixture="obstacles_sphere_sphere_moving")


def three_spheres_nofriction():
    """test/test_gradient.cpp:377-439: 3 spheres, vap_fpd, apgd tol 1e-7, alpha 0.1, 60 steps."""
    c = Container(2)
    c.sphere((-1.7, 1.3), 0.5, Property(mass=1, moment_inertia=1, force=(1.7, -1.3)))
    c.sphere((0.5, 1.7), 0.5, Property(mass=1, moment_inertia=1, force=(-0.5, -1.7)))
    c.sphere((4.5, 1.3), 0.5, Property(mass=1, moment_inertia=1, force=(-4.5, -1.3)))
    return dict(scene=c.arrays(), dt=0.1, total_it=60, vap=FPD,
                params=dict(tolerance=1e-7, max_ite=10000, alpha=0.1, dynamic_descent=1),
                fixture="3spheres_nofriction")


def three_spheres_friction_fp():
    """test/test_gradient.cpp:505-572: the 3 spheres of 3spheres_nofriction under FrictionFixedPoint
    (mu 0.5, fixed_point_tol 1e-6, fixed_point_max_iter 1000), apgd tol 1e-7, alpha 0.1, 70 steps."""
    case = three_spheres_nofriction()
    case.update(total_it=70, fixture="3spheres_friction_fp",
                params=dict(tolerance=1e-7, max_ite=10000, alpha=0.1, dynamic_descent=1, problem=2, mu=0.5,
                            fixed_point_tol=1e-6, fixed_point_max_iter=1000))
    return case


def three_spheres_viscous():
    """test/test_gradient.cpp:441-503: the 3 spheres of 3spheres_nofriction under Viscous (default
    gamma 0, gamma_min -2, gamma_tol 1e-6), apgd tol 1e-7, alpha 0.1, 70 steps."""
    case = three_spheres_nofriction()
    case.update(total_it=70, fixture="3spheres_viscous",
                params=dict(tolerance=1e-7, max_ite=10000, alpha=0.1, dynamic_descent=1, problem=3))
    return case


def three_spheres_viscous_friction():
    """test/test_gradient.cpp:574-640: ViscousFriction, mu 0.5, fixed_point_tol 1e-3, max_iter 1000,
    gamma 0, gamma_min -1.4, gamma_tol 1e-6, apgd tol 1e-7, alpha 0.1, 70 steps."""
    case = three_spheres_nofriction()
    case.update(total_it=70, fixture="3spheres_viscous_friction",
                params=dict(tolerance=1e-7, max_ite=10000, alpha=0.1, dynamic_descent=1, problem=4, mu=0.5,
                            fixed_point_tol=1e-3, fixed_point_max_iter=1000, gamma=0.0, gamma_min=-1.4,
                            gamma_tol=1e-6))
    return case


def sphere_plane_friction_mu01():
    """test/test_gradient.cpp:156-213: Friction, plane at pi/2-pi/6, mu 0.1, dt 0.1, 200 steps."""
    c = Container(2)
    c.plane((0.0, 0.0), PI / 2 - PI / 6)
    dd, rr = 1.0, 2.0
    c.sphere((0.0, (rr + dd) / math.cos(PI / 6)), rr, Property(mass=1, moment_inertia=0.5 * rr * rr, force=(0, -1)))
    return dict(scene=c.arrays(), dt=0.1, total_it=200, vap=FPD,
                params=dict(tolerance=1e-7, max_ite=10000, alpha=0.1, dynamic_descent=1, problem=1, mu=0.1),
                fixture="sphere_plane_friction_mu01")


def inclined_plane(alpha):
    """test/test_obstacles.cpp:256-313: sphere on an inclined plane, 1000 steps of dt 0.01, vap_fpd."""
    radius, g, h = 1.0, 1.0, 2.0
    c = Container(2)
    c.plane((0.0, 0.0), PI / 2.0 - alpha)
    c.sphere((h * math.sin(alpha), h * math.cos(alpha)), radius,
             Property(mass=1.0, moment_inertia=radius * radius / 2.0, force=(0.0, -g)))
    return dict(scene=c.arrays(), dt=0.01, total_it=1000, vap=FPD, params={})


# test_obstacles.cpp:258-260: (alpha, expected error_pos, expected error_q, expected error_v)
INCLINED_EXPECTED = [(PI / 6.0, 0.000998789, 0.0, 0.0010002), (PI / 4.0, 0.00100008, 0.0, 0.00100119),
                     (PI / 3.0, 0.00100105, 0.0, 0.0010027)]


def analytical_solution_sphere_plane(alpha, mu, t, r, g, y0):
    """test/analytical_solution.cpp / test/utils.cpp:77-110 (position, rotation angle)."""
    t_impact = math.sqrt(2 * (y0 - r) / (g * math.cos(alpha)))
    if t > t_impact:
        v_t_m = g * t_impact * math.sin(alpha)
        v_n_m = -g * t_impact * math.cos(alpha)
        t2 = t - t_impact
        x_impact = g * math.sin(alpha) * t_impact * t_impact / 2.0
        if math.tan(alpha) <= 3 * mu:
            x_normal = g * math.sin(alpha) * t2 * t2 / 3.0 + 2.0 * v_t_m * t2 / 3.0 + x_impact
            theta = -g * math.sin(alpha) * t2 * t2 / (3.0 * r) - 2 * v_t_m * t2 / (3.0 * r)
        else:
            x_normal = g * (math.sin(alpha) - mu * math.cos(alpha)) * t2 * t2 / 2.0 + (v_t_m + mu * v_n_m) * t2 + x_impact
            theta = -mu * g * math.cos(alpha) * t2 * t2 / r + 2 * mu * v_n_m * t2 / r
        x = np.array([x_normal * math.cos(alpha) + r * math.sin(alpha), -x_normal * math.sin(alpha) + r * math.cos(alpha)])
        return x, theta
    return np.array([y0 * math.sin(alpha), y0 * math.cos(alpha) - g * t * t / 2.0]), 0.0


def analytical_solution_sphere_plane_velocity(alpha, mu, t, r, g, y0):
    """test/utils.cpp:112-140 (velocity, rotation velocity)."""
    t_impact = math.sqrt(2 * (y0 - r) / (g * math.cos(alpha)))
    if t > t_impact:
        v_t_m = g * t_impact * math.sin(alpha)
        v_n_m = -g * t_impact * math.cos(alpha)
        t2 = t - t_impact
        if math.tan(alpha) <= 3 * mu:
            v_normal = 2.0 * g * math.sin(alpha) * t2 / 3.0 + 2.0 * v_t_m / 3.0
            omega = -2.0 * g * math.sin(alpha) * t2 / (3.0 * r) - 2 * v_t_m / (3.0 * r)
        else:
            v_normal = g * (math.sin(alpha) - mu * math.cos(alpha)) * t2 + (v_t_m + mu * v_n_m)
            omega = -2.0 * mu * g * math.cos(alpha) * t2 / r + 2 * mu * v_n_m / r
        return np.array([v_normal * math.cos(alpha), -v_normal * math.sin(alpha)]), omega
    return np.array([0.0, -g * t]), 0.0


def compare_objects(fixture, dim, state, n, tol):
    """diffFile (test/utils.cpp:14-75) restricted to the dynamic fields of "objects".

    Returns the largest absolute difference over position, quaternion, velocity
    and omega of every object.
    """
    objs = fixture["objects"]
    assert len(objs) == n
    worst = 0.0
    for k, o in enumerate(objs):
        pos = np.asarray(o["position"], dtype=float)
        worst = max(worst, float(np.max(np.abs(pos - state["pos"][k, :dim]))))
        qq = np.asarray(o["quaternion"], dtype=float)
        worst = max(worst, float(np.max(np.abs(qq - state["q"][k]))))
        vel = np.asarray(o["properties"]["velocity"], dtype=float)
        worst = max(worst, float(np.max(np.abs(vel - state["v"][k, :dim]))))
        om = o["properties"]["omega"]
        om = np.atleast_1d(np.asarray(om, dtype=float))
        mine = state["omega"][k, 2:3] if dim == 2 else state["omega"][k]
        worst = max(worst, float(np.max(np.abs(om - mine))))
    return worst


def sphere_plane_tilted_gravity(mu, problem, total_it=500, dt=0.01):
    """test/test_gradient.cpp:215-282: a horizontal plane (angle pi/2) and a sphere of radius 1 at
    height 2 under the force (g sin a, -g cos a), a = pi/4, g = 1; apgd tol 1e-7, alpha 0.01,
    fixed_point_tol 1e-6, fixed_point_max_iter 1000, vap_fpd."""
    radius, g, h, alpha = 1.0, 1.0, 2.0, PI / 4
    c = Container(2)
    c.plane((0.0, 0.0), PI / 2)
    c.sphere((0.0, h), radius, Property(mass=1.0, moment_inertia=radius * radius / 2.0,
                                        force=(g * math.sin(alpha), -g * math.cos(alpha))))
    return dict(scene=c.arrays(), dt=dt, total_it=total_it, vap=FPD, alpha=alpha, radius=radius, g=g, h=h, mu=mu,
                params=dict(tolerance=1e-7, max_ite=10000, alpha=0.01, dynamic_descent=1, problem=problem, mu=mu,
                            fixed_point_tol=1e-6, fixed_point_max_iter=1000))


def analytical_sphere_plane_friction(alpha, mu, t, r, g, y0):
    """test/analytical_solution.cpp:130-191: position, rotation angle, velocity and rotation velocity
    of a sphere dropped on a horizontal plane under tilted gravity, rolling without slipping when
    tan(alpha) <= 3 mu, else sliding with Coulomb friction."""
    t_impact = math.sqrt(2 * (y0 - r) / (g * math.cos(alpha)))
    if t <= t_impact:
        return (np.array([g * t * t * math.sin(alpha) / 2, -g * t * t * math.cos(alpha) / 2 + y0]), 0.0,
                np.array([g * t * math.sin(alpha), -g * t * math.cos(alpha)]), 0.0)
    v_t_m = g * t_impact * math.sin(alpha)
    v_n_m = -g * t_impact * math.cos(alpha)
    t2 = t - t_impact
    xt_impact = g * t_impact * t_impact * math.sin(alpha) / 2
    if math.tan(alpha) <= 3 * mu:
        x_t = g * math.sin(alpha) * t2 * t2 / 3.0 + 2.0 * v_t_m * t2 / 3.0 + xt_impact
        theta = -g * math.sin(alpha) * t2 * t2 / (3.0 * r) - 2 * v_t_m * t2 / (3.0 * r)
        v_t = 2.0 * g * math.sin(alpha) * t2 / 3.0 + 2.0 * v_t_m / 3.0
        omega = -2.0 * g * math.sin(alpha) * t2 / (3.0 * r) - 2 * v_t_m / (3.0 * r)
    else:
        x_t = g * (math.sin(alpha) - mu * math.cos(alpha)) * t2 * t2 / 2.0 + (v_t_m + mu * v_n_m) * t2 + xt_impact
        theta = -mu * g * math.cos(alpha) * t2 * t2 / r + 2 * mu * v_n_m * t2 / r
        v_t = g * (math.sin(alpha) - mu * math.cos(alpha)) * t2 + (v_t_m + mu * v_n_m)
        omega = -2.0 * mu * g * math.cos(alpha) * t2 / r + 2 * mu * v_n_m / r
    return np.array([x_t, r]), theta, np.array([v_t, 0.0]), omega


def sphere_plane_errors(case, state):
    """The four errors of test/test_gradient.cpp:264-277 (position, quaternion, velocity: relative
    norms; omega: std::norm of a double is its square, so (dw)^2 / w^2)."""
    t = case["dt"] * case["total_it"]
    pa, th, va, wa = analytical_sphere_plane_friction(case["alpha"], case["mu"], t, case["radius"], case["g"], case["h"])
    qa = np.array([math.cos(th / 2), 0.0, 0.0, math.sin(th / 2)])
    return dict(pos=np.linalg.norm(state["pos"][1, :2] - pa) / np.linalg.norm(pa),
                q=np.linalg.norm(state["q"][1] - qa) / np.linalg.norm(qa),
                v=np.linalg.norm(state["v"][1, :2] - va) / np.linalg.norm(va),
                omega=(state["omega"][1, 2] - wa) ** 2 / wa ** 2)


def doctest_le(x, bound, eps=float(np.finfo(np.float32).eps) * 100):
    """doctest's `x <= Approx(bound)` (epsilon 100 float eps, scale 1)."""
    return x < bound or abs(x - bound) < eps * (1.0 + max(abs(x), abs(bound)))


def analytical_sphere_plane_viscous(alpha, t, r, g, y0, gamma_min, t_inv):
    """test/analytical_solution.cpp:39-129: a sphere dropped on a horizontal plane under gravity
    tilted by alpha, Viscous, the gravity reversed at t_inv; position and velocity (no rotation)."""
    t_impact = math.sqrt(2 * (y0 - r) / (g * math.cos(alpha)))
    t_gamma_min = -gamma_min / (g * math.cos(alpha))
    gamma_inv = gamma_min if t_inv > t_gamma_min else -g * t_inv * math.cos(alpha)
    t_unstick = t_inv - gamma_inv / (g * math.cos(alpha))
    sa, ca = math.sin(alpha), math.cos(alpha)
    if t < t_impact:
        return np.array([g * t * t * sa / 2, -g * t * t * ca / 2 + y0]), np.array([g * t * sa, -g * t * ca])
    if t < t_inv:
        return np.array([g * t * t * sa / 2, r]), np.array([g * t * sa, 0.0])
    x0 = g * t_inv ** 2 * sa / 2 + g * t_inv * sa * (t - t_inv) - g * (t - t_inv) ** 2 * sa / 2
    v0 = g * t_inv * sa - g * (t - t_inv) * sa
    if t < t_unstick:
        return np.array([x0, r]), np.array([v0, 0.0])
    return np.array([x0, r + g * (t - t_unstick) ** 2 * ca / 2]), np.array([v0, g * (t - t_unstick) * ca])


def analytical_sphere_plane_viscous_friction(alpha, mu, t, r, g, y0, gamma_min, t_inv):
    """test/analytical_solution.cpp:193-419: the same drop under ViscousFriction; position, rotation
    angle, velocity and rotation velocity."""
    sa, ca, ta = math.sin(alpha), math.cos(alpha), math.tan(alpha)
    t_impact = math.sqrt(2 * (y0 - r) / (g * ca))
    t_gamma_min = -gamma_min / (g * ca)
    gamma_inv = gamma_min if t_inv > t_gamma_min else -g * t_inv * ca
    if t < t_impact:
        return (np.array([g * t * t * sa / 2, -g * t * t * ca / 2 + y0]), 0.0, np.array([g * t * sa, -g * t * ca]), 0.0)
    if t < t_gamma_min:
        return np.array([g * t * t * sa / 2, r]), 0.0, np.array([g * t * sa, 0.0]), 0.0
    v_t_gm = -gamma_min * ta
    x_t_gm = g * t_gamma_min ** 2 * sa / 2

    def sliding_then_rolling(tt):
        """(x_t, angle, v_t, omega) at tt >= t_gamma_min before the reversal."""
        t2 = tt - t_gamma_min
        slide = (g * (sa - mu * ca) * t2 ** 2 / 2 + v_t_gm * t2 + x_t_gm, -2. * mu * g * ca * t2 ** 2 / (2 * r),
                 g * (sa - mu * ca) * t2 + v_t_gm, -2. * mu * g * ca * t2 / r)
        if ta <= 3 * mu:
            t_ns = -v_t_gm / (g * ca * (ta - 3 * mu)) + t_gamma_min
            if tt > t_ns:
                vt_ns = g * (sa - mu * ca) * (t_ns - t_gamma_min) + v_t_gm
                om_ns = -2. * mu * g * ca * (t_ns - t_gamma_min) / r
                xt_ns = g * (sa - mu * ca) * (t_ns - t_gamma_min) ** 2 / 2 + v_t_gm * (t_ns - t_gamma_min) + x_t_gm
                an_ns = -2. * mu * g * ca * (t_ns - t_gamma_min) ** 2 / (2 * r)
                return (2. * g * sa * (tt - t_ns) ** 2 / 6. + vt_ns * (tt - t_ns) + xt_ns,
                        -2. * g * sa * (tt - t_ns) ** 2 / (6. * r) + om_ns * (tt - t_ns) + an_ns,
                        2. * g * sa * (tt - t_ns) / 3. + vt_ns, -2. * g * sa * (tt - t_ns) / (3. * r) + om_ns)
        return slide

    if t < t_inv:
        x, ang, v, om = sliding_then_rolling(t)
        return np.array([x, r]), ang, np.array([v, 0.0]), om
    x_i, ang_i, v_i, om_i = sliding_then_rolling(t_inv)
    t_unstick = t_inv - gamma_inv / (g * ca)
    x = np.array([v_i * (t - t_inv) - g * sa * (t - t_inv) ** 2 / 2 + x_i, r])
    v = np.array([v_i - g * sa * (t - t_inv), 0.0])
    if t >= t_unstick:
        x[1] = g * ca * (t - t_unstick) ** 2 / 2 + r
        v[1] = g * ca * (t - t_unstick)
    return x, om_i * (t - t_inv) + ang_i, v, om_i


def sphere_plane_reversed_gravity(problem, steps1, steps2, **prm):
    """test/test_gradient.cpp:86-155 / 284-373: a sphere of radius 1 at height 2 over a horizontal plane,
    force (g sin a, -g cos a), a = pi/4, g = 1, dt 0.01, steps1 steps, then the force reversed for
    steps2 steps; apgd tol 1e-7, alpha 0.1, vap_fpd."""
    radius, g, h, alpha = 1.0, 1.0, 2.0, PI / 4
    c = Container(2)
    c.plane((0.0, 0.0), PI / 2)
    c.sphere((0.0, h), radius, Property(mass=1.0, moment_inertia=radius * radius / 2.0,
                                        force=(g * math.sin(alpha), -g * math.cos(alpha))))
    return dict(scene=c.arrays(), dt=0.01, steps=(steps1, steps2), vap=FPD, alpha=alpha, radius=radius, g=g, h=h,
                reversed_force=(-g * math.sin(alpha), g * math.cos(alpha)),
                params=dict(tolerance=1e-7, max_ite=10000, alpha=0.1, dynamic_descent=1, problem=problem, **prm))


def pile_2d(seed=3):
    """18 disks (r 0.049) in 3 rows over a floor plane, seeded horizontal velocities, force (0, -10)."""
    rng = np.random.default_rng(seed)
    c = Container(2)
    c.plane((0.0, 0.0), PI / 2)
    for i in range(6):
        for j in range(3):
            p = (0.1 * i + 0.05 + rng.uniform(-0.005, 0.005), 0.05 + 0.1 * j + rng.uniform(0, 0.005))
            c.sphere(p, 0.049, Property(mass=1.0, moment_inertia=0.5 * 0.049 ** 2, velocity=(rng.uniform(-1, 1), 0.0),
                                        force=(0.0, -10.0)))
    return c.arrays()
