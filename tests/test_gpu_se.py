"""GPU parity of the superellipsoid path (closest points, contact set, friction solve, steps).

The superellipsoid closest-point solve (grid start + MINPACK hybrd1) is chaotic at
the rounding level: one ulp in an input can move the end point by ~1e-6 in the
distance.  The oracle therefore runs with its portable elementary functions
(oracle.set_se_math(True)), which the device evaluates with the same operations, and
the bar for closest points and contact sets is bit-for-bit equality.  Solves and
steps are held to the solver bars of the sphere tests.
"""
import math
import sys

import numpy as np
import pytest

from scopi_amd import hip, scenes
from scopi_amd.scenes import quaternion
from tests import oracle_problem as op

pytestmark = pytest.mark.gpu

KD3 = scenes.CONFIG3["kd_tree_radius"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture
def se_oracle(oracle_lib):
    oracle_lib.set_se_math(True)
    yield oracle_lib
    oracle_lib.set_se_math(False)


@pytest.fixture(scope="module")
def ctx():
    g = hip.HipScene(scenes.superellipsoids_3d(2, 2, 2, seed=1))
    yield g
    g.close()


def random_pairs(n, seed, gap_lo=-0.01, gap_hi=0.003):
    rng = np.random.default_rng(seed)
    qa = rng.normal(size=(n, 4)); qa /= np.linalg.norm(qa, axis=1, keepdims=True)
    qb = rng.normal(size=(n, 4)); qb /= np.linalg.norm(qb, axis=1, keepdims=True)
    ra = rng.uniform(0.008, 0.015, (n, 3)); rb = rng.uniform(0.008, 0.015, (n, 3))
    ea = rng.uniform(0.4, 1.6, (n, 2)); eb = rng.uniform(0.4, 1.6, (n, 2))
    ca = rng.uniform(-1, 1, (n, 3))
    cb = ca + np.array([0.035, 0.0, 0.0]) + rng.uniform(gap_lo, gap_hi, (n, 3))
    return ca, qa, ra, ea, cb, qb, rb, eb


def oracle_pairs(se, args):
    n = len(args[0])
    out = {k: [] for k in ("dij", "nij", "pi", "pj", "u0", "u", "info", "nfev")}
    for k in range(n):
        r = se.se_closest_points(*(a[k] for a in args))
        for key in out:
            out[key].append(r[key])
    return {k: np.array(v) for k, v in out.items()}


def test_se_pairs_bit_exact(se_oracle, ctx):
    """Random pairs (separated and overlapping) and the reference's known answers."""
    args = random_pairs(48, 5)
    c30, s30 = math.sqrt(3.0) / 2.0, 0.5
    known = [((-0.2, 0, 0), quaternion(0.0), (0.1, 0.2, 0.3), (1, 1), (0.2, 0, 0), quaternion(0.0), (0.1, 0.2, 0.3),
              (1, 1)),
             ((-0.4 * c30, 0.4 * s30, 0), quaternion(-math.pi / 6), (0.1, 0.4, 0.5), (1, 1),
              (0.4 * c30, -0.4 * s30, 0), quaternion(math.pi - math.pi / 6), (0.2, 0.3, 0.6), (1, 1))]
    args = tuple(np.concatenate([a, np.array([kk[m] for kk in known], np.float64)]) for m, a in enumerate(args))
    say(f"[se-pairs] n={len(args[0])}")
    g = ctx.se_closest_points(*args)
    o = oracle_pairs(se_oracle, args)
    say(f"  info counts gpu={np.bincount(g['info'])} nfev max={g['nfev'].max()}")
    for k in ("u0", "u", "info", "nfev", "dij", "nij", "pi", "pj"):
        assert np.array_equal(g[k], o[k]), (k, np.abs(np.asarray(g[k], float) - np.asarray(o[k], float)).max())
    assert abs(g["dij"][-2] - 0.2) < 1e-6 and abs(g["dij"][-1] - 0.5) < 1e-6


def hip_contacts(sc, prm):
    g = hip.HipScene(sc)
    nc = g.compute_contacts(prm)
    c = g.contacts()
    g.close()
    return nc, c


@pytest.mark.parametrize("method", [hip.KDTREE, hip.BRUTE_FORCE], ids=["kdtree", "brute_force"])
def test_se_contact_set_bit_exact(se_oracle, method):
    sc = scenes.superellipsoids_3d(5, 5, 4, seed=7) if method == hip.KDTREE else scenes.superellipsoids_3d(3, 3, 3,
                                                                                                           seed=8)
    prm = hip.params(problem=hip.FRICTION, mu=0.3, dmax=0.02, kd_tree_radius=KD3, contact_method=method)
    say(f"[se-contacts] method={method} n={sc.n}")
    nc, g = hip_contacts(sc, prm)
    o = se_oracle.Oracle(sc, problem=1, mu=0.3, dmax=0.02, kd_tree_radius=KD3, contact_method=method)
    assert o.compute_contacts() == nc and nc > 0
    oc = o.contacts()
    assert np.array_equal(g["ij"], oc["ij"])
    for k in ("dij", "nij", "pi", "pj"):
        assert np.array_equal(g[k], oc[k]), k
    say(f"  nc={nc}")


def se_case(seed=11):
    return scenes.superellipsoids_3d(4, 4, 4, seed=seed, velocity_scale=3.0)


def se_first_step(sc, prm, dt):
    g = hip.HipScene(sc)
    nc = g.compute_contacts(prm)
    g.set_a_priori_velocity(prm, dt)
    info = g.solve(prm, dt)
    sol = g.solution()
    g.close()
    return np.concatenate([sol["u"].ravel(), sol["w"].ravel()]), info, nc


def exact_optimum(se, sc, dt):
    o = se.Oracle(sc, problem=1, mu=0.3, dmax=0.02, kd_tree_radius=KD3, vap=1)
    o.compute_contacts()
    o.set_a_priori_velocity(dt)
    A, invM, c, U, nij = op.sparse_problem(o, sc, dt, friction=True)
    # a-priori velocities: vap_fpd with zero force keeps v, so U is the scene's velocity
    lam, u, L = op.exact_solution(A, invM, c, dt, U, friction=True, nij=nij, mu=0.3, dim=3)
    na = sc.n_active
    u_exact = np.concatenate([u[:3 * na], u[3 * na:]])
    return o, u_exact, L


@pytest.mark.parametrize("precision", [hip.FP64, hip.FP32_REFINE], ids=["fp64", "fp32_refine"])
def test_se_friction_solve(se_oracle, precision):
    """Friction mu 0.3 on superellipsoid contacts: the fixed-step path agrees with the oracle and the
    exact optimum to 1e-8; the adaptive path is within max(3 e_oracle, 1e-6) of the optimum."""
    dt = 1e-3
    sc = se_case()
    o, u_exact, L = exact_optimum(se_oracle, sc, dt)
    na = sc.n_active
    prm = hip.params(problem=hip.FRICTION, mu=0.3, dmax=0.02, kd_tree_radius=KD3, vap=hip.VAP_FPD,
                     dynamic_descent=0, alpha=1.0 / L, tolerance=1e-10, max_ite=200000, precision=precision)
    u, info, nc = se_first_step(sc, prm, dt)
    say(f"[se-solve] precision={precision} nc={nc} iterations={info['iterations']}")
    assert info["converged"] == 1 and nc == len(o.contacts()["dij"])
    np.testing.assert_allclose(u, u_exact, atol=1e-8)
    o.set_params(problem=1, mu=0.3, dmax=0.02, kd_tree_radius=KD3, vap=1, dynamic_descent=0, alpha=1.0 / L,
                 tolerance=1e-10, max_ite=200000)
    r = o.solve(dt)
    np.testing.assert_allclose(u, np.concatenate([r["u"].ravel(), r["w"].ravel()]), atol=1e-8)
    # reference defaults (dynamic descent), tol 1e-9.  Where the reference's stopping test
    # |l+ - l| < tol fires depends on the line-search path; its distance to the optimum over
    # initial steps alpha in [0.03, 0.07] calibrates the bar.
    prm = hip.params(problem=hip.FRICTION, mu=0.3, dmax=0.02, kd_tree_radius=KD3, vap=hip.VAP_FPD, tolerance=1e-9,
                     max_ite=200000, precision=precision)
    u, info, _ = se_first_step(sc, prm, dt)
    e_orc = []
    for a in (0.03, 0.04, 0.045, 0.05, 0.055, 0.06, 0.07):
        o.set_params(problem=1, mu=0.3, dmax=0.02, kd_tree_radius=KD3, vap=1, tolerance=1e-9, max_ite=200000, alpha=a)
        r = o.solve(dt)
        e_orc.append(float(np.abs(np.concatenate([r["u"].ravel(), r["w"].ravel()]) - u_exact).max()))
    e_gpu = float(np.abs(u - u_exact).max())
    say(f"  adaptive e_gpu={e_gpu:.3e} e_oracle(alpha 0.05)={e_orc[3]:.3e} max over alpha={max(e_orc):.3e} "
        f"iterations={info['iterations']}")
    assert info["converged"] == 1 and e_gpu <= max(3 * max(e_orc), 1e-6)
    assert na == 64


def test_se_steps_against_oracle(se_oracle):
    """One and four full steps (Friction mu 0.3, fixed step so both sides iterate the same map).

    One step starts from the same state, so the contacts are bit-identical and the positions
    agree to 1e-10 (dt x 1e-7 in velocity).  Later steps start from states that differ by the
    solver's rounding, and the superellipsoid solve can turn an ulp into a different end point
    (rarely a spurious root and a large impulse, as in the reference).  Per particle, the bar is
    calibrated on the oracle itself: |gpu - oracle| <= 20 x the oracle's own change under a
    one-ulp perturbation of the initial positions, + 1e-9."""
    dt = 1e-3
    sc = scenes.superellipsoids_3d(3, 3, 4, seed=13, velocity_scale=3.0)
    _, _, L = exact_optimum(se_oracle, sc, dt)
    kw = dict(problem=1, mu=0.3, dmax=0.02, kd_tree_radius=KD3, vap=1, dynamic_descent=0, alpha=0.5 / L,
              tolerance=1e-11, max_ite=400000)
    hprm = hip.params(problem=hip.FRICTION, mu=0.3, dmax=0.02, kd_tree_radius=KD3, vap=hip.VAP_FPD,
                      dynamic_descent=0, alpha=0.5 / L, tolerance=1e-11, max_ite=400000)

    def oracle_run(scene, steps):
        o = se_oracle.Oracle(scene, **kw)
        o.run(dt, steps)
        return o.state()

    def gpu_run(steps):
        g = hip.HipScene(sc)
        info = g.run(hprm, dt, steps)
        st = g.state()
        g.close()
        return st, info
    sg, _ = gpu_run(1)
    d1 = float(np.abs(sg["pos"] - oracle_run(sc, 1)["pos"]).max())
    say(f"[se-steps] 1 step max|dpos|={d1:.3e}")
    assert d1 < 1e-10  # dt x 1e-7 in velocity
    steps = 4
    ref = oracle_run(sc, steps)["pos"]
    spread = np.zeros(sc.n)
    for sign in (1.0, -1.0):
        pert = scenes.superellipsoids_3d(3, 3, 4, seed=13, velocity_scale=3.0)
        pert.pos[:] = np.nextafter(pert.pos, pert.pos + sign)
        spread = np.maximum(spread, np.abs(oracle_run(pert, steps)["pos"] - ref).max(axis=1))
    sg, info = gpu_run(steps)
    dpos = np.abs(sg["pos"] - ref).max(axis=1)
    say(f"[se-steps] {steps} steps max|dpos|={dpos.max():.3e} oracle 1-ulp spread max={spread.max():.3e}, "
        f"{int(np.count_nonzero(spread > 1e-9))}/{sc.n} particles with spread > 1e-9, contacts(last)={info['n_contacts']}")
    assert info["n_contacts"] > 0
    assert np.all(dpos <= 20.0 * spread + 1e-9)


# ------------------------------------------------------------------ sphere / plane against a superellipsoid
SPHERE, PLANE, SE = 0, 1, 3


def newton_pair_cases(n, seed):
    """(shape_i, ..., shape_j, ...) rows: sphere-SE, SE-sphere, plane-SE, SE-plane, in turn."""
    rng = np.random.default_rng(seed)
    qp = quaternion(-math.pi / 2, (0.0, 1.0, 0.0))
    rows = []
    for k in range(n):
        q = rng.normal(size=4); q /= np.linalg.norm(q)
        se = (SE, rng.uniform(-0.5, 0.5, 3), q, rng.uniform(0.008, 0.015, 3), rng.uniform(0.4, 1.6, 2))
        if k % 4 < 2:
            qs = rng.normal(size=4); qs /= np.linalg.norm(qs)
            other = (SPHERE, se[1] + np.array([0.03, 0, 0]) + rng.uniform(-0.005, 0.005, 3), qs,
                     (rng.uniform(0.008, 0.015), 0, 0), (1.0, 1.0))
        else:
            other = (PLANE, se[1] - np.array([0, 0, rng.uniform(0.01, 0.05)]), qp, (0, 0, 0), (1.0, 1.0))
        rows.append((other, se) if k % 2 == 0 else (se, other))
    cols = list(zip(*[a + b for a, b in rows]))
    return [np.array(c) for c in cols]


def test_newton_pairs_bit_exact(se_oracle, ctx):
    """Sphere / plane against superellipsoid pairs in both orders: records, starts, end points and
    Newton counts bitwise equal to the oracle's (NaN where the reference's Newton diverges)."""
    cols = newton_pair_cases(24, 21)
    say(f"[newton-pairs] n={len(cols[0])}")
    g = ctx.closest_points(*cols)
    for k in range(len(cols[0])):
        o = se_oracle.newton_closest_points(*(c[k] for c in cols))
        if not np.array_equal(g["u"][k], o["u"], equal_nan=True):
            say(f"  pair {k} shapes {cols[0][k]},{cols[5][k]}: u0 gpu {g['u0'][k]} oracle {o['u0']}; "
                f"info gpu {g['info'][k]} oracle {o['iterations']}; u gpu {g['u'][k]} oracle {o['u']}")
        for key, gk in (("dij", g["dij"][k]), ("nij", g["nij"][k]), ("pi", g["pi"][k]), ("pj", g["pj"][k]),
                        ("u0", g["u0"][k]), ("u", g["u"][k])):
            assert np.array_equal(np.asarray(gk), np.asarray(o[key]), equal_nan=True), (k, key, gk, o[key])
        assert g["info"][k] == o["iterations"], k
    say(f"  converged {int(np.count_nonzero(g['info'] >= 0))}/{len(cols[0])}")


def mixed_scene(seed=31, n=4):
    """A floor plane (obstacle), then a lattice alternating spheres and superellipsoids."""
    from scopi_amd.scenes import Container, Property
    rng = np.random.default_rng(seed)
    c = Container(3)
    c.plane((0.0, 0.0, 0.0), q=quaternion(-math.pi / 2, (0.0, 1.0, 0.0)))
    k = 0
    for i in range(n):
        for j in range(n):
            for l in range(n):
                pos = np.array([i, j, l + 0.6]) * 0.035 + rng.uniform(-0.003, 0.003, 3)
                prop = Property(mass=1.0, moment_inertia=(0.1, 0.1, 0.1), velocity=rng.uniform(-0.5, 0.5, 3))
                if k % 2 == 0:
                    c.sphere(pos, rng.uniform(0.008, 0.015), prop)
                else:
                    q = rng.normal(size=4); q /= np.linalg.norm(q)
                    c.superellipsoid(pos, q, rng.uniform(0.008, 0.015, 3), rng.uniform(0.4, 1.6, 2), prop)
                k += 1
    return c.arrays()


@pytest.mark.parametrize("method", [hip.KDTREE, hip.BRUTE_FORCE], ids=["kdtree", "brute_force"])
def test_mixed_scene_contacts_bit_exact(se_oracle, method):
    """Plane obstacle + spheres + superellipsoids: every shape pair kind in one contact set."""
    sc = mixed_scene()
    prm = hip.params(problem=hip.FRICTION, mu=0.3, dmax=0.02, kd_tree_radius=KD3, contact_method=method)
    say(f"[mixed-contacts] method={method} n={sc.n}")
    nc, g = hip_contacts(sc, prm)
    o = se_oracle.Oracle(sc, problem=1, mu=0.3, dmax=0.02, kd_tree_radius=KD3, contact_method=method)
    assert o.compute_contacts() == nc and nc > 0
    oc = o.contacts()
    assert np.array_equal(g["ij"], oc["ij"])
    for k in ("dij", "nij", "pi", "pj"):
        assert np.array_equal(g[k], oc[k]), k
    kinds = {(int(sc.shape[i]), int(sc.shape[j])) for i, j in g["ij"]}
    say(f"  nc={nc} pair kinds={sorted(kinds)}")


@pytest.mark.parametrize("fn", ["cos", "sin", "asin", "log", "exp", "pow"])
def test_portable_math_bit_exact(se_oracle, ctx, fn):
    """The device's elementary functions equal the oracle's portable ones bit for bit, over
    ordinary, large, tiny and special arguments."""
    rng = np.random.default_rng(6)
    if fn in ("cos", "sin"):
        x = np.concatenate([rng.uniform(-10, 10, 2000), rng.uniform(-1e13, 1e13, 500),
                            rng.choice([-1, 1], 500) * 10.0 ** rng.uniform(13, 15, 500), 10.0 ** rng.uniform(-320, -1, 200),
                            [0.0, -0.0, np.pi / 2, 1e15, -8.13671627e+14, 7.58242260e+12, np.inf, np.nan]])
    elif fn == "asin":
        x = np.concatenate([rng.uniform(-1, 1, 2000), [0.5, -0.5, 1.0, -1.0, 0.0, 1.5, np.nan]])
    elif fn == "log":
        x = np.concatenate([10.0 ** rng.uniform(-320, 308, 2500), [0.0, 1.0, np.inf, -1.0, np.nan, 5e-324]])
    elif fn == "exp":
        x = np.concatenate([rng.uniform(-750, 715, 2500), [0.0, -0.0, 709.78, 709.79, -745.1, -746.0, np.nan]])
    else:
        x = np.concatenate([10.0 ** rng.uniform(-320, 300, 2500), [0.0, 0.0, 0.0, 1.0, np.inf, np.nan]])
    y = rng.uniform(-1.5, 2.6, len(x)) if fn == "pow" else None
    if fn == "pow":
        y[-6:-3] = [0.5, 0.0, -0.5]
    g = ctx.portable_math(fn, x, y)
    o = np.array([se_oracle.portable_math(fn, a, 0.0 if y is None else b) for a, b in
                  zip(x, y if y is not None else np.zeros(len(x)))])
    bad = np.flatnonzero(~((g == o) | (np.isnan(g) & np.isnan(o))))
    assert bad.size == 0, [(x[k], None if y is None else y[k], g[k], o[k]) for k in bad[:5]]


def test_newton_residual_and_solve_bit_exact(se_oracle, ctx):
    """The device's residual, Jacobian and 4x4 solve of the Newton pairs equal the oracle's bit for bit
    (ordinary and diverged, very large arguments)."""
    rng = np.random.default_rng(8)
    qp = quaternion(-math.pi / 2, (0.0, 1.0, 0.0))
    bad = []
    for t in range(60):
        kind = t % 2
        q = rng.normal(size=4); q /= np.linalg.norm(q)
        se = (rng.uniform(-0.5, 0.5, 3), q, rng.uniform(0.008, 0.015, 3), rng.uniform(0.4, 1.6, 2))
        if kind == 0:
            qs = rng.normal(size=4); qs /= np.linalg.norm(qs)
            other = (se[0] + np.array([0.03, 0, 0]), qs, rng.uniform(0.008, 0.015))
        else:
            other = (se[0] - np.array([0, 0, 0.03]), qp, 0.0)
        u = rng.uniform(-3, 3, 4) * (10.0 ** rng.uniform(0, 14, 4) if t % 3 == 0 else 1.0)
        A = rng.normal(size=(4, 4)); b = rng.normal(size=4)
        F, J, x = ctx.newton_probe(kind, se, other, u, A, b)
        Fo, Jo = se_oracle.newton_residual(kind, se[0], se[1], se[2], se[3], other[0], other[1], (other[2], 0, 0), u,
                                           jacobian=True)
        xo = se_oracle.lu_solve4(A, b)
        for name, g, o in (("F", F, Fo), ("J", J, Jo), ("x", x, xo)):
            if not np.array_equal(g, o, equal_nan=True):
                bad.append((t, kind, name, u, g, o))
    for b_ in bad[:4]:
        say(f"  mismatch t={b_[0]} kind={b_[1]} {b_[2]} u={b_[3]}\n   gpu {b_[4]}\n   orc {b_[5]}")
    assert not bad
