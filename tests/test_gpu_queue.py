"""Per-scene work-queue kernel (SCOPI_TUNE_QUEUE_MAX_CONTACTS): a batch's small scenes solved whole
by one workgroup each, in one launch, through the C ABI.

Bars:
  * the first step's solve, reference defaults (adaptive line search), tol 1e-9: u no farther from
    the exact optimum (tests/golden/solve_cases.npz) than 3x the reference's own APGD gets, or 1e-6
    (the bar of test_gpu_parity.test_solve_adaptive_vs_exact_optimum), FP64 and FP32 + FP64, at
    every workgroup size;
  * no line search (alpha = 1/L): u within 1e-8 of the oracle's and of the exact optimum, APGD and
    PGD (both codes iterate the same map);
  * 20 steps of 8x8x4 spheres in the 5-plane box, reference defaults: positions within 1e-8 of the
    oracle (the bar of test_gpu_parity.test_trajectory_3d_small);
  * a queued scene's trajectory is bitwise the same alone (a batch of one) and in any mix, with
    scenes added and removed between steps, FP64 and FP32 + FP64, cold and warm start;
  * a batch whose scenes are split between the queue and the two-launch rounds leaves every scene
    bitwise as it is alone: the large scenes as in the one-scene context.
"""
import os
import sys

import numpy as np
import pytest

from scopi_amd import hip, scenes

from tests import solve_cases

pytestmark = pytest.mark.gpu

GOLDEN = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "solve_cases.npz"))
QUEUE_ALL = 1 << 30  # every scene is queued


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def queued_batch(prm, dt, threads=256, max_contacts=QUEUE_ALL):
    b = hip.HipBatch(prm, dt)
    b.set_tuning(hip.TUNE_QUEUE_MAX_CONTACTS, max_contacts)
    b.set_tuning(hip.TUNE_QUEUE_THREADS, threads)
    return b


def same_state(a, b, tag):
    for k in ("pos", "q", "v", "omega"):
        assert np.array_equal(a[k], b[k]), f"{tag}: {k} differs (max {np.abs(a[k] - b[k]).max():.3e})"


def case_params(case, **over):
    kw = dict(problem=case["problem"], vap=hip.VAP_FPD, dmax=case["dmax"], kd_tree_radius=case["kd"], mu=case["mu"],
              tolerance=1e-9, max_ite=200000)
    kw.update(over)
    return hip.params(**kw)


def first_step_u(case, prm, threads):
    b = queued_batch(prm, case["dt"], threads)
    sid = b.add(case["scene"]())
    b.step(1)
    sol, info = b.solution(sid), b.info(sid)
    b.close()
    return np.concatenate([sol["u"].ravel(), sol["w"].ravel()]), info


CASES_3D = [n for n in solve_cases.CASES if "_3d_" in n]


@pytest.mark.parametrize("threads", [256, 512, 1024])
@pytest.mark.parametrize("precision", [hip.FP64, hip.FP32_REFINE], ids=["fp64", "fp32_refine"])
@pytest.mark.parametrize("name", CASES_3D)
def test_queue_solve_vs_exact_optimum(name, precision, threads):
    case = solve_cases.CASES[name]
    prm = case_params(case, precision=precision)
    say(f"[queue-adaptive] {name} precision={precision} threads={threads}")
    u, info = first_step_u(case, prm, threads)
    assert info["n_contacts"] == int(GOLDEN[name + "/nc"]) and info["converged"] == 1, info
    e_gpu = float(np.abs(u - GOLDEN[name + "/u_exact"]).max())
    e_orc = float(GOLDEN[name + "/e_oracle"])
    say(f"  e_gpu={e_gpu:.3e} e_oracle={e_orc:.3e} iterations={info['iterations']} fp32={info['iterations_fp32']}")
    assert e_gpu <= max(3.0 * e_orc, 1e-6)
    if precision == hip.FP32_REFINE:
        assert 0 < info["iterations_fp32"] <= info["iterations"]


@pytest.mark.parametrize("method", [hip.APGD, hip.PGD], ids=["apgd", "pgd"])
@pytest.mark.parametrize("name", CASES_3D)
def test_queue_fixed_step_exact_path(oracle_lib, name, method):
    case = solve_cases.CASES[name]
    alpha = 1.0 / float(GOLDEN[name + "/L"])
    prm = case_params(case, dynamic_descent=0, alpha=alpha, tolerance=1e-10, method=method)
    say(f"[queue-fixed] {name} method={method}")
    u, info = first_step_u(case, prm, 256)
    assert info["converged"] == 1
    o = oracle_lib.Oracle(case["scene"](), **solve_cases.oracle_kw(case, dynamic_descent=0, alpha=alpha,
                                                                     tolerance=1e-10, method=method))
    o.compute_contacts()
    o.set_a_priori_velocity(case["dt"])
    r = o.solve(case["dt"])
    u_o = np.concatenate([r["u"].ravel(), r["w"].ravel()])
    say(f"  gpu iterations={info['iterations']} oracle iterations={o.stats()['iterations']}")
    np.testing.assert_allclose(u, u_o, atol=1e-8)
    np.testing.assert_allclose(u, GOLDEN[name + "/u_exact"], atol=1e-8)


def test_queue_trajectory_3d_small(oracle_lib):
    sc = scenes.spheres_3d(8, 8, 4)
    prm = hip.params(vap=hip.VAP_FPD, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-6)
    say("[queue-trajectory] 8x8x4 spheres, 20 steps")
    b = queued_batch(prm, 1e-3)
    sid = b.add(sc)
    b.step(20)
    o = oracle_lib.Oracle(sc, problem=prm.problem, contact_method=prm.contact_method, vap=prm.vap,
                          method=prm.method, dmax=prm.dmax, kd_tree_radius=prm.kd_tree_radius, alpha=prm.alpha,
                          max_ite=prm.max_ite, tolerance=prm.tolerance, dynamic_descent=prm.dynamic_descent,
                          mu=prm.mu)
    o.run(1e-3, 20)
    np.testing.assert_allclose(b.state(sid)["pos"], o.state()["pos"], atol=1e-8)


def mix(sizes, seed=5):
    return [scenes.spheres_3d_count(n, seed=seed + k, exact=True) for k, n in enumerate(sizes)]


@pytest.mark.parametrize("warm", [0, 1], ids=["cold", "warm"])
@pytest.mark.parametrize("precision", [hip.FP64, hip.FP32_REFINE], ids=["fp64", "fp32_refine"])
def test_queue_scene_independent_of_mix(precision, warm):
    prm = hip.params(vap=hip.VAP_FPD, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-6, precision=precision,
                     warm_start=warm)
    scs = mix([300, 1234, 2500, 777])
    target = scs[2]
    say(f"[queue] alone vs mix precision={precision} warm={warm}")
    alone = queued_batch(prm, 1e-3)
    a = alone.add(target)
    alone.step(4)
    m = queued_batch(prm, 1e-3)
    ids = [m.add(sc) for sc in scs[:2]]
    t = m.add(target)
    m.step(2)
    m.remove(ids[0])
    extra = [m.add(sc) for sc in mix([640, 2000], seed=11)]
    m.step(1)
    m.remove(extra[1])
    m.step(1)
    same_state(alone.state(a), m.state(t), "alone vs mix")
    ia, im = alone.info(a), m.info(t)
    for k in ("iterations", "iterations_fp32", "backtracks", "n_contacts", "steps"):
        assert ia[k] == im[k], k
    assert ia["converged"] == 1
    sa, sm = alone.solution(a), m.solution(t)
    for k in ("u", "w", "lam"):
        assert np.array_equal(sa[k], sm[k]), k


@pytest.mark.parametrize("warm", [0, 1], ids=["cold", "warm"])
@pytest.mark.parametrize("precision", [hip.FP64, hip.FP32_REFINE], ids=["fp64", "fp32_refine"])
def test_queue_split_batch(precision, warm):
    """Scenes below the threshold are queued (on the queue stream), the others take the two-launch
    rounds at the same time; every scene stays bitwise as it is alone, the large ones as in the
    one-scene context."""
    prm = hip.params(vap=hip.VAP_FPD, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-6, precision=precision,
                     warm_start=warm)
    sizes = [300, 4096, 1234, 2500]
    scs = mix(sizes)
    # contacts of each scene's first step, from a batch without the queue
    probe = hip.HipBatch(prm, 1e-3)
    pids = [probe.add(sc) for sc in scs]
    probe.step(1)
    ncs = [probe.info(p)["n_contacts"] for p in pids]
    probe.close()
    thr = int(sorted(ncs)[1]) + 1  # the two smallest scenes are queued (contact counts change little)
    say(f"[queue] split batch precision={precision} warm={warm} contacts={ncs} threshold={thr}")
    b = queued_batch(prm, 1e-3, max_contacts=thr)
    ids = [b.add(sc) for sc in scs]
    b.step(3)
    for sid, sc, nc in zip(ids, scs, ncs):
        if nc < thr:
            alone = queued_batch(prm, 1e-3, max_contacts=thr)
            a = alone.add(sc)
            alone.step(3)
            same_state(alone.state(a), b.state(sid), f"queued n={sc.n}")
            assert alone.info(a)["iterations"] == b.info(sid)["iterations"]
            alone.close()
        else:
            g = hip.HipScene(sc)
            g.run(prm, 1e-3, 3)
            same_state(g.state(), b.state(sid), f"rounds n={sc.n}")
            g.close()
        assert b.info(sid)["converged"] == 1


@pytest.mark.parametrize("name", ["lattice_3d_nofriction", "random_3d_nofriction"])
def test_queue_general_jacobian_layout(name):
    """The rigid layout (rotational blocks kept, general_jacobian=1) through the queue: the fixed-step
    solution within 1e-8 of the exact optimum, as test_gpu_parity.test_general_jacobian_layout holds
    the two-launch rounds."""
    case = solve_cases.CASES[name]
    alpha = 1.0 / float(GOLDEN[name + "/L"])
    prm = case_params(case, dynamic_descent=0, alpha=alpha, tolerance=1e-10, general_jacobian=1)
    say(f"[queue-rigid] {name}")
    u, info = first_step_u(case, prm, 512)
    assert info["converged"] == 1
    np.testing.assert_allclose(u, GOLDEN[name + "/u_exact"], atol=1e-8)


def test_queue_superellipsoid_scenes(oracle_lib):
    """Superellipsoid scenes (Friction mu 0.3, the cone layout over rotating bodies) in the queue: each
    scene bitwise the same alone and in a mix; its contacts bitwise those of the two-launch rounds
    (the same state); its multipliers meet the bars test_gpu_fullsize holds the 32k-superellipsoid
    step to: in the Coulomb cone (1e-9 of their scale), the gradient (formed by the oracle on the
    device's contacts) in the dual cone (1e-5 of its scale), |lam - P(lam - step g)| <= 10 tol."""
    from scopi_amd import scenes, serve
    from tests.cones import cone_projection
    kinds, _, prm = serve.hip_kinds(warm_start=False)
    p = prm["superellipsoids"]
    reqs = [serve.Request(rid=k, kind="superellipsoids", n=n, seed=30 + k, steps=1, arrival=0.0)
            for k, n in enumerate([125, 216, 343])]
    say("[queue] superellipsoid scenes")
    q = queued_batch(p, 1e-3, threads=512)
    ids = [q.add(serve.served_scene(r)) for r in reqs]
    q.step(1)
    alone = queued_batch(p, 1e-3, threads=512)
    a = alone.add(serve.served_scene(reqs[1]))
    alone.step(1)
    same_state(alone.state(a), q.state(ids[1]), "superellipsoids alone vs mix")
    rounds = hip.HipBatch(p, 1e-3)
    rids = [rounds.add(serve.served_scene(r)) for r in reqs]
    rounds.step(1)
    cfg = scenes.CONFIG3
    for sid, rid, r in zip(ids, rids, reqs):
        sc = serve.served_scene(r)
        ct, cr = q.contacts(sid), rounds.contacts(rid)
        for k in ct:
            assert np.array_equal(ct[k], cr[k]), f"n={r.n}: contacts {k}"
        info = q.info(sid)
        assert info["converged"] == 1, info
        nc = info["n_contacts"]
        lam = q.solution(sid)["lam"]
        o = oracle_lib.Oracle(sc, vap=1, problem=oracle_lib.FRICTION, mu=cfg["mu"], dmax=cfg["dmax"],
                              kd_tree_radius=cfg["kd_tree_radius"], tolerance=cfg["tolerance"])
        o.set_contacts(ct)
        o.set_a_priori_velocity(1e-3)
        gr = o.gradient(1e-3, lam).reshape(nc, 3)
        lam = lam.reshape(nc, 3)
        n = ct["nij"]
        ln = (lam * n).sum(1)
        lt = np.linalg.norm(lam - ln[:, None] * n, axis=1)
        gn = (gr * n).sum(1)
        gt = np.linalg.norm(gr - gn[:, None] * n, axis=1)
        ls, gs = float(np.abs(lam).max()), float(np.abs(gr).max())
        cone = float(np.max(lt - cfg["mu"] * ln))
        dual = float(np.max(cfg["mu"] * gt - gn))
        nat = float(np.linalg.norm(lam - cone_projection(lam - info["step"] * gr, n, cfg["mu"])))
        say(f"  n={r.n}: contacts {nc}, iterations {info['iterations']}; cone {cone:.3e}, dual {dual:.3e}, "
            f"|lam - P(lam - step g)| {nat:.3e}")
        assert cone <= 1e-9 * max(ls, 1e-300)
        assert dual <= 1e-5 * max(gs, 1e-300)
        assert nat <= 10 * cfg["tolerance"]
