"""FrictionFixedPoint on the GPU (SCOPI_PROBLEM_FRICTION_FIXED_POINT): the outer fixed point on
s_ij (OptimGradient.hpp:86-166) around the device projected-gradient solve.

* against the oracle on small 2D and 3D scenes with sliding contacts, adaptive APGD tol 1e-10,
  fixed_point_tol 1e-8, 20 steps.  Stated tolerance: positions within 1e-8, velocities within
  1e-6, rotation velocities within 1e-5 (the oracle's own trajectories at tol 1e-9 and 1e-10
  differ by 2.2e-9 / 9.4e-8 / 3.1e-6 on the 2D pile, whose disks have J = 1.2e-3, and by
  2.6e-10 / 7.9e-9 / 1.5e-8 on the 3D box: both codes stop each solve at the tolerance);
* Coulomb's closed form for a sliding sphere (mu 0.1, test/analytical_solution.cpp:130-191) and
  the reference's bounds for the rolling one (test/test_gradient.cpp:215-282);
* the outer-loop bound and the error paths.
"""
import math

import numpy as np
import pytest

from scopi_amd import hip, scenes
from scopi_amd.scenes import Container, Property
from tests import fixture_scenes as fs

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


CASES = {
    "pile_2d": (fs.pile_2d, dict(vap=hip.VAP_FPD, dmax=0.02, kd_tree_radius=0.2 ** 2), 0.01),
    "box_3d": (lambda: scenes.spheres_3d(3, 3, 2, velocity_scale=1.0),
               dict(vap=hip.VAP_FPD, dmax=0.02, kd_tree_radius=0.042 ** 2), 1e-3),
}


@pytest.mark.parametrize("name", list(CASES))
def test_trajectory_against_oracle(oracle_lib, name):
    make, kw, dt = CASES[name]
    sc = make()
    prm = hip.params(problem=hip.FRICTION_FIXED_POINT, mu=0.3, tolerance=1e-10, max_ite=100000,
                     fixed_point_tol=1e-8, **kw)
    say(f"[fixed-point] {name}")
    g = hip.HipScene(sc)
    info = g.run(prm, dt, 20)
    o = oracle_lib.Oracle(sc, problem=oracle_lib.FRICTION_FIXED_POINT, mu=0.3, tolerance=1e-10, max_ite=100000,
                          fixed_point_tol=1e-8, **kw)
    o.run(dt, 20)
    sg, so = g.state(), o.state()
    dpos = float(np.abs(sg["pos"] - so["pos"]).max())
    say(f"  contacts={info['n_contacts']} outer={info['outer_iterations']} max|dpos|={dpos:.3e} "
        f"max|dv|={float(np.abs(sg['v'] - so['v']).max()):.3e}")
    assert info["n_contacts"] > 0 and info["outer_iterations"] >= 20
    np.testing.assert_allclose(sg["pos"], so["pos"], atol=1e-8)
    np.testing.assert_allclose(sg["v"], so["v"], atol=1e-6)
    np.testing.assert_allclose(sg["omega"], so["omega"], atol=1e-5)


@pytest.mark.parametrize("mu", [0.5, 0.1])
def test_tilted_gravity_analytic(mu):
    case = fs.sphere_plane_tilted_gravity(mu, hip.FRICTION_FIXED_POINT)
    g = hip.HipScene(case["scene"])
    g.run(hip.params(vap=case["vap"], **case["params"]), case["dt"], case["total_it"])
    st = g.state()
    e = fs.sphere_plane_errors(case, st)
    say(f"[fixed-point] tilted gravity mu={mu}: {e}")
    assert fs.doctest_le(e["pos"], 1e-2) and fs.doctest_le(e["q"], 1e-2)
    if mu == 0.5:  # rolling: the reference test's bounds
        assert fs.doctest_le(e["v"], 1e-12) and fs.doctest_le(e["omega"], 1e-12)
    else:  # sliding: Coulomb's closed form, on the plane
        assert e["v"] < 1e-8 and abs(st["pos"][1, 1] - 1.0) < 1e-8


@pytest.mark.parametrize("max_iter", [1, 3])
def test_outer_loop_bound(oracle_lib, max_iter):
    """fixed_point_max_iter caps the passes per step (the sliding contact needs more than 3)."""
    case = fs.sphere_plane_tilted_gravity(0.1, hip.FRICTION_FIXED_POINT, total_it=300)
    case["params"]["fixed_point_max_iter"] = max_iter
    g = hip.HipScene(case["scene"])
    g.run(hip.params(vap=case["vap"], **case["params"]), case["dt"], case["total_it"] - 1)
    info = g.run(hip.params(vap=case["vap"], **case["params"]), case["dt"], case["total_it"],
                 initial_iter=case["total_it"] - 1)
    o = oracle_lib.Oracle(case["scene"], vap=case["vap"], **case["params"])
    o.run(case["dt"], case["total_it"])
    assert info["outer_iterations"] == max_iter == o.stats()["fixed_point_iterations"]
    np.testing.assert_allclose(g.state()["pos"], o.state()["pos"], atol=1e-8)


def test_batch_refuses_fixed_point():
    with pytest.raises(hip.ScopiHipError):
        hip.HipBatch(hip.params(problem=hip.FRICTION_FIXED_POINT, mu=0.3), 1e-3)
