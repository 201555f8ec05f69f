"""Warm start in the oracle (oracle_params.warm_start; not in the reference, whose APGD starts every
solve from lambda = 0, solvers/apgd.hpp:152-159): a NoFriction or Friction solve starts from the
last solve's multipliers of the contacts found again, matched by (i, j) as ScopiSolver::transfer
matches contacts (solver.hpp:241-259), new contacts at 0.  This is the checker of the device's warm
start, so it is pinned here by what the option must do:

* the first step has no history: warm and cold are bitwise the same;
* a resting stack (zero gaps, gravity, no motion) poses the same problem every step, so a solve
  started from the last solution stops at its first iteration; a transfer that gave a contact
  another contact's multiplier, or dropped the tangential rows, would not;
* the warm-started trajectory stays within the solver's reach of the cold one;
* warm_start = 0 is the reference's behaviour, bit for bit.
"""
import numpy as np
import pytest

from scopi_amd import scenes


def stack():
    # 4 x 4 x 4 spheres of radius 0.01 on a 0.02 lattice: every neighbour, the floor and the walls
    # touch at zero gap
    return scenes.spheres_3d(4, 4, 4, spacing=0.02, jitter=0.0, rmin=0.01, rmax=0.01)


CASES = {"nofriction": {}, "friction": dict(mu=0.3)}


def run_steps(oracle_lib, sc, case, warm, steps, **kw):
    pb = oracle_lib.FRICTION if case == "friction" else oracle_lib.NOFRICTION
    o = oracle_lib.Oracle(sc, vap=oracle_lib.VAP_FPD, problem=pb, warm_start=warm, **kw, **CASES[case])
    its = []
    for k in range(steps):
        o.run(1e-3, k + 1, k)
        its.append(o.stats()["iterations"])
    return o, its


STACK_KW = dict(dmax=0.005, kd_tree_radius=0.03 ** 2, tolerance=1e-8, max_ite=100000)


@pytest.mark.parametrize("case", list(CASES))
def test_first_step_is_cold(oracle_lib, case):
    sc = scenes.spheres_3d(4, 4, 2, velocity_scale=1.0)
    kw = dict(dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-8, max_ite=100000)
    a, ia = run_steps(oracle_lib, sc, case, 1, 1, **kw)
    b, ib = run_steps(oracle_lib, sc, case, 0, 1, **kw)
    assert ia == ib
    for k in ("pos", "v", "omega", "q"):
        np.testing.assert_array_equal(a.state()[k], b.state()[k])


@pytest.mark.parametrize("case", list(CASES))
def test_resting_stack_stops_at_first_iteration(oracle_lib, case):
    sc = stack()
    warm, iw = run_steps(oracle_lib, sc, case, 1, 8, **STACK_KW)
    cold, ic = run_steps(oracle_lib, sc, case, 0, 8, **STACK_KW)
    assert iw[0] == ic[0] and ic[0] > 100
    assert iw[1:] == [1] * 7, iw
    assert min(ic[1:]) > 100, ic
    # the same motion to the solver's reach (tolerance 1e-8 on the multiplier step)
    np.testing.assert_allclose(warm.state()["pos"], cold.state()["pos"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(warm.state()["v"], cold.state()["v"], rtol=0, atol=1e-6)


def test_warm_start_off_is_the_default(oracle_lib):
    sc = scenes.spheres_3d(4, 4, 2, velocity_scale=1.0)
    kw = dict(vap=oracle_lib.VAP_FPD, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-8, max_ite=100000)
    a = oracle_lib.Oracle(sc, **kw)
    b = oracle_lib.Oracle(sc, warm_start=0, **kw)
    a.run(1e-3, 5)
    b.run(1e-3, 5)
    for k in ("pos", "v", "omega", "q"):
        np.testing.assert_array_equal(a.state()[k], b.state()[k])
