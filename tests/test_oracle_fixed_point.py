"""FrictionFixedPoint in the CPU restatement, pinned by the reference's own tests and by the
closed form of Coulomb sliding.

* test/test_gradient.cpp:215-282 ("sphere - plane friction fixed point"): mu 0.5 on a plane under
  gravity tilted by pi/4 rolls without slipping (tan a <= 3 mu); the bounds are the reference's.
* With mu 0.1 the sphere slides (tan a > 3 mu).  The De Saxce cone of Friction lifts a sliding
  contact off the plane (normal velocity mu |u_t|), the fixed point on s_ij removes that lift, so
  FrictionFixedPoint follows the Coulomb solution of test/analytical_solution.cpp:130-191 while
  Friction does not: a dropped or wrongly scaled s_ij fails the sliding check.
* 3spheres_friction_fp.json (the reference's fixture) is covered with the other fixtures.
"""
import pytest

from tests import fixture_scenes as fs


def run(oracle_lib, case):
    o = oracle_lib.Oracle(case["scene"], vap=case["vap"], **case["params"])
    o.run(case["dt"], case["total_it"])
    return o


def test_rolling_matches_reference_bounds(oracle_lib):
    case = fs.sphere_plane_tilted_gravity(0.5, oracle_lib.FRICTION_FIXED_POINT)
    e = fs.sphere_plane_errors(case, run(oracle_lib, case).state())
    assert fs.doctest_le(e["pos"], 1e-2) and fs.doctest_le(e["q"], 1e-2)
    assert fs.doctest_le(e["v"], 1e-12) and fs.doctest_le(e["omega"], 1e-12)


def test_sliding_follows_coulomb(oracle_lib):
    case = fs.sphere_plane_tilted_gravity(0.1, oracle_lib.FRICTION_FIXED_POINT)
    o = run(oracle_lib, case)
    st = o.state()
    e = fs.sphere_plane_errors(case, st)
    assert e["v"] < 1e-8 and e["omega"] < 1e-12
    assert e["pos"] < 1e-2 and e["q"] < 1e-2
    assert abs(st["pos"][1, 1] - 1.0) < 1e-8      # stays on the plane
    assert abs(st["v"][1, 1]) < 1e-8
    assert o.stats()["fixed_point_iterations"] > 1  # the last step needed the outer loop


def test_convexified_friction_lifts_off_when_sliding(oracle_lib):
    """The same scene under Friction: the contact separates at rate ~mu |u_t| dt per step."""
    case = fs.sphere_plane_tilted_gravity(0.1, oracle_lib.FRICTION)
    st = run(oracle_lib, case).state()
    e = fs.sphere_plane_errors(case, st)
    assert e["v"] > 1e-5
    assert st["pos"][1, 1] > 1.0 + 1e-3


@pytest.mark.parametrize("max_iter", [1, 2])
def test_outer_loop_bounded_by_max_iter(oracle_lib, max_iter):
    case = fs.sphere_plane_tilted_gravity(0.1, oracle_lib.FRICTION_FIXED_POINT, total_it=300)
    case["params"]["fixed_point_max_iter"] = max_iter
    o = run(oracle_lib, case)
    assert o.stats()["fixed_point_iterations"] == max_iter
