"""The C++ slot API mirror (scopi_amd/host) driven like the reference's tests.

tests/cpp/test_slots (built by __graft_entry__.build()) sets up the fixture
scenes with ScopiSolver<dim, problem, OptimGradient<apgd>, contact_kdtree |
contact_brute_force, vap_fixed | vap_fpd> exactly as the reference's tests do,
runs them on the GPU and prints the final objects; they must match
test/references/*.json with the reference tests' tolerances.
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from tests import fixture_scenes as fs

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_slots")
TOL = {"two_spheres_asymmetrical": 1e-6, "two_spheres_symmetrical": 1e-6, "2d_case_spheres": 1e-6,
       "obstacles_sphere_sphere_moving": 1e-6, "3spheres_nofriction": 1e-5, "sphere_plane_friction_mu01": 1e-5,
       "3spheres_friction_fp": 1e-5, "3spheres_viscous": 1e-5, "3spheres_viscous_friction": 1e-5}


@pytest.fixture(scope="module")
def output(tmp_path_factory):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (run __graft_entry__.build())")
    cwd = tmp_path_factory.mktemp("slots")  # the output files land here, as in the reference's test runs
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=240, cwd=str(cwd))
    assert res.returncode == 0, res.stderr
    out = {d["case"]: d for d in (json.loads(l) for l in res.stdout.splitlines() if l.startswith("{"))}
    for d in out.values():
        if "output" in d:
            d["output"] = os.path.join(str(cwd), d["output"])
    return out


def same_within(ref, got, tol, where="objects"):
    """diffFile (test/utils.cpp:14-75): the same structure, floats within tol, everything else equal."""
    if isinstance(ref, float) and isinstance(got, float):
        assert abs(ref - got) <= tol, f"{where}: {got} vs {ref}"
    elif isinstance(ref, dict):
        assert isinstance(got, dict) and sorted(ref) == sorted(got), f"{where}: keys {sorted(got)} vs {sorted(ref)}"
        for k in ref:
            same_within(ref[k], got[k], tol, f"{where}/{k}")
    elif isinstance(ref, list):
        assert isinstance(got, list) and len(ref) == len(got), f"{where}: length"
        for k, (a, b) in enumerate(zip(ref, got)):
            same_within(a, b, tol, f"{where}/{k}")
    else:
        assert type(ref) is type(got) and ref == got, f"{where}: {got!r} vs {ref!r}"


@pytest.mark.parametrize("name", list(TOL))
def test_output_file_against_reference_file(output, name):
    """The file the mirror's ScopiSolver writes after the last step (its output path, set up as the
    reference's tests set it) against the reference's file: check_reference_file's comparison of
    "objects", plus the same contact records (keys and pairs)."""
    path = output[name]["output"]
    assert os.path.exists(path), path
    first = os.path.join(os.path.dirname(path), os.path.basename(path).rsplit("_", 1)[0] + "_0000.json")
    assert os.path.exists(first), first
    got = json.load(open(path))
    ref = fs.load_fixture(name)
    same_within(ref["objects"], got["objects"], TOL[name])
    rc, gc = ref["contacts"] or [], got["contacts"] or []
    assert [(c["i"], c["j"]) for c in rc] == [(c["i"], c["j"]) for c in gc]
    for a, b in zip(rc, gc):
        assert sorted(a) == sorted(b) and a["property"] == b["property"]


@pytest.mark.parametrize("name", list(TOL))
def test_fixture_through_slot_api(output, name):
    got = output[name]["objects"]
    n = len(got)
    st = dict(pos=np.zeros((n, 3)), q=np.zeros((n, 4)), v=np.zeros((n, 3)), omega=np.zeros((n, 3)))
    for k, o in enumerate(got):
        st["pos"][k, :2] = o["position"]
        st["q"][k] = o["quaternion"]
        st["v"][k, :2] = o["velocity"]
        st["omega"][k, 2] = o["omega"]
    err = fs.compare_objects(fs.load_fixture(name), 2, st, n, TOL[name])
    assert err <= TOL[name], f"{name}: {err}"


def test_contact_slot_then_optim_slot(output):
    c = output["contacts_kdtree"]
    r2 = math.sqrt(2.0) / 2.0
    assert c["n"] == 1 and (c["i"], c["j"]) == (0, 1)
    assert abs(c["dij"] - (math.sqrt(2.0) - 0.3)) < 1e-12
    np.testing.assert_allclose(c["nij"], [-r2, -r2], atol=1e-12)
    np.testing.assert_allclose(c["pi"], [0.1 * r2, 0.1 * r2], atol=1e-12)
    np.testing.assert_allclose(c["pj"], [1 - 0.2 * r2, 1 - 0.2 * r2], atol=1e-12)
    # closed form: lambda = -c / (dt^2 (1/m_i + 1/m_j)), c = dij + dt n.(v_i - v_j)
    o = output["optim_slot"]
    cval = (math.sqrt(2.0) - 0.3) - math.sqrt(2.0)
    lam = -cval / 2.0
    assert abs(o["lambda"] - lam) < 1e-8
    np.testing.assert_allclose(o["u0"], [1 - lam * r2, 1 - lam * r2], atol=1e-8)
    np.testing.assert_allclose(o["u1"], [lam * r2, lam * r2], atol=1e-8)


def test_superellipsoid_contact_slot(output):
    """The reference's SE3-SE3 known answers (test_closest_points.cpp:1632-1700) through the C++
    contact slot: superellipsoid<3> objects in a scopi_container, contact_kdtree<NoFriction>::run."""
    c30, s30 = math.sqrt(3.0) / 2.0, 0.5
    known = {"superellipsoids_3d_0": ((-0.1, 0, 0), (0.1, 0, 0), (-1.0, 0, 0), 0.2),
             "superellipsoids_3d_1": ((-0.3 * c30, 0.3 * s30, 0), (0.2 * c30, -0.2 * s30, 0), (-c30, s30, 0), 0.5),
             # superellipsoid<2> (test_closest_points.cpp:1502-1562)
             "superellipsoids_2d_0": ((-0.1, 0, 0), (0.1, 0, 0), (-1.0, 0, 0), 0.2),
             "superellipsoids_2d_1": ((-0.3 * c30, 0.3 * s30, 0), (0.2 * c30, -0.2 * s30, 0), (-c30, s30, 0), 0.5)}
    for name, (pi, pj, nij, dij) in known.items():
        c = output[name]
        assert c["n"] == 1
        # doctest::Approx default tolerance of the reference's test
        tol = lambda ref: np.finfo(np.float32).eps * 100 * (1.0 + abs(ref))
        assert abs(c["dij"] - dij) < tol(dij)
        for got, ref in ((c["pi"], pi), (c["pj"], pj), (c["nij"], nij)):
            for g, r in zip(got, ref):
                assert abs(g - r) < tol(r), (name, got, ref)


def test_periodic_box_through_slot_api(output, oracle_lib):
    """ScopiSolver(box, particles) with BoxDomain::with_periodicity(0): the two spheres meet across the
    periodic bound; the mirror's end state equals the oracle's (same box, same slots) to 1e-9 modulo L."""
    from scopi_amd.scenes import Container, Property
    got = output["periodic_pair"]["objects"]
    c = Container(2)
    c.sphere((0.012, 0.5), 0.01, Property(mass=1.0, moment_inertia=0.1, desired_velocity=(-0.5, 0.0)))
    c.sphere((0.985, 0.503), 0.012, Property(mass=1.0, moment_inertia=0.1, desired_velocity=(0.5, 0.0)))
    o = oracle_lib.Oracle(c.arrays(), dmax=0.02, kd_tree_radius=0.1 ** 2)
    o.set_box([0.0, 0.0], [1.0, 1.0], [True, False])
    o.run(0.005, 40)
    so = o.state()
    for k, ob in enumerate(got):
        d = np.array(ob["position"]) - so["pos"][k, :2]
        d[0] -= np.round(d[0])
        assert np.abs(d).max() <= 1e-9, (k, ob["position"], so["pos"][k])
    # the spheres met across the bound: the contact pushed them apart in y
    assert abs(got[0]["position"][1] - 0.5) > 1e-3 and abs(got[1]["position"][1] - 0.503) > 1e-3


@pytest.mark.parametrize("mu", [0.5, 0.1])
def test_fixed_point_tilted_gravity_through_solver(output, mu):
    """ScopiSolver<2, FrictionFixedPoint, ...> on test_gradient.cpp:215-282's scene: rolling (mu 0.5,
    the reference's bounds) and sliding (mu 0.1, Coulomb's closed form to 1e-8)."""
    d = output[f"sphere_plane_fp_mu{mu:g}"]
    o = d["objects"][0]
    st = dict(pos=np.zeros((2, 3)), q=np.zeros((2, 4)), v=np.zeros((2, 3)), omega=np.zeros((2, 3)))
    st["pos"][1, :2], st["q"][1], st["v"][1, :2], st["omega"][1, 2] = o["position"], o["quaternion"], o["velocity"], o["omega"]
    case = fs.sphere_plane_tilted_gravity(mu, 2)
    e = fs.sphere_plane_errors(case, st)
    assert fs.doctest_le(e["pos"], 1e-2) and fs.doctest_le(e["q"], 1e-2)
    if mu == 0.5:
        assert fs.doctest_le(e["v"], 1e-12) and fs.doctest_le(e["omega"], 1e-12)
    else:
        assert e["v"] < 1e-8 and abs(o["position"][1] - 1.0) < 1e-8
    assert d["outer_iterations"] > 300  # one or more outer passes per step in contact (from step 169 on)


def test_fixed_point_optim_slot_closed_form(output):
    d = output["fixed_point_slot"]
    assert d["n"] == 1 and d["outer_iterations"] > 1
    np.testing.assert_allclose(d["u"], [0.8, 0.0], atol=1e-7)
    assert abs(d["omega"] + 0.4) < 1e-7


def test_viscous_reversed_gravity_through_solver(output):
    """ScopiSolver<2, Viscous | ViscousFriction, ...> run, force changed, run again
    (test_gradient.cpp:86-155, 284-373): the reference's bounds against its analytical solutions."""
    o = output["reversed_viscous"]["objects"][0]
    pa, va = fs.analytical_sphere_plane_viscous(math.pi / 4, 10.0, 1.0, 1.0, 2.0, -2.0, 5.0)
    assert fs.doctest_le(np.linalg.norm(np.array(o["position"]) - pa) / np.linalg.norm(pa), 1e-3)
    assert fs.doctest_le(np.linalg.norm(np.array(o["velocity"]) - va) / np.linalg.norm(va), 1e-3)
    o = output["reversed_viscous_friction"]["objects"][0]
    pa, th, va, wa = fs.analytical_sphere_plane_viscous_friction(math.pi / 4, 0.5, 6.0, 1.0, 1.0, 2.0, -1.4, 5.0)
    assert fs.doctest_le(np.linalg.norm(np.array(o["position"]) - pa) / np.linalg.norm(pa), 1e-1)
    assert fs.doctest_le(np.linalg.norm(np.array(o["velocity"]) - va) / np.linalg.norm(va), 1e-3)
    assert fs.doctest_le((o["omega"] - wa) ** 2 / wa ** 2, 1e-3)


def test_viscous_slots_closed_form(output):
    d = output["viscous_slot"]
    assert d["n"] == 1 and d["rows"] == 2
    np.testing.assert_allclose(d["u"], [0.0, 0.0], atol=1e-7)
    assert abs(d["gamma"] + 1.0) < 1e-7
