"""The reference's demos written against the C++ slot mirror (tests/cpp/test_slots.cpp, `demos` group) and
run on the GPU: the demo code's scene set-up, slots and parameters, as a user of the reference writes them.

  * demos/box_spheres_3d.cpp: 1,000 spheres (n = 10, radii U(0.8 r0, 0.9 r0), masses U(1, 2), centres
    jittered by U(-0.05, 0.05)) falling in a box of 5 planes, force (0, -1, 0), NoFriction, APGD tol 1e-3
    with alpha = 0.2 / dt^2, vap_fpd, dmax 0.9 r0, kd_tree_radius dmax + 1.8 r0; 30 steps of dt = 0.045
    (the demo runs 1000; the spheres land after about 10 steps and are piled up by step 30);
  * demos/segment.cpp: 50 spheres driven through the gap between two segment obstacles at their desired
    velocity (default slots: vap_fixed), dmax 2 dt, kd_tree_radius 4 max_radius; 300 steps of dt = 0.001
    (the spheres reach the segments after about 220);
  * demos/sphere_plane.cpp (2D): a sphere on a 30-degree incline, FrictionFixedPoint mu 0.1, vap_fpd,
    dt 0.05; demos/two_ellipsoids.cpp (2D superellipsoids, default slots), dt 0.005;
    demos/pile_of_sand_spheres.cpp (27 spheres on a plane, FrictionFixedPoint mu 0.1, vap_fpd, masses from
    std::default_random_engine), dt 0.01; 200 steps each (the demos run 200, 1000 and 1000).

Bars:
  * the scene each program builds is bitwise the one scopi_amd.scenes builds from the oracle's copy of
    the demo's std::default_random_engine draws (so the oracle runs the same scene);
  * box_spheres_3d, 30 steps against the oracle: positions within 1e-8 (the trajectory bar of
    test_gpu_parity.test_trajectory_3d_small); the device and the oracle run the same iterations of the
    same line search, apart from the order of their sums; every sphere stays inside the box;
  * segment, 300 steps against the oracle (reference default tol 1e-6): positions within 1e-8 (the bar
    of test_gpu_segment's trajectory test);
  * sphere_plane, two_ellipsoids and pile_of_sand, 200 steps against the oracle: positions within 1e-8.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from scopi_amd import scenes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_slots")
BOX_STEPS, SEG_STEPS, MORE_STEPS = 30, 300, 200
N_SIDE, R0, G = 10, 0.5, 1.0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def demo_runs(tmp_path_factory):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (run __graft_entry__.build())")
    cwd = tmp_path_factory.mktemp("demos")  # the solver's output files land in ./Results here
    say(f"[demos] {BIN} demos {BOX_STEPS} {SEG_STEPS} {MORE_STEPS}")
    res = subprocess.run([BIN, "demos", str(BOX_STEPS), str(SEG_STEPS), str(MORE_STEPS)], capture_output=True,
                         text=True, timeout=240, cwd=str(cwd))
    assert res.returncode == 0, res.stderr
    out = {d["case"]: d for d in (json.loads(l) for l in res.stdout.splitlines() if l.startswith("{"))}
    out["_results"] = os.path.join(str(cwd), "Results")
    return out


def positions(d, dim):
    return np.asarray(d["pos"], np.float64).reshape(-1, dim)


def box_scene(oracle_lib):
    r, m, dx, dy, dz = oracle_lib.gen_box_spheres_demo(N_SIDE, R0)
    return scenes.box_spheres_demo(N_SIDE, R0, G, r, m, dx, dy, dz)


def test_box_spheres_3d_demo(oracle_lib, demo_runs):
    sc = box_scene(oracle_lib)
    init = positions(demo_runs["box_spheres_3d_initial"], 3)
    assert np.array_equal(init, sc.pos), "the program's scene differs from the builder's"
    dt = 0.2 * 0.9 * R0 / (2.0 * G)
    dmax = 0.9 * R0
    o = oracle_lib.Oracle(sc, problem=0, vap=1, dmax=dmax, kd_tree_radius=dmax + 2.0 * 0.9 * R0,
                          alpha=0.2 / (dt * dt), tolerance=1e-3)
    o.run(dt, BOX_STEPS)
    got = positions(demo_runs["box_spheres_3d"], 3)
    ref = o.state()["pos"]
    d = float(np.abs(got - ref).max())
    say(f"[demos] box_spheres_3d {BOX_STEPS} steps: max|pos_gpu - pos_oracle| = {d:.3e}")
    assert d <= 1e-8
    act = got[sc.n_inactive:]
    rad = sc.size[sc.n_inactive:, 0]
    w = 2.0 * N_SIDE * R0 + 2.0 * R0
    slack = 0.05 * R0
    assert np.all(act[:, 0] - rad > -slack) and np.all(act[:, 0] + rad < w + slack)
    assert np.all(act[:, 1] - rad > -slack)
    assert np.all(act[:, 2] - rad > -slack) and np.all(act[:, 2] + rad < w + slack)
    # the solver wrote an output file per step, as the demo's does (default output frequency 1)
    files = sorted(os.listdir(demo_runs["_results"]))
    assert len(files) >= BOX_STEPS + 1


def test_segment_demo(oracle_lib, demo_runs):
    x, y, r = oracle_lib.gen_segment_demo(50, 0.06)
    sc = scenes.segment_demo(x, y, r)
    init = positions(demo_runs["segment_initial"], 2)
    assert np.array_equal(init, sc.pos[:, :2]), "the program's scene differs from the builder's"
    o = oracle_lib.Oracle(sc, problem=0, vap=0, dmax=2 * 0.001, kd_tree_radius=4 * 0.06)
    o.run(0.001, SEG_STEPS)
    got = positions(demo_runs["segment"], 2)
    ref = o.state()["pos"][:, :2]
    d = float(np.abs(got - ref).max())
    say(f"[demos] segment {SEG_STEPS} steps: max|pos_gpu - pos_oracle| = {d:.3e}")
    assert d <= 1e-8


def demo_against_oracle(oracle_lib, demo_runs, name, sc, dt, **prm):
    dim = sc.dim
    init = positions(demo_runs[name + "_initial"], dim)
    assert np.array_equal(init, sc.pos[:, :dim]), "the program's scene differs from the builder's"
    o = oracle_lib.Oracle(sc, **prm)
    o.run(dt, MORE_STEPS)
    got = positions(demo_runs[name], dim)
    d = float(np.abs(got - o.state()["pos"][:, :dim]).max())
    say(f"[demos] {name} {MORE_STEPS} steps: max|pos_gpu - pos_oracle| = {d:.3e}")
    assert d <= 1e-8


def test_sphere_plane_demo(oracle_lib, demo_runs):
    demo_against_oracle(oracle_lib, demo_runs, "sphere_plane", scenes.sphere_plane_demo(), 0.05,
                        problem=oracle_lib.FRICTION_FIXED_POINT, vap=1, mu=0.1)


def test_two_ellipsoids_demo(oracle_lib, demo_runs):
    demo_against_oracle(oracle_lib, demo_runs, "two_ellipsoids", scenes.two_ellipsoids_demo(), 0.005)


def test_pile_of_sand_demo(oracle_lib, demo_runs):
    masses = oracle_lib.gen_uniform(27, 1.0, 2.0)
    demo_against_oracle(oracle_lib, demo_runs, "pile_of_sand", scenes.pile_of_sand_demo(masses), 0.01,
                        problem=oracle_lib.FRICTION_FIXED_POINT, vap=1, mu=0.1)
