"""Periodic boxes in the oracle (BoxDomain with periodic axes): pins from closed forms and from the
reference's own rules, which no reference test exercises (test/ has no periodic case).

  * a pair across a periodic boundary: the copy of the particle near the lower bound meets the
    particle near the upper bound; distance, normal and points equal the sphere-sphere closed form
    on the image, the copy's point moved back by L (contact/base.hpp:116-139);
  * the copy's point moves back along every periodic axis, also axes it was not copied along;
  * an obstacle near a lower bound is copied too, and meets itself and other obstacles as a zero
    record (closest_points(plane, plane), closest_points.hpp:750-753);
  * an object leaving the box is wrapped (solver.hpp:482-522), obstacles included;
  * a box without periodic axes changes nothing.
"""
import math

import numpy as np
import pytest

import oracle
from scopi_amd.scenes import Container, Property, quaternion


def two_spheres_2d(xa, xb, y=0.5, ra=0.01, rb=0.012):
    c = Container(2)
    c.sphere((xa, y), ra, Property(mass=1.0, moment_inertia=0.1))
    c.sphere((xb, y + 0.001), rb, Property(mass=1.0, moment_inertia=0.1))
    return c.arrays()


def test_pair_across_the_boundary_closed_form(oracle_lib):
    sc = two_spheres_2d(0.005, 0.99)
    o = oracle_lib.Oracle(sc, dmax=0.02, kd_tree_radius=0.1 ** 2)
    o.set_box([0.0, 0.0], [1.0, 1.0], [True, False])
    o.compute_contacts()
    c = o.contacts()
    assert [tuple(x) for x in c["ij"]] == [(1, 0)]  # i = the original, j = the copy's original
    ca = np.array([0.005 + 1.0, 0.5]); cb = np.array([0.99, 0.501])
    d = ca - cb
    u = d / np.linalg.norm(d)
    np.testing.assert_allclose(c["dij"][0], np.linalg.norm(d) - 0.01 - 0.012, rtol=0, atol=1e-15)
    np.testing.assert_allclose(c["pi"][0][:2], cb + 0.012 * u, rtol=0, atol=1e-15)
    np.testing.assert_allclose(c["pj"][0][:2], ca - 0.01 * u - np.array([1.0, 0.0]), rtol=0, atol=1e-15)
    assert abs(abs(c["nij"][0][0]) - abs(u[0])) < 1e-15


def test_no_copy_beyond_dmax_of_the_lower_bound(oracle_lib):
    sc = two_spheres_2d(0.03, 0.99)  # 0.03 - dmax >= 0: no copy, no contact
    o = oracle_lib.Oracle(sc, dmax=0.02, kd_tree_radius=0.1 ** 2)
    o.set_box([0.0, 0.0], [1.0, 1.0], [True, False])
    assert o.compute_contacts() == 0


def test_copy_point_moves_back_along_every_periodic_axis(oracle_lib):
    sc = two_spheres_2d(0.005, 0.99)
    o = oracle_lib.Oracle(sc, dmax=0.02, kd_tree_radius=0.1 ** 2)
    o.set_box([0.0, 0.0], [1.0, 2.0], [True, True])  # copied along x only (y = 0.5 is far from 0)
    o.compute_contacts()
    c = o.contacts()
    assert [tuple(x) for x in c["ij"]] == [(1, 0)]
    ca = np.array([1.005, 0.5]); cb = np.array([0.99, 0.501])
    u = (ca - cb) / np.linalg.norm(ca - cb)
    np.testing.assert_allclose(c["pj"][0][:2], ca - 0.01 * u - np.array([1.0, 2.0]), rtol=0, atol=1e-15)


def test_obstacle_copies_meet_as_zero_records(oracle_lib):
    c = Container(2)
    c.plane((0.3, 0.01), math.pi / 2)  # obstacle, y-normal plane 0.01 above y = 0
    c.sphere((0.3, 0.5), 0.01, Property(mass=1.0, moment_inertia=0.1))
    sc = c.arrays()
    o = oracle_lib.Oracle(sc, dmax=0.02, kd_tree_radius=0.1 ** 2)
    o.set_box([0.0, 0.0], [1.0, 1.0], [False, True])
    o.compute_contacts()
    ct = o.contacts()
    pairs = [tuple(x) for x in ct["ij"]]
    assert (0, 0) in pairs  # the plane and its own copy
    k = pairs.index((0, 0))
    assert ct["dij"][k] == 0.0 and not np.any(ct["nij"][k]) and not np.any(ct["pi"][k])
    np.testing.assert_array_equal(ct["pj"][k][:2], [0.0, -1.0])  # the copy's zero point, moved back by L


def test_wrap_after_the_move(oracle_lib):
    c = Container(2)
    c.sphere((0.995, 0.5), 0.01, Property(mass=1.0, moment_inertia=0.1, velocity=(1.0, 0.0),
                                          desired_velocity=(1.0, 0.0)))
    sc = c.arrays()
    o = oracle_lib.Oracle(sc, dmax=0.02, kd_tree_radius=0.1 ** 2)
    o.set_box([0.0, 0.0], [1.0, 1.0], [True, False])
    o.run(0.01, 1)
    x = o.state()["pos"][0, 0]
    assert x == (0.995 + 0.01 * 1.0) - 1.0


def test_box_without_periodic_axes_changes_nothing(oracle_lib):
    rng = np.random.default_rng(3)
    c = Container(2)
    for _ in range(80):
        c.sphere(rng.uniform(0, 1, 2), rng.uniform(0.02, 0.05), Property(mass=1.0, moment_inertia=0.1))
    sc = c.arrays()
    a = oracle_lib.Oracle(sc, dmax=0.05, kd_tree_radius=0.15 ** 2)
    b = oracle_lib.Oracle(sc, dmax=0.05, kd_tree_radius=0.15 ** 2)
    b.set_box([0.0, 0.0], [1.0, 1.0], [False, False])
    a.compute_contacts(); b.compute_contacts()
    ca, cb = a.contacts(), b.contacts()
    for k in ca:
        assert np.array_equal(ca[k], cb[k])
