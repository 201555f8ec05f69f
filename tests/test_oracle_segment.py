"""Pins the CPU restatement of closest_points(sphere<2>, segment<2>) and (segment<2>, sphere<2>)
(closest_points.hpp:824-881, with segment::normal / tangent / extrema, segment.hpp:154-177).

The reference's tests hold no segment case, so the restatement is pinned by what the geometry
fixes:
  * brute force: the distance from the sphere's centre to a dense sampling of the segment equals
    dij + r, and pj is the sampled closest point, in all three regions of the closed form (beyond
    either end, along the interior), for rotated segments;
  * the record's own relations: nij a unit vector along c - pj, pi = c - r nij, dij = (pi - pj).nij;
  * the argument order: (segment, sphere) negates nij and swaps the points (:874-881);
  * sign(0): a centre on the segment's line gives the value-initialised normal and dij = 0;
  * the two-point constructor (segment.hpp:133-138): its extrema are the two points.
"""
import math

import numpy as np
import pytest

from scopi_amd.scenes import quaternion

SPHERE, SEGMENT = 0, 2


def segment_frame(angle):
    """Normal (first column of the 2D rotation) and tangent (-n1, n0) of a segment at `angle`."""
    n = np.array([math.cos(angle), math.sin(angle)])
    return n, np.array([-n[1], n[0]])


def pair(oracle_lib, c, r, pos, angle, length, order="sphere_segment"):
    q = quaternion(angle)
    if order == "sphere_segment":
        return oracle_lib.closest_points(2, SPHERE, c, quaternion(0.0), (r, 0, 0), SEGMENT, pos, q, (length, 0, 0))
    return oracle_lib.closest_points(2, SEGMENT, pos, q, (length, 0, 0), SPHERE, c, quaternion(0.0), (r, 0, 0))


@pytest.mark.parametrize("region", ["end0", "end1", "interior"])
def test_sphere_segment_brute_force(oracle_lib, region):
    rng = np.random.default_rng({"end0": 1, "end1": 2, "interior": 3}[region])
    for _ in range(40):
        pos = rng.uniform(-1, 1, 2)
        angle = rng.uniform(-math.pi, math.pi)
        length = rng.uniform(0.2, 1.0)
        r = rng.uniform(0.01, 0.1)
        n, t = segment_frame(angle)
        along = {"end0": rng.uniform(-1.5, -0.55), "end1": rng.uniform(0.55, 1.5),
                 "interior": rng.uniform(-0.45, 0.45)}[region] * length
        c = pos + along * t + rng.uniform(-0.3, 0.3) * n
        rec = pair(oracle_lib, c, r, pos, angle, length)
        s = np.linspace(-0.5 * length, 0.5 * length, 200001)
        pts = pos[None, :] + s[:, None] * t[None, :]
        d = np.linalg.norm(pts - c[None, :], axis=1)
        k = int(np.argmin(d))
        # dense sampling: the closest sample is within half a step of the closed form
        step = length / 200000
        assert abs((rec["dij"] + r) - d[k]) < step, (region, rec["dij"] + r, d[k])
        assert np.linalg.norm(rec["pj"] - pts[k]) < step
        if region != "interior":
            end = pos + (0.5 if region == "end1" else -0.5) * length * t
            assert np.allclose(rec["pj"], end, atol=1e-15)
        # relations of the record
        u = c - rec["pj"]
        assert abs(np.linalg.norm(rec["nij"]) - 1) < 1e-14
        np.testing.assert_allclose(rec["nij"], u / np.linalg.norm(u), atol=1e-14)
        np.testing.assert_allclose(rec["pi"], c - r * rec["nij"], atol=1e-15)
        assert abs(rec["dij"] - (rec["pi"] - rec["pj"]) @ rec["nij"]) < 1e-15
        # the other argument order (closest_points.hpp:874-881)
        rev = pair(oracle_lib, c, r, pos, angle, length, order="segment_sphere")
        assert rev["dij"] == rec["dij"]
        assert np.array_equal(rev["nij"], -rec["nij"])
        assert np.array_equal(rev["pi"], rec["pj"]) and np.array_equal(rev["pj"], rec["pi"])


def test_sphere_on_segment_line(oracle_lib):
    """sign(0) = 0 (src/utils.cpp:47-50): a centre on the segment's line (interior) gives the zero
    normal, pi = pj = the centre and dij = 0; just off the line the normal points to the sphere."""
    pos = np.array([0.25, -0.5])
    for along in (-0.3, 0.0, 0.2):
        c = pos + np.array([0.0, along])  # angle 0: normal (1, 0), tangent (0, 1)
        rec = pair(oracle_lib, c, 0.05, pos, 0.0, 1.0)
        assert rec["dij"] == 0.0 and np.array_equal(rec["nij"], [0.0, 0.0])
        assert np.array_equal(rec["pi"], c) and np.array_equal(rec["pj"], c)
        off = pair(oracle_lib, c + np.array([1e-3, 0.0]), 0.05, pos, 0.0, 1.0)
        np.testing.assert_allclose(off["nij"], [1.0, 0.0], atol=1e-15)
        assert abs(off["dij"] - (1e-3 - 0.05)) < 1e-15


def test_two_point_segment(oracle_lib):
    """segment(pos1, pos2): centre 0.5 (pos1 + pos2), rotation quaternion(atan2(x1 - x2, y2 - y1)),
    length |pos2 - pos1| (segment.hpp:133-138); its extrema R (0, -+L/2) + pos (:170-177) are the two
    points: a sphere beyond either end touches that end."""
    from scopi_amd.scenes import Container
    rng = np.random.default_rng(4)
    for _ in range(30):
        p1, p2 = rng.uniform(-1, 1, 2), rng.uniform(-1, 1, 2)
        c = Container(2)
        c.segment_points(p1, p2)
        sc = c.arrays()
        pos, q, length = sc.pos[0, :2], sc.q[0], sc.size[0, 0]
        d = (p2 - p1) / np.linalg.norm(p2 - p1)
        for end, other, sign in ((p1, p2, -1.0), (p2, p1, 1.0)):
            cs = end + sign * 0.3 * d
            rec = oracle_lib.closest_points(2, SPHERE, cs, quaternion(0.0), (0.01, 0, 0), SEGMENT, pos, q,
                                            (length, 0, 0))
            np.testing.assert_allclose(rec["pj"], end, atol=1e-14)
            assert abs(rec["dij"] - (0.3 - 0.01)) < 1e-14
