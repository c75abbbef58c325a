"""Five-axis scan paths: quaternions, the rotated beam and the rotated
deposition boxes, pinned by the reference's known answers
(tests/test_quaternion.cc, tests/test_scan_path.cc:106-153,
tests/test_material_deposition.cc:697-908) at their tolerances, and by the
properties a rotation has (orthogonality, slerp end points, unit norm)."""
import math
import os

import numpy as np
import pytest

from adamantine_amd.deposition import deposition_along_scan_path
from adamantine_amd.problem import (SOURCE_GOLDAK, Beam, Problem,
                                    read_scan_path)
from oracle import oracle as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def close(a, b, tol):
    """BOOST_TEST(a == b) under utf::tolerance(tol): relative on both sides,
    absolute when one side is zero."""
    if a == b:
        return True
    if a == 0.0 or b == 0.0:
        return abs(a - b) <= tol
    d = abs(a - b)
    return d <= tol * abs(a) and d <= tol * abs(b)


def desc_of(segs):
    p = Problem(beams=[Beam(type=SOURCE_GOLDAK, absorption_efficiency=0.3,
                            max_power=400.0, radius=0.3, depth=0.5,
                            segments=segs)])
    d, keep = p.to_c()
    return d.beams[0], keep


def test_quaternion_rotation():
    """tests/test_quaternion.cc:9-37: a rotation of 2 pi / 3 about
    (1, 1, 1) cycles the axes; pow(0.5) squared and q / q are exact."""
    q = [0.5, 0.5, 0.5, 0.5]
    R = orc.quaternion_matrix(q)
    p1 = np.array([1.0, 0.0, 0.0])
    r1 = orc.rotate(R, p1)
    assert list(r1) == [0.0, 1.0, 0.0]
    r2 = orc.rotate(R, r1)
    assert list(r2) == [0.0, 0.0, 1.0]
    assert list(orc.rotate(R, r2)) == list(p1)
    assert list(orc.rotate(R, r1, inverse=True)) == list(p1)
    h = orc.quaternion_op("pow", q, e=0.5)
    assert list(orc.quaternion_op("mul", h, h)) == q
    assert list(orc.quaternion_op("div", q, q)) == [1.0, 0.0, 0.0, 0.0]


def test_quaternion_properties():
    """R is orthogonal with det 1 and turns the axis (i, j, k) into itself
    by the angle 2 atan2(|v|, r); the Hamilton product composes rotations."""
    rng = np.random.default_rng(5)
    for _ in range(20):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        R = orc.quaternion_matrix(q).reshape(3, 3)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(R) - 1.0) < 1e-14
        axis = q[1:] / np.linalg.norm(q[1:])
        assert np.allclose(R @ axis, axis, atol=1e-14)
        ang = 2.0 * math.atan2(np.linalg.norm(q[1:]), q[0])
        # a vector normal to the axis turns by the angle
        n = np.cross(axis, [1.0, 0.0, 0.0])
        n /= np.linalg.norm(n)
        assert abs(float(n @ (R @ n)) - math.cos(ang)) < 1e-13
        p = rng.normal(size=4)
        p /= np.linalg.norm(p)
        Rp = orc.quaternion_matrix(p).reshape(3, 3)
        Rqp = orc.quaternion_matrix(orc.quaternion_op("mul", q, p)).reshape(3, 3)
        assert np.allclose(Rqp, R @ Rp, atol=1e-14)
        # (q^(1/3))^3 = q
        c = orc.quaternion_op("pow", q, e=1.0 / 3.0)
        c3 = orc.quaternion_op("mul", orc.quaternion_op("mul", c, c), c)
        assert np.allclose(c3, q, atol=1e-14)


def test_event_series_with_rotations():
    """the 9-column event series (source/ScanPath.cc:143-195): time, point,
    quaternion, and a power that applies to the next segment."""
    segs = read_scan_path(os.path.join(GOLD, "scan_path_arcs.txt"),
                          "event_series")
    assert len(segs) == 4
    assert all(len(s) == 4 for s in segs)
    assert segs[0][1] == 0.0 and segs[1][1] == 1.0
    assert segs[2][3] == (0.70710678118, 0.0, 0.70710678118, 0.0)
    assert segs[2][2] == (0.5, 0.0, 0.0)


def test_five_axis_scan_path():
    """tests/test_scan_path.cc:106-153 (tolerance 1e-9): the path position
    rotated by the slerped quaternion."""
    segs = read_scan_path(os.path.join(GOLD, "scan_path_arcs.txt"),
                          "event_series")
    bd, keep = desc_of(segs)
    ref = {0.5: (0.2222222222, 0.0, 0.0), 1.0: (0.5, 0.0, 0.0),
           1.5: (0.35355339059327373, 0.0, -0.35355339059327373),
           2.0: (0.0, 0.0, -0.5), 2.5: (0.5, 0.25, -0.5),
           3.0: (1.0, 0.0, 0.0)}
    seg = 0
    for t, want in ref.items():
        s = orc.beam_state(bd, t, seg)
        p = np.array(s.center[:])
        R = orc.quaternion_matrix(orc.scan_path_quaternion(bd, t, seg))
        assert list(R) == list(s.rotation[:]) and s.rotated == 1
        got = orc.rotate(R, p)
        for a, b in zip(got, want):
            assert close(a, b, 1e-9), (t, got, want)


def test_slerp_end_points():
    """at a segment's end time the slerp returns that segment's rotation,
    and it stays on the unit sphere in between"""
    segs = read_scan_path(os.path.join(GOLD, "scan_path_5_axis.txt"),
                          "event_series")
    bd, keep = desc_of(segs)
    for k in range(1, len(segs)):
        q = orc.scan_path_quaternion(bd, segs[k][0])
        want = np.array(segs[k][3])
        assert np.allclose(q, want, atol=1e-12) or np.allclose(q, -want,
                                                               atol=1e-12)
        tm = 0.5 * (segs[k - 1][0] + segs[k][0])
        # the inputs are normed to 1e-10 (Quaternion::reinit)
        assert abs(np.linalg.norm(orc.scan_path_quaternion(bd, tm)) - 1) < 1e-10


def test_three_axis_beam_not_rotated(golden):
    segs = read_scan_path(golden("scan_path.txt"), "segment")
    bd, keep = desc_of(segs)
    s = orc.beam_state(bd, 1e-6)
    assert s.rotated == 0
    assert list(orc.scan_path_quaternion(bd, 1e-6)) == [1.0, 0.0, 0.0, 0.0]


def test_rotated_beam_value():
    """the Goldak value at point p on a five-axis path is the unrotated
    value at R p (source/GoldakHeatSource.cc:46-89)"""
    segs = read_scan_path(os.path.join(GOLD, "scan_path_arcs.txt"),
                          "event_series")
    bd, keep = desc_of(segs)
    t = 2.5
    s = orc.beam_state(bd, t)
    R = np.array(s.rotation[:])
    flat = orc.BeamState()
    C = type(s.center)
    for name, _ in orc.BeamState._fields_:
        if name not in ("rotated", "rotation"):
            v = getattr(s, name)
            setattr(flat, name, C(*v) if hasattr(v, "__len__") else v)
    rng = np.random.default_rng(2)
    c = np.array(s.center[:])
    n_nonzero = 0
    for _ in range(200):
        x = orc.rotate(R, c + rng.normal(scale=0.3, size=3), inverse=True)
        v = orc.beam_value(s, 3, x)
        assert v == orc.beam_value(flat, 3, orc.rotate(R, x))
        n_nonzero += v > 0
    assert n_nonzero > 20


# tests/test_material_deposition.cc:712-888: three directions of travel, each
# with the rotations (1,0,0,0), pi/2 about y and pi/2 about x
_L = 0.1 * math.sqrt(2.0) / 2.0
_O1, _O2 = 0.03786796564, 0.04289321882


def _reference_boxes():
    b = []
    for k in range(6):  # x direction
        b.append(((-0.05 + 0.1 * k, -0.05, -0.1), (0.05 + 0.1 * k, 0.05, 0.0)))
    for k in range(6):
        b.append(((-0.1, -0.05, -0.05 + 0.1 * k), (0.0, 0.05, 0.05 + 0.1 * k)))
    for k in range(6):
        b.append(((-0.05 + 0.1 * k, 0.0, -0.05), (0.05 + 0.1 * k, 0.1, 0.05)))
    for k in range(6):  # y direction
        b.append(((-0.05, -0.05 + 0.1 * k, -0.1), (0.05, 0.05 + 0.1 * k, 0.0)))
    for k in range(6):
        b.append(((-0.1, -0.05 + 0.1 * k, -0.05), (0.0, 0.05 + 0.1 * k, 0.05)))
    for k in range(6):
        b.append(((-0.05, 0.0, -0.05 - 0.1 * k), (0.05, 0.1, 0.05 - 0.1 * k)))
    lo = [(k - 1) * _L for k in range(8)] + [6 * _L + _O1]
    hi = [(k + 1) * _L for k in range(8)] + [7 * _L + _O2]
    for a, c in zip(lo, hi):  # y = x direction
        b.append(((a, a, -0.1), (c, c, 0.0)))
    for a, c in zip(lo, hi):
        b.append(((-0.1, a, a), (0.0, c, c)))
    for a, c in zip(lo, hi):
        b.append(((a, 0.0, -c), (c, 0.1, -a)))
    return b


def test_deposition_quaternion_straight_line():
    """tests/test_material_deposition.cc:697-908 (tolerance 1e-9)"""
    segs = read_scan_path(os.path.join(GOLD, "scan_path_5_axis.txt"),
                          "event_series")
    db = {"deposition_length": 0.1, "deposition_height": 0.1,
          "deposition_width": 0.1, "deposition_lead_time": 0.0}
    boxes, t, c, s = deposition_along_scan_path(db, segs, 3)
    ref = _reference_boxes()
    assert len(boxes) == len(ref) == 63
    for i, (lo, hi) in enumerate(ref):
        for j in range(3):
            assert close(boxes[i][j], lo[j], 1e-9), (i, j, boxes[i], lo)
            assert close(boxes[i][3 + j], hi[j], 1e-9), (i, j, boxes[i], hi)


def test_unnormed_quaternion_rejected(tmp_path):
    f = tmp_path / "bad.txt"
    f.write_text("0.1, 0, 0, 0, 1.0, 0.1, 0, 0, 1.0\n")
    with pytest.raises(ValueError):
        read_scan_path(str(f), "event_series")
