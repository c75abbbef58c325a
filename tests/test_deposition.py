"""Deposition boxes against the reference's known answers
(tests/test_material_deposition.cc), at its tolerances."""
import math
import os

import numpy as np
import pytest

from adamantine_amd.deposition import (deposition_along_scan_path,
                                       merge_deposition_paths,
                                       read_material_deposition)
from adamantine_amd.problem import read_scan_path

GOLD = os.path.join(os.path.dirname(__file__), "golden")
DB = {"deposition_length": 0.0005, "deposition_height": 0.1,
      "deposition_width": 0.1, "deposition_lead_time": 0.0}


def close(a, b, tol):
    """BOOST_TEST(a == b) under utf::tolerance(tol)"""
    if a == b:
        return True
    d = abs(a - b)
    return d <= tol * abs(a) and d <= tol * abs(b)


def path(name):
    return read_scan_path(os.path.join(GOLD, name), "segment")


@pytest.mark.parametrize("dim", [2, 3])
def test_read_deposition_file(dim):
    """read_material_deposition_file_{2d,3d} (:50-132)"""
    boxes, t, c, s = read_material_deposition(
        os.path.join(GOLD, "material_deposition_%dd.txt" % dim), dim)
    if dim == 2:
        ref = [((0., 0.), (1., 1.)), ((3., 2.4), (7., 10.)),
               ((4., 2.4), (7., 9.)), ((4., 3.4), (6., 9.))]
        ang = [0., 3.14, 1.57, -1.57]
    else:
        ref = [((0., 0., 0.), (1., 1., 1.)), ((3., 2.4, 1.), (7., 10., 2.)),
               ((4., 2.4, 2.), (7., 9., 4.)), ((4., 3.4, 3.), (6., 9., 8.))]
        ang = [0., -3.14, 1.57, -1.57]
    assert list(t) == [1., 2., 3.4, 4.98]
    assert all(close(x, math.cos(a), 1e-13) for x, a in zip(c, ang))
    assert all(close(x, math.sin(a), 1e-13) for x, a in zip(s, ang))
    assert len(boxes) == 4
    for b, (lo, hi) in zip(boxes, ref):
        # the reference compares the first two coordinates
        for d in range(2):
            assert close(b[d], lo[d], 1e-13) and close(b[dim + d], hi[d], 1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_deposition_along_straight_path(dim):
    """deposition_from_scan_path_{2d,3d} (:395-490)"""
    boxes, t, c, s = deposition_along_scan_path(DB, path("scan_path.txt"), dim)
    if dim == 2:
        first = ([-0.00025, 0.0], [0.00025, 0.1])
        fifth = ([0.002 - 0.00025, 0.0], [0.002 + 0.00025, 0.1])
    else:
        first = ([-0.00025, 0.05, 0.0], [0.00025, 0.15, 0.1])
        fifth = ([0.002 - 0.00025, 0.05, 0.0], [0.002 + 0.00025, 0.15, 0.1])
    for k, (lo, hi) in ((0, first), (4, fifth)):
        for d in range(dim):
            assert close(boxes[k][d], lo[d], 1e-13)
            assert close(boxes[k][dim + d], hi[d], 1e-13)
    for got, ref in zip(t, [1.0e-6, 6.26e-4, 1.251e-3, 1.876e-3, 2.501e-3]):
        assert close(got, ref, 1e-13)
    assert list(c[:5]) == [1.] * 5 and list(s[:5]) == [0.] * 5


def test_deposition_along_L_path():
    """deposition_from_L_scan_path_3d (:492-585)"""
    boxes, t, c, s = deposition_along_scan_path(DB, path("scan_path_L.txt"))
    ref = {0: ([-0.00025, -0.05, 0.0], [0.00025, 0.05, 0.1]),
           4: ([0.002 - 0.00025, -0.05, 0.0], [0.002 + 0.00025, 0.05, 0.1]),
           5: ([0.002 - 0.05, -0.00025, 0.0], [0.002 + 0.05, 0.00025, 0.1]),
           9: ([0.002 - 0.05, 0.00165, 0.0], [0.002 + 0.05, 0.00195, 0.1])}
    for k, (lo, hi) in ref.items():
        for d in range(3):
            assert close(boxes[k][d], lo[d], 1e-13), (k, d, boxes[k][d])
            assert close(boxes[k][3 + d], hi[d], 1e-13), (k, d, boxes[k][3 + d])
    tref = [1.0e-6, 6.26e-4, 1.251e-3, 1.876e-3, 2.501e-3, 2.501e-3, 3.126e-3,
            3.751e-3, 4.376e-3, 0.004751]
    for got, r in zip(t, tref):
        assert close(got, r, 1e-13)
    assert list(c[:10]) == [1.] * 5 + [0.] * 5
    assert list(s[:10]) == [0.] * 5 + [1.] * 5


def test_deposition_along_diagonal_path():
    """deposition_from_diagonal_scan_path_3d (:587-640), 1e-10"""
    boxes, t, c, s = deposition_along_scan_path(DB,
                                                path("scan_path_diagonal.txt"))
    ref = {0: ([-0.022584286572747879, -0.04483316294887079, 0.0],
               [0.022584286572747879, 0.04483316294887079, 0.1]),
           5: ([-0.020466252583997986, -0.043774145954495844, 0.0],
               [0.024466252583997983, 0.045774145954495846, 0.1])}
    for k, (lo, hi) in ref.items():
        for d in range(3):
            assert close(boxes[k][d], lo[d], 1e-10)
            assert close(boxes[k][3 + d], hi[d], 1e-10)
    tref = [1.0e-6, 6.26e-4, 1.251e-3, 1.876e-3, 2.501e-3,
            0.0027960849718747364]
    for got, r in zip(t, tref):
        assert close(got, r, 1e-10)
    assert close(c[0], 2. / math.sqrt(5.), 1e-10)
    assert close(s[0], 1. / math.sqrt(5.), 1e-10)


def test_merge_paths():
    """merge_multiple_deposition_paths (:642-700)"""
    b = [np.array([[0., 0., 1., 1.], [2., 2., 3., 3.]]),
         np.array([[1., 1., 2., 2.], [3., 3., 4., 4.]])]
    p1 = (b[0], np.array([0., 2.]), np.array([0., 0.2]), np.array([-0., -0.2]))
    p2 = (b[1], np.array([1., 3.]), np.array([0.1, 0.3]),
          np.array([-0.1, -0.3]))
    boxes, t, c, s = merge_deposition_paths([p1, p2])
    np.testing.assert_array_equal(boxes, [[0., 0., 1., 1.], [1., 1., 2., 2.],
                                          [2., 2., 3., 3.], [3., 3., 4., 4.]])
    assert list(t) == [0., 1., 2., 3.]
    assert list(c) == [0., 0.1, 0.2, 0.3] and list(s) == [-0., -0.1, -0.2, -0.3]
