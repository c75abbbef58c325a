"""Scan path and heat sources of the oracle, pinned by the reference's own
closed-form tests (tests/test_scan_path.cc, tests/test_heat_source.cc)."""
import math

import numpy as np
import pytest

from adamantine_amd.problem import (SOURCE_ELECTRON_BEAM, SOURCE_GOLDAK, Beam,
                                    Problem, read_scan_path)
from oracle import oracle as orc

REL = 1e-12


def close(a, b, tol=REL):
    return abs(a - b) <= tol * max(abs(a), abs(b)) or a == b


def beam_desc(kind, golden, depth=0.1, absorption=0.1, diameter=1.0,
              power=10.0):
    segs = read_scan_path(golden("scan_path.txt"), "segment")
    b = Beam(type=kind, absorption_efficiency=absorption, max_power=power,
             radius=diameter / 2.0, depth=depth, segments=segs)
    p = Problem(beams=[b])
    d, keep = p.to_c()
    return d.beams[0], keep


def test_segment_file(golden):
    # tests/test_scan_path.cc:36-57
    segs = read_scan_path(golden("scan_path.txt"), "segment")
    assert len(segs) == 2
    assert close(segs[0][0], 1.0e-6) and segs[0][2][:2] == (0.0, 0.1)
    assert segs[0][1] == 0.0
    assert close(segs[1][0], 1.0e-6 + 0.002 / 0.8)
    assert segs[1][2][:2] == (0.002, 0.1) and segs[1][1] == 1.0


def test_event_series_file(golden):
    # tests/test_scan_path.cc:59-77
    segs = read_scan_path(golden("scan_path_event_series.inp"), "event_series")
    assert len(segs) == 3
    assert [s[0] for s in segs] == [0.1, 1.0, 2.0]
    assert [s[1] for s in segs] == [0.0, 1.0, 0.5]
    assert [s[2][:2] for s in segs] == [(0.0, 0.0), (0.5, 0.0), (0.5, 0.5)]


def test_scan_path_location(golden):
    # tests/test_scan_path.cc:80-103 (tolerance 1e-10)
    bd, keep = beam_desc(SOURCE_GOLDAK, golden)
    import ctypes as C
    out = np.zeros(3)
    seg = C.c_int(0)
    L = orc.lib()
    L.oracle_scan_path_position(C.byref(bd), 1.0e-7, C.byref(seg),
                                out.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.allclose(out, [0.0, 0.1, 0.1], rtol=1e-10, atol=0)
    L.oracle_scan_path_position(C.byref(bd), 0.001001, C.byref(seg),
                                out.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.allclose(out, [8.0e-4, 0.1, 0.1], rtol=1e-10, atol=0)
    L.oracle_scan_path_position(C.byref(bd), 100.0, C.byref(seg),
                                out.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.all(out == np.finfo(np.float64).min)
    assert L.oracle_scan_path_power_modifier(C.byref(bd), 100.0,
                                             C.byref(seg)) == 0.0


@pytest.mark.parametrize("dim", [2, 3])
def test_heat_source_values(golden, dim):
    # tests/test_heat_source.cc:19-93 (2D) and 220-284 (3D), tolerance 1e-12
    g, kg = beam_desc(SOURCE_GOLDAK, golden)
    e, ke = beam_desc(SOURCE_ELECTRON_BEAM, golden)

    def pt(x, z, y=0.1):
        return [x, z] if dim == 2 else [x, y, z]

    p1 = pt(0.0, 0.05) if dim == 2 else [0.0, 0.0, 0.05]
    for b in (g, e):
        assert orc.beam_value(orc.beam_state(b, 1.0e-7), dim, p1) == 0.0
    p2 = pt(10.0, 0.0) if dim == 2 else [10.0, 0.0, 0.0]
    for b in (g, e):
        assert orc.beam_value(orc.beam_state(b, 5.0e-7), dim, p2) == 0.0
    pi_over_3_to_1p5 = math.pow(math.pi / 3.0, 1.5)
    sg = orc.beam_state(g, 0.001001)
    se = orc.beam_state(e, 0.001001)
    p3 = pt(8.0e-4, 0.1)
    assert close(orc.beam_value(sg, dim, p3),
                 2.0 * 0.1 * 10.0 / (0.5 * 0.5 * 0.1 * pi_over_3_to_1p5))
    assert close(orc.beam_value(se, dim, p3),
                 -0.1 * 10. * 1.0 * math.log(0.1) /
                 (math.pi * 0.5 * 0.5 * 0.1))
    p4 = pt(7.0e-4, 0.09)
    exp_g = 2.0 * 0.1 * 10.0 / (0.5 * 0.5 * 0.1 * pi_over_3_to_1p5) * math.exp(
        -3.0 * 1.0e-4 * 1.0e-4 / 0.25 - 3.0 * 0.01 * 0.01 / 0.1 / 0.1)
    assert close(orc.beam_value(sg, dim, p4), exp_g)
    exp_e = (-0.1 * 10. * 1.0 * math.log(0.1) / (math.pi * 0.5 * 0.5 * 0.1) *
             math.exp(math.log(0.1) * 1.0e-4 * 1.0e-4 / 0.25) *
             (-3.0 * 0.01 * 0.01 / 0.1 / 0.1 + 2.0 * 0.01 / 0.1 + 1.0))
    assert close(orc.beam_value(se, dim, p4), exp_e)
    for b in (g, e):
        assert orc.beam_value(orc.beam_state(b, 100.0), dim, p4) == 0.0


def test_goldak_cutoff_is_per_point(golden):
    """One-lane semantics: a point beyond sqrt(5) radii is zero even when
    evaluated next to points inside (the vectorized reference form only
    returns early when every lane is outside, GoldakHeatSource.cc:129-147)."""
    g, kg = beam_desc(SOURCE_GOLDAK, golden)
    s = orc.beam_state(g, 0.001001)
    r = 0.5
    inside = [8.0e-4 + 2.2 * r, 0.1, 0.1]
    outside = [8.0e-4 + 2.3 * r, 0.1, 0.1]
    assert orc.beam_value(s, 3, inside) > 0.0
    assert orc.beam_value(s, 3, outside) == 0.0
