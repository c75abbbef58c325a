"""Host readers of experimental data against the reference's own known
answers (tests/test_experimental_data.cc)."""
import os

import numpy as np
import pytest

from adamantine_amd.experimental_data import (PointCloud, _strtod_prefix,
                                              read_frame_timestamps)

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_point_cloud_frame():
    """read_experimental_data_point_cloud_from_file (:22-73): 9 points and
    values of experimental_data_0_0.csv, compared exactly."""
    pc = PointCloud(os.path.join(GOLD, "experimental_data_#camera_#frame.csv"),
                    first_camera_id=0, last_camera_id=0)
    assert pc.read_next_frame() == 0
    pv = pc.get_points_values()
    values_ref = [1.2, 1.3, 1.4, 1.5, 1.6, 1.7, 1.8, 1.9, 2.]
    points_ref = [(0, 0, 1.), (0, 0.5, 1.), (0, 1., 1.), (0.5, 0., 1.),
                  (0.5, 0.5, 1.), (0.5, 1., 1.), (1., 0., 1.), (1., 0.5, 1.),
                  (1., 1., 1.)]
    assert len(pv.points) == 9 and len(pv.values) == 9
    assert pv.values == values_ref
    assert [tuple(p) for p in pv.points] == points_ref


def test_point_cloud_waits_then_times_out(tmp_path):
    pc = PointCloud(str(tmp_path / "missing_#camera_#frame.csv"), 0, 0,
                    timeout=0.2)
    with pytest.raises(FileNotFoundError):
        pc.read_next_frame()


def test_point_cloud_fields(tmp_path):
    """Empty fields are skipped; the last field after the coordinates is the
    value; two cameras are concatenated; frames advance."""
    for cam, txt in ((0, '"x","y","z","T"\n1,,2,3,4\n'),
                     (1, 'h\n 5, 6 ,7,8,9\n')):
        (tmp_path / f"d_{cam}_3.csv").write_text(txt)
    pc = PointCloud(str(tmp_path / "d_#camera_#frame.csv"), 0, 1,
                    first_frame=3)
    assert pc.read_next_frame() == 3
    pv = pc.get_points_values()
    assert pv.points == [(1.0, 2.0, 3.0), (5.0, 6.0, 7.0)]
    assert pv.values == [4.0, 9.0]


def test_strtod_prefix():
    assert _strtod_prefix(" 1.5e-3x") == 1.5e-3
    assert _strtod_prefix("abc") == 0.0
    assert _strtod_prefix("-.5") == -0.5


def test_frame_timestamps():
    """timestamp (:199-223): frames 1..3, two cameras, offset 0.1, 1e-12."""
    ts = read_frame_timestamps(os.path.join(GOLD, "experiment_log_test.txt"),
                               first_camera_id=0, last_camera_id=1,
                               last_frame=3, first_frame=1,
                               first_frame_temporal_offset=0.1)
    assert len(ts) == 2 and len(ts[0]) == 3 and len(ts[1]) == 3
    for got, ref in zip(ts[0], [0.1105, 0.124, 0.145]):
        assert abs(got - ref) < 1e-12
    for got, ref in zip(ts[1], [0.1106, 0.1242, 0.1454]):
        assert abs(got - ref) < 1e-12


def test_frame_timestamps_rejects_gaps(tmp_path):
    f = tmp_path / "log.txt"
    f.write_text("0, 0.1\n2, 0.2\n")
    with pytest.raises(ValueError):
        read_frame_timestamps(str(f), 0, 0, last_frame=2)


def test_ray_tracing_reference_case():
    """read_experimental_data_ray_tracing_from_file
    (tests/test_experimental_data.cc:138-197): five rays against the unit
    cube in 2x2x2 cells; four hit, at exactly these points."""
    from adamantine_amd.experimental_data import RayTracing
    from adamantine_amd.problem import Problem
    p = Problem(n_cells=(2, 2, 2), extent=(1.0, 1.0, 1.0))
    rt = RayTracing(os.path.join(GOLD, "raytracing_experimental_data_"
                                 "#camera_#frame.csv"), 0, 0, p,
                    lambda: np.ones(8))
    assert rt.read_next_frame() == 0
    pv = rt.get_points_values()
    assert pv.values == [1.0, 2.0, 3.0, 5.0]
    assert [tuple(x) for x in pv.points] == [
        (0.0, 0.1, 0.2), (1.0, 0.1, 0.001), (1.0, 1.0, 0.1), (1.0, 0.0, 0.5)]


def _slab_first_hit(o, r, lo, hi):
    """first t >= 0 at which the ray meets the closed box, by the slab
    method (entry if outside, exit if inside), or inf"""
    with np.errstate(divide="ignore", invalid="ignore"):
        t1 = (lo - o) / r
        t2 = (hi - o) / r
    tmin = np.where(r != 0, np.minimum(t1, t2), -np.inf)
    tmax = np.where(r != 0, np.maximum(t1, t2), np.inf)
    outside = (r == 0) & ((o < lo) | (o > hi))
    if outside.any():
        return np.inf
    t_in, t_out = tmin.max(), tmax.min()
    if t_in > t_out or t_out < 0:
        return np.inf
    return t_in if t_in >= 0 else t_out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_ray_tracing_against_slab_search(seed):
    """random rays, from outside and from inside, against a random set of
    active cells: the first hit equals the minimum over every active cell
    of the slab method's first crossing (an independent ray-box test)"""
    from adamantine_amd.experimental_data import RayTracing
    from adamantine_amd.problem import Problem
    rng = np.random.default_rng(seed)
    p = Problem(n_cells=(4, 3, 5), extent=(2.0, 1.5, 1.0),
                origin=(-0.5, 0.25, 0.0))
    active = (rng.random(p.n_cells_total) < 0.4).astype(np.uint8)
    rt = RayTracing("unused", 0, 0, p, lambda: active)
    n = 400
    centre = np.array([0.5, 1.0, 0.5])
    o = centre + rng.normal(scale=1.5, size=(n, 3))
    # aimed at random points of the box, with some rays starting inside
    target = np.array(p.origin) + rng.random((n, 3)) * np.array(p.extent)
    o[: n // 4] = target[n // 2: n // 2 + n // 4]
    r = target - o + 0.2 * rng.normal(size=(n, 3))
    rt._origins, rt._directions = o, r
    rt._values = np.arange(n, dtype=float)
    pv = rt.get_points_values()
    got = {int(v): np.array(x) for x, v in zip(pv.points, pv.values)}
    h = np.array(p.extent) / np.array(p.n_cells)
    hits = 0
    for i in range(n):
        best = np.inf
        for c in np.nonzero(active)[0]:
            idx = np.array([c % 4, (c // 4) % 3, c // 12])
            lo = np.array(p.origin) + idx * h
            best = min(best, _slab_first_hit(o[i], r[i], lo, lo + h))
        if np.isinf(best):
            assert i not in got
        else:
            hits += 1
            np.testing.assert_allclose(got[i], o[i] + best * r[i], rtol=0,
                                       atol=1e-12)
    assert hits > n // 4
