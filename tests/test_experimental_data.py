"""Host readers of experimental data against the reference's own known
answers (tests/test_experimental_data.cc)."""
import os

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
