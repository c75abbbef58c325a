"""Experimental data on the device path: point-cloud frames, frame time
stamps, the observation operator and the write of observed values into a
temperature vector.

Host readers follow the reference's file formats:
  PointCloud             source/PointCloud.cc:14-105
  read_frame_timestamps  source/experimental_data_utils.cc:139-207
  wait_for_file          source/utils.hh:24-36
The node lookup and the vector write are libadm.so calls
(adm_nearest_dofs, adm_set_values).
"""
from __future__ import annotations

import ctypes as C
import os
import re
import time
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from .thermal import _dev_ptr, _dp, _stream_ptr, check, load_library

_ip64 = C.POINTER(C.c_int64)


def _lib():
    L = load_library()
    if not getattr(L, "_xd_bound", False):
        L.adm_nearest_dofs.argtypes = [C.c_void_p, C.c_int64, _dp, _ip64]
        L.adm_set_values.argtypes = [C.c_void_p, C.c_int64, _ip64, _dp,
                                     C.c_void_p, C.c_void_p]
        L._xd_bound = True
    return L


def wait_for_file(filename: str, message: str, poll: float = 0.1,
                  timeout: Optional[float] = None):
    """Wait until `filename` exists (the reference spins; this polls and
    prints `message` once). A timeout in seconds raises FileNotFoundError."""
    t0 = time.monotonic()
    told = False
    while not os.path.exists(filename):
        if not told:
            print(message, flush=True)
            told = True
        if timeout is not None and time.monotonic() - t0 > timeout:
            raise FileNotFoundError(filename)
        time.sleep(poll)


def _strtod_prefix(s: str) -> float:
    """std::strtod on a field: leading white space, the longest numeric
    prefix, 0.0 when there is none."""
    m = re.match(r"\s*([+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan))",
                 s, re.IGNORECASE)
    return float(m.group(1)) if m else 0.0


@dataclass
class PointsValues:
    points: List[Tuple[float, ...]] = field(default_factory=list)
    values: List[float] = field(default_factory=list)


class PointCloud:
    """PointCloud<dim> (source/PointCloud.cc:14-105). File names replace
    "#camera" by the camera id, then "#frame" by the frame number; each file
    has one header line, then "x,y[,z],value" lines: empty fields are
    skipped, the first `dim` fields are the coordinates and the last one the
    value."""

    def __init__(self, file: str, first_camera_id: int, last_camera_id: int,
                 first_frame: int = 0, dim: int = 3,
                 timeout: Optional[float] = None):
        self._data_filename = file
        self._next_frame = first_frame
        self._first_camera_id = first_camera_id
        self._last_camera_id = last_camera_id
        self._dim = dim
        self._timeout = timeout
        self._current = PointsValues()

    @classmethod
    def from_database(cls, db: dict, dim: int = 3, **kw):
        """experiment.file / first_frame / first_camera_id / last_camera_id"""
        return cls(db["file"], int(db["first_camera_id"]),
                   int(db["last_camera_id"]), int(db.get("first_frame", 0)),
                   dim, **kw)

    def read_next_frame(self) -> int:
        pv = PointsValues()
        for cam in range(self._first_camera_id, self._last_camera_id + 1):
            name = self._data_filename.replace("#camera", str(cam)).replace(
                "#frame", str(self._next_frame))
            wait_for_file(name, "Waiting for the next frame: " + name,
                          timeout=self._timeout)
            with open(name) as f:
                lines = f.read().split("\n")
            if lines and lines[-1] == "":
                lines.pop()
            for line in lines[1:]:
                point = [0.0] * self._dim
                value = 0.0
                i = 0
                for fld in line.split(","):
                    if fld == "":
                        continue
                    if i < self._dim:
                        point[i] = _strtod_prefix(fld)
                    else:
                        value = _strtod_prefix(fld)
                    i += 1
                pv.points.append(tuple(point))
                pv.values.append(value)
        self._current = pv
        frame = self._next_frame
        self._next_frame += 1
        return frame

    def get_points_values(self) -> PointsValues:
        return self._current


def read_frame_timestamps(log_filename: str, first_camera_id: int,
                          last_camera_id: int, last_frame: int,
                          first_frame: int = 0,
                          first_frame_temporal_offset: float = 0.0,
                          timeout: Optional[float] = None
                          ) -> List[List[float]]:
    """read_frame_timestamps (source/experimental_data_utils.cc:139-207):
    "frame, t_camera0, t_camera1, ..." lines with consecutive frame indices;
    for frames in [first_frame, last_frame] the non-empty entries plus the
    offset, one list per camera."""
    wait_for_file(log_filename, "Waiting for frame time stamps: " +
                  log_filename, timeout=timeout)
    n_cam = last_camera_id - first_camera_id + 1
    stamps: List[List[float]] = [[] for _ in range(n_cam)]
    prev = None
    with open(log_filename) as f:
        for line in f.read().split("\n"):
            if line == "" :
                continue
            entries = [e.strip() for e in line.split(",")]
            frame = int(entries[0])
            if prev is not None and frame - prev != 1:
                raise ValueError("The file " + log_filename +
                                 " does not have consecutive frame indices.")
            prev = frame
            if not (first_frame <= frame <= last_frame):
                continue
            for k, e in enumerate(entries[1:]):
                if e:
                    if k >= n_cam:
                        raise ValueError("more time stamps than cameras in " +
                                         log_filename)
                    stamps[k].append(float(e) + first_frame_temporal_offset)
    return stamps


def get_expt_to_dof_mapping(operator, points_values: PointsValues
                            ) -> Tuple[np.ndarray, np.ndarray]:
    """get_expt_to_dof_mapping (source/experimental_data_utils.cc:70-110):
    (point indices, nearest active node of each point)."""
    pts = np.zeros((len(points_values.points), 3))
    for k, p in enumerate(points_values.points):
        if len(p) == 2:   # 2-D points (x, z) of the reference's dim = 2
            pts[k] = (p[0], 0.0, p[1])
        else:
            pts[k] = p
    dofs = np.zeros(len(pts), dtype=np.int64)
    check(_lib().adm_nearest_dofs(operator.handle, len(pts),
                                  pts.ctypes.data_as(_dp),
                                  dofs.ctypes.data_as(_ip64)))
    return np.arange(len(pts), dtype=np.int64), dofs


def set_with_experimental_data(operator, points_values: PointsValues,
                               expt_to_dof_mapping, temperature, stream=None):
    """set_with_experimental_data (source/experimental_data_utils.cc:
    112-137): temperature[dof_i] = value_i in point order."""
    dofs = np.ascontiguousarray(expt_to_dof_mapping[1], dtype=np.int64)
    vals = np.ascontiguousarray(points_values.values, dtype=np.float64)
    if dofs.shape != vals.shape:
        raise ValueError("mapping and values differ in length")
    check(_lib().adm_set_values(operator.handle, dofs.size,
                                dofs.ctypes.data_as(_ip64),
                                vals.ctypes.data_as(_dp), _dev_ptr(temperature),
                                _stream_ptr(stream)))
