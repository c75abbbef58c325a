"""Experimental data on the device path: point-cloud frames, frame time
stamps, the observation operator and the write of observed values into a
temperature vector.

Host readers follow the reference's file formats:
  PointCloud             source/PointCloud.cc:14-105
  RayTracing             source/RayTracing.cc:111-416
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


class RayTracing:
    """RayTracing (source/RayTracing.cc:111-416) on the structured box.
    Each line of a frame file is "x0,y0,z0,x1,y1,z1,value": the ray starts
    at the first point and goes through the second. A ray's observed point
    is its nearest intersection (distance d >= 0 along the direction) with a
    face of an active cell; rays that hit nothing are dropped.

    The reference finds candidate cells with a bounding-volume search and
    then tests every face of each candidate: the ray-plane distance by
    Cramer's rule and the point in either of the face's two triangles by
    barycentric coordinates. On the box the candidate faces are those of
    the grid planes the ray crosses, taken in the order the ray meets them;
    each is tested with the reference's own face arithmetic, so the points
    are those the reference computes. `active` returns the current active
    flags per cell (cell order x fastest)."""

    def __init__(self, file: str, first_camera_id: int, last_camera_id: int,
                 problem, active, first_frame: int = 0,
                 timeout: Optional[float] = None):
        self._data_filename = file
        self._next_frame = first_frame
        self._first_camera_id = first_camera_id
        self._last_camera_id = last_camera_id
        self._problem = problem
        self._active = active
        self._timeout = timeout
        self._origins = np.zeros((0, 3))
        self._directions = np.zeros((0, 3))
        self._values = np.zeros(0)

    def read_next_frame(self) -> int:
        """read_next_frame (source/RayTracing.cc:126-187)"""
        o, d, v = [], [], []
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
                point, direction, value = [0.0] * 3, [0.0] * 3, 0.0
                i = 0
                for fld in line.split(","):
                    if fld == "":
                        continue
                    x = _strtod_prefix(fld)
                    if i < 3:
                        point[i] = x
                    elif i < 6:
                        direction[i - 3] = x - point[i - 3]
                    else:
                        value = x
                    i += 1
                o.append(point)
                d.append(direction)
                v.append(value)
        self._origins = np.array(o, dtype=float).reshape(-1, 3)
        self._directions = np.array(d, dtype=float).reshape(-1, 3)
        self._values = np.array(v, dtype=float)
        frame = self._next_frame
        self._next_frame += 1
        return frame

    @staticmethod
    def _face_hit(o, r, p0, p1, p2, p3):
        """The reference's face test (source/RayTracing.cc:290-356) for rays
        o + d r against the faces (p0, p1, p2, p3), elementwise: returns
        (d, hit)."""
        def dot(a, b):
            return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + \
                a[..., 2] * b[..., 2]

        e01, e02 = p1 - p0, p2 - p0
        m0 = -r
        # determinant of the rows (-r, e01, e02)
        c0 = e01[..., 1] * e02[..., 2] - e01[..., 2] * e02[..., 1]
        c1 = e01[..., 2] * e02[..., 0] - e01[..., 0] * e02[..., 2]
        c2 = e01[..., 0] * e02[..., 1] - e01[..., 1] * e02[..., 0]
        det = (m0[..., 0] * c0 + m0[..., 1] * c1) + m0[..., 2] * c2
        area = np.zeros(det.shape)
        for e in range(3):
            for k in range(3):
                area = area + np.abs(e01[..., e] * e02[..., k])
        ok = ~(np.abs(det) < 1e-10 * area)
        cross = np.stack([c0, c1, c2], -1)
        with np.errstate(divide="ignore", invalid="ignore"):
            d = dot(cross, o - p0) / det
        ok &= d >= 0
        x = o + d[..., None] * r

        def in_triangle(a, b, c):
            ab, ac, ap = b - a, c - a, x - a
            ab_ab, ab_ac, ac_ac = dot(ab, ab), dot(ab, ac), dot(ac, ac)
            ap_ab, ap_ac = dot(ap, ab), dot(ap, ac)
            denom = ab_ab * ac_ac - ab_ac * ab_ac
            with np.errstate(divide="ignore", invalid="ignore"):
                v = (ac_ac * ap_ab - ab_ac * ap_ac) / denom
                w = (ab_ab * ap_ac - ab_ac * ap_ab) / denom
            u = 1.0 - v - w
            return (u >= 0.0) & (v >= 0.0) & (w >= 0.0)

        ok &= in_triangle(p0, p1, p2) | in_triangle(p1, p2, p3)
        return d, ok

    def get_points_values(self) -> PointsValues:
        """get_points_values (source/RayTracing.cc:189-416)"""
        p = self._problem
        n = np.array(p.n_cells)
        org = np.array(p.origin, dtype=float)
        h = np.array(p.extent, dtype=float) / n
        act = np.asarray(self._active(), bool).reshape(n[2], n[1], n[0])
        O, R = self._origins, self._directions
        nr = len(O)
        best = np.full(nr, np.inf)
        point = np.zeros((nr, 3))
        # deal.II's face vertex order: the two tangential axes, lower first
        tang = {0: (1, 2), 1: (0, 2), 2: (0, 1)}
        for a in (2, 1, 0):
            b, c = tang[a]
            go = R[:, a] != 0.0
            with np.errstate(divide="ignore", invalid="ignore"):
                # plane index of the first crossing at t >= 0, ray order
                f = (O[:, a] - org[a]) / h[a]
                step = np.where(R[:, a] > 0, 1, -1)
                k = np.where(step > 0, np.ceil(f), np.floor(f))
            k = np.nan_to_num(k, nan=-1.0, posinf=-1.0, neginf=-1.0)
            # a ray that starts before the first plane it meets enters at it
            k = np.where(step > 0, np.maximum(k, 0), np.minimum(k, n[a]))
            go &= (k >= 0) & (k <= n[a])
            idx = np.nonzero(go)[0]
            k, step = k[idx].astype(int), step[idx]
            while idx.size:
                inside = (k >= 0) & (k <= n[a])
                P = org[a] + k * h[a]
                with np.errstate(divide="ignore", invalid="ignore"):
                    t = (P - O[idx, a]) / R[idx, a]
                keep = inside & (t <= best[idx] * (1 + 1e-12) + 1e-300)
                idx, k, step, P, t = (idx[keep], k[keep], step[keep],
                                      P[keep], t[keep])
                if not idx.size:
                    break
                q = O[idx] + t[:, None] * R[idx]
                jb0 = np.floor((q[:, b] - org[b]) / h[b]).astype(int)
                jc0 = np.floor((q[:, c] - org[c]) / h[c]).astype(int)
                for db in (-1, 0, 1):
                    for dc in (-1, 0, 1):
                        jb, jc = jb0 + db, jc0 + dc
                        ok = (jb >= 0) & (jb < n[b]) & (jc >= 0) & (jc < n[c])
                        cell = np.zeros((idx.size, 3), int)
                        cell[:, b], cell[:, c] = jb, jc
                        on = np.zeros(idx.size, bool)
                        for side in (k - 1, k):
                            cs = cell.copy()
                            cs[:, a] = side
                            ins = ok & (side >= 0) & (side < n[a])
                            cz = np.clip(cs, 0, n - 1)
                            on |= ins & act[cz[:, 2], cz[:, 1], cz[:, 0]]
                        if not on.any():
                            continue
                        sel = np.nonzero(on)[0]
                        v = np.zeros((sel.size, 4, 3))
                        for m, (ub, uc) in enumerate(((0, 0), (1, 0), (0, 1),
                                                      (1, 1))):
                            v[:, m, a] = P[sel]
                            v[:, m, b] = org[b] + (jb[sel] + ub) * h[b]
                            v[:, m, c] = org[c] + (jc[sel] + uc) * h[c]
                        d, hit = self._face_hit(O[idx[sel]], R[idx[sel]],
                                                v[:, 0], v[:, 1], v[:, 2],
                                                v[:, 3])
                        rows = idx[sel]
                        better = hit & (d < best[rows])
                        r2 = rows[better]
                        best[r2] = d[better]
                        point[r2] = O[r2] + d[better][:, None] * R[r2]
                k = k + step
        found = np.isfinite(best)
        return PointsValues([tuple(x) for x in point[found]],
                            list(self._values[found]))


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
                                vals.ctypes.data_as(_dp),
                                _dev_ptr(temperature, operator.m()),
                                _stream_ptr(stream)))
