"""Material deposition boxes: from a deposition file or along the scan paths
(source/material_deposition.cc:20-375).

Each result is (boxes [n][2*dim] as min corner then max corner, times [n],
cos [n], sin [n]); these are the deposition boxes, times and directions that
the time loop hands to cells_to_activate / add_material.
"""
from __future__ import annotations

import math
import os
from typing import Dict, List, Sequence, Tuple

import numpy as np

from .problem import _get

Deposition = Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]


def _empty(dim: int) -> Deposition:
    return (np.zeros((0, 2 * dim)), np.zeros(0), np.zeros(0), np.zeros(0))


def read_material_deposition(filename: str, dim: int = 3) -> Deposition:
    """read_material_deposition (source/material_deposition.cc:59-128): a
    first line with the dimension, then per box "center size time angle"
    (space separated); times must not decrease."""
    boxes, times, cos, sin = [], [], [], []
    with open(filename) as f:
        lines = f.read().split("\n")
    if int(lines[0]) != dim:
        raise ValueError("Dimension in %s does not match %d ." % (filename, dim))
    for line in lines[1:]:
        if not line.strip():
            continue
        v = [float(x) for x in line.split()]
        center, size = v[:dim], v[dim:2 * dim]
        t = v[2 * dim]
        if times and times[-1] > t:
            raise ValueError("Time stamp not increasing.")
        times.append(t)
        angle = v[2 * dim + 1]
        cos.append(math.cos(angle))
        sin.append(math.sin(angle))
        boxes.append([center[d] - 0.5 * size[d] for d in range(dim)] +
                     [center[d] + 0.5 * size[d] for d in range(dim)])
    return (np.array(boxes, dtype=float).reshape(-1, 2 * dim),
            np.array(times), np.array(cos), np.array(sin))


def quaternion_matrix(q) -> np.ndarray:
    """Rotation matrix of the unit quaternion (r, i, j, k)
    (Quaternion::build_rotation_matrices, source/Quaternion.cc:120-134)."""
    r, i, j, k = (float(v) for v in q)
    return np.array([
        [1.0 - 2.0 * (j * j + k * k), 2.0 * (i * j - k * r),
         2.0 * (i * k + j * r)],
        [2.0 * (i * j + k * r), 1.0 - 2.0 * (i * i + k * k),
         2.0 * (j * k - i * r)],
        [2.0 * (i * k - j * r), 2.0 * (j * k + i * r),
         1.0 - 2.0 * (i * i + j * j)]])


def _apply(R: np.ndarray, p) -> np.ndarray:
    """R p with the sum over columns in order (Tensor<2,3> * Point<3>)."""
    out = np.empty(3)
    for a in range(3):
        acc = R[a, 0] * p[0]
        acc += R[a, 1] * p[1]
        acc += R[a, 2] * p[2]
        out[a] = acc
    return out


def deposition_along_scan_path(geometry: Dict, segments: Sequence[tuple],
                               dim: int = 3) -> Deposition:
    """deposition_along_scan_path (source/material_deposition.cc:130-300):
    boxes of deposition_length along every powered segment (the last one
    shortened to the segment end), deposition_width across it and
    deposition_height below the path, each deposited deposition_lead_time
    before the beam reaches its centre.

    segments: (end_time, power_modifier, end_point[, end_rotation]) as
    read_scan_path gives them. On a five-axis path (segments carrying a
    quaternion) the box centres walk from the inverse-rotated start point to
    the inverse-rotated end point, and each box is the axis-aligned hull of
    the segment's end rotation applied to the box's two corners around the
    centre, as the reference does."""
    length = float(geometry["deposition_length"])
    height = float(geometry["deposition_height"])
    width = float(_get(geometry, "deposition_width", 0.0))
    lead = float(geometry["deposition_lead_time"])
    eps, eps_time = 1.0e-10, 1.0e-10
    boxes, times, coss, sins = [], [], [], []

    def rot(seg):
        return quaternion_matrix(seg[3]) if len(seg) > 3 else None

    def build_ref(R, point):
        # inv_rotate: the transpose applied to the point
        return _apply(R.T, point) if R is not None else point.copy()

    start_time = 0.0
    start = np.array(segments[0][2], dtype=float)
    ref_start = build_ref(rot(segments[0]), start)
    for seg in segments:
        end_time, pmod = seg[0], seg[1]
        end = np.array(seg[2], dtype=float)
        R = rot(seg)
        if pmod > eps:
            ref_end = build_ref(R, end)
            seg_len = float(np.linalg.norm(end - start))
            center = ref_start.copy()
            velocity = seg_len / (end_time - start_time)
            c = (end[0] - start[0]) / seg_len if seg_len != 0.0 else 1.0
            s = (end[1] - start[1]) / seg_len if seg_len != 0.0 else 0.0
            box_len = length
            while True:
                dist = float(np.linalg.norm(center - ref_start))
                t_center = dist / velocity if velocity != 0.0 else 0.0
                if R is not None:
                    if dim != 3:
                        raise NotImplementedError(
                            "five-axis deposition needs dim 3")
                    xl = abs(c) * box_len + abs(s) * width
                    yl = abs(c) * width + abs(s) * box_len
                    a = _apply(R, [xl / 2.0, yl / 2.0, 0.0])
                    b = _apply(R, [-xl / 2.0, -yl / 2.0, -height])
                    lo = list(center + np.minimum(a, b))
                    hi = list(center + np.maximum(a, b))
                else:
                    if dim == 2:
                        size = [box_len, height]
                    else:
                        size = [abs(c) * box_len + abs(s) * width,
                                abs(c) * width + abs(s) * box_len, height]
                    lo = [center[d] - 0.5 * size[d] for d in range(dim - 1)]
                    hi = [center[d] + 0.5 * size[d] for d in range(dim - 1)]
                    lo.append(center[dim - 1] - size[dim - 1])
                    hi.append(center[dim - 1])
                boxes.append(lo + hi)
                times.append(max(start_time + t_center - lead, eps_time))
                coss.append(c)
                sins.append(s)
                if dist + eps > seg_len:
                    break
                incr = length
                if dist + length > seg_len:
                    incr = seg_len - dist
                    box_len = seg_len - dist
                if R is not None:
                    # deal.II's Point /= scalar multiplies by the reciprocal
                    direction = ref_end - center
                    direction = direction * (1.0 / np.sqrt(np.sum(
                        direction * direction)))
                    center = center + incr * direction
                else:
                    center[0] += c * incr
                    center[1] += s * incr
        start = end
        start_time = end_time
        ref_start = build_ref(R, start)
    if not boxes:
        return _empty(dim)
    return (np.array(boxes, dtype=float), np.array(times), np.array(coss),
            np.array(sins))


def merge_deposition_paths(paths: Sequence[Deposition]) -> Deposition:
    """merge_deposition_paths (source/material_deposition.cc:302-353): the
    concatenated paths in chronological order (stable for equal times)."""
    boxes = np.concatenate([p[0] for p in paths])
    times = np.concatenate([p[1] for p in paths])
    cos = np.concatenate([p[2] for p in paths])
    sin = np.concatenate([p[3] for p in paths])
    perm = np.argsort(times, kind="stable")
    return boxes[perm], times[perm], cos[perm], sin[perm]


def create_material_deposition_boxes(geometry: Dict, beams, base_dir: str = ".",
                                     dim: int = 3) -> Deposition:
    """create_material_deposition_boxes (source/material_deposition.cc:
    20-57): none, the deposition file, or the boxes along every beam's scan
    path, merged in time order."""
    if str(_get(geometry, "material_deposition", "false")) != "true":
        return _empty(dim)
    method = str(geometry["material_deposition_method"])
    if method == "file":
        return read_material_deposition(
            os.path.join(base_dir, str(geometry["material_deposition_file"])),
            dim)
    return merge_deposition_paths(
        [deposition_along_scan_path(geometry, b.segments, dim) for b in beams])
