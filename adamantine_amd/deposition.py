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


def deposition_along_scan_path(geometry: Dict, segments: Sequence[tuple],
                               dim: int = 3) -> Deposition:
    """deposition_along_scan_path (source/material_deposition.cc:130-300)
    for a three-axis path: boxes of deposition_length along every powered
    segment (the last one shortened to the segment end), deposition_width
    across it and deposition_height below the path, each deposited
    deposition_lead_time before the beam reaches its centre.
    segments: (end_time, power_modifier, end_point) as read_scan_path gives
    them."""
    length = float(geometry["deposition_length"])
    height = float(geometry["deposition_height"])
    width = float(_get(geometry, "deposition_width", 0.0))
    lead = float(geometry["deposition_lead_time"])
    eps, eps_time = 1.0e-10, 1.0e-10
    boxes, times, coss, sins = [], [], [], []
    start_time = 0.0
    start = np.array(segments[0][2], dtype=float)
    for end_time, pmod, end_point in segments:
        end = np.array(end_point, dtype=float)
        if pmod > eps:
            seg_len = float(np.linalg.norm(end - start))
            center = start.copy()
            velocity = seg_len / (end_time - start_time)
            c = (end[0] - start[0]) / seg_len if seg_len != 0.0 else 1.0
            s = (end[1] - start[1]) / seg_len if seg_len != 0.0 else 0.0
            box_len = length
            while True:
                dist = float(np.linalg.norm(center - start))
                t_center = dist / velocity if velocity != 0.0 else 0.0
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
                center[0] += c * incr
                center[1] += s * incr
        start = end
        start_time = end_time
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
