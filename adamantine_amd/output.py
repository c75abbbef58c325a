"""Thermal output as VTU/PVTU/PVD files (source/PostProcessor.cc:83-227,
source/PostProcessor.hh:218-322).

Like deal.II's DataOut with build_patches(n), every cell is written as its
own patch of n^3 hexahedra (n^2 quadrilaterals in 2-D; n = max(1,
additional_output_refinement)) on points of its own. The temperature is
the FE_Q(p) field evaluated at the patch's equispaced points; cells without FE_Q (inactive) get zero, as
FE_Nothing does. The cell data are the material state ratios "powder",
"liquid" and "solid" and "subdomain" (0 here, one process); TIME is the
field data of VtkFlags(time). Files are
<output_dir><prefix>.<step>.0.vtu, <prefix>.<step>.pvtu and <prefix>.pvd,
named as the reference names them. Arrays are inline base64 binary
(UInt64 byte-count header, uncompressed).
"""
from __future__ import annotations

import base64
import os
from typing import Dict, List, Optional, Tuple

import numpy as np

from .basis import gauss_lobatto, lagrange_matrix
from .problem import Problem

_VTK_HEXAHEDRON = 12
_VTK_QUAD = 9


def _b64(a: np.ndarray) -> str:
    data = np.ascontiguousarray(a).tobytes()
    return base64.b64encode(np.uint64(len(data)).tobytes() + data).decode()


def _array(name: Optional[str], a: np.ndarray, vtype: str, ncomp: int = 1,
           extra: str = "") -> str:
    nm = ' Name="%s"' % name if name else ""
    nc = ' NumberOfComponents="%d"' % ncomp if ncomp != 1 else ""
    return ('<DataArray type="%s"%s%s format="binary"%s>\n%s\n</DataArray>\n'
            % (vtype, nm, nc, extra, _b64(a)))


class PostProcessor:
    """Thermal part of PostProcessor: write_thermal_output and write_pvd."""

    def __init__(self, database: Dict, problem: Problem,
                 ensemble_member_index: int = -1):
        prefix = str(database["filename_prefix"])
        if ensemble_member_index >= 0:
            prefix += "_m%d" % ensemble_member_index
        self.filename_prefix = prefix
        self.output_dir = str(database.get("output_dir", ""))
        self.n_sub = max(1, int(database.get("additional_output_refinement",
                                             0)))
        self.problem = problem
        self.times_filenames: List[Tuple[float, str]] = []
        self._build_patches()

    def _build_patches(self):
        p = self.problem
        dim = p.dim
        ax = p.axes()       # (x, y, z), or (x, height) in 2-D
        n, deg = self.n_sub, p.fe_degree
        nc = [p.n_cells[a] for a in ax]
        h = [p.extent[a] / p.n_cells[a] for a in ax]
        org = [p.origin[a] for a in ax]
        xi = np.arange(n + 1) / n
        # equispaced patch points from the FE_Q(p) Gauss-Lobatto support
        # points, one direction at a time
        self._L = lagrange_matrix(gauss_lobatto(deg + 1)[0], xi)
        k = np.arange(n + 1)
        # patch points in x-fastest order, per cell
        pi = np.stack(np.meshgrid(*([k] * dim), indexing="ij"), -1)[..., ::-1]
        pi = pi.reshape(-1, dim)                        # (n+1)^dim x (i, ...)
        cells = np.arange(p.n_cells_total)
        cc = np.stack([(cells // int(np.prod(nc[:d]))) % nc[d]
                       for d in range(dim)], 1)
        # points have three components; a 2-D point is (x, height, 0), as
        # deal.II writes it
        pts = np.zeros((p.n_cells_total, (n + 1) ** dim, 3))
        for d in range(dim):
            pts[..., d] = (org[d] + (cc[:, None, d] + xi[pi[None, :, d]])
                           * h[d])
        self.points = pts.reshape(-1, 3)
        # hexahedra (quadrilaterals in 2-D) of each patch, VTK corner order
        s = [1, n + 1, (n + 1) ** 2]
        base = []
        if dim == 3:
            for kz in range(n):
                for ky in range(n):
                    for kx in range(n):
                        o = kx * s[0] + ky * s[1] + kz * s[2]
                        base.append([o, o + s[0], o + s[0] + s[1], o + s[1],
                                     o + s[2], o + s[0] + s[2],
                                     o + s[0] + s[1] + s[2], o + s[1] + s[2]])
        else:
            for kz in range(n):
                for kx in range(n):
                    o = kx * s[0] + kz * s[1]
                    base.append([o, o + s[0], o + s[0] + s[1], o + s[1]])
        base = np.array(base, dtype=np.int64)
        npp = (n + 1) ** dim
        self.connectivity = (base[None, :, :] +
                             npp * cells[:, None, None]).reshape(-1)
        self.n_hex = p.n_cells_total * n ** dim
        self.cell_type = _VTK_HEXAHEDRON if dim == 3 else _VTK_QUAD
        self.corners = 2 ** dim
        # node index of every (cell, local node) for the gather
        m = deg + 1
        nn = p.n_nodes
        loc = np.arange(m)
        grid = np.meshgrid(*([loc] * dim), indexing="ij")[::-1]  # x first
        node = np.zeros((p.n_cells_total, m ** dim), np.int64)
        stride = 1
        for d in range(dim):
            node += (cc[:, None, d] * deg + grid[d].ravel()[None, :]) * stride
            stride *= nn[d]
        self._cell_nodes = node

    def patch_temperature(self, temperature, active) -> np.ndarray:
        """The field at the patch points, cell by cell; zero on inactive
        cells."""
        T = temperature.detach().cpu().numpy() if hasattr(
            temperature, "cpu") else np.asarray(temperature, np.float64)
        m = self.problem.fe_degree + 1
        L = self._L
        if self.problem.dim == 3:
            u = T[self._cell_nodes].reshape(-1, m, m, m)    # [c][z][y][x]
            v = np.einsum("ax,czyx->czya", L, u)
            v = np.einsum("by,czya->czba", L, v)
            v = np.einsum("cz,dzba->dcba", L, v)
        else:
            u = T[self._cell_nodes].reshape(-1, m, m)       # [c][z][x]
            v = np.einsum("ax,czx->cza", L, u)
            v = np.einsum("bz,cza->cba", L, v)
        v = v.reshape(v.shape[0], -1)
        v[~np.asarray(active, bool)] = 0.0
        return v.reshape(-1)

    def write_thermal_output(self, time_step: int, time: float, temperature,
                             cell_state: np.ndarray, active) -> str:
        """write_thermal_output (source/PostProcessor.hh:218-234): the
        temperature, the material state and the subdomain, then
        write_pvtu; returns the .pvtu path."""
        ns = cell_state.shape[1]
        n3 = self.n_sub ** self.problem.dim
        solid = np.repeat(cell_state[:, 0], n3)
        liquid = np.repeat(cell_state[:, 1] if ns > 1 else
                           np.zeros(len(cell_state)), n3)
        powder = np.repeat(cell_state[:, 2] if ns > 2 else
                           np.zeros(len(cell_state)), n3)
        Tp = self.patch_temperature(temperature, active)
        name = "%s.%d.0.vtu" % (self.filename_prefix, time_step)
        path = os.path.join(self.output_dir, name) if self.output_dir else name
        offsets = self.corners * np.arange(1, self.n_hex + 1, dtype=np.int64)
        types = np.full(self.n_hex, self.cell_type, np.uint8)
        with open(path, "w") as f:
            f.write('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" '
                    'version="1.0" byte_order="LittleEndian" '
                    'header_type="UInt64">\n<UnstructuredGrid>\n'
                    '<FieldData>\n')
            f.write(_array("TIME", np.array([time], np.float64), "Float64",
                           extra=' NumberOfTuples="1"'))
            f.write('</FieldData>\n<Piece NumberOfPoints="%d" '
                    'NumberOfCells="%d">\n<Points>\n' %
                    (len(self.points), self.n_hex))
            f.write(_array(None, self.points, "Float64", 3))
            f.write('</Points>\n<Cells>\n')
            f.write(_array("connectivity", self.connectivity, "Int64"))
            f.write(_array("offsets", offsets, "Int64"))
            f.write(_array("types", types, "UInt8"))
            f.write('</Cells>\n<PointData Scalars="temperature">\n')
            f.write(_array("temperature", Tp, "Float64"))
            f.write('</PointData>\n<CellData>\n')
            for nm, a in (("powder", powder), ("liquid", liquid),
                          ("solid", solid),
                          ("subdomain", np.zeros(self.n_hex, np.float32))):
                f.write(_array(nm, a, "Float64" if a.dtype == np.float64
                               else "Float32"))
            f.write('</CellData>\n</Piece>\n</UnstructuredGrid>\n'
                    '</VTKFile>\n')
        pvtu = "%s.%d.pvtu" % (self.filename_prefix, time_step)
        pvtu_path = (os.path.join(self.output_dir, pvtu) if self.output_dir
                     else pvtu)
        with open(pvtu_path, "w") as f:
            f.write('<?xml version="1.0"?>\n<VTKFile type="PUnstructuredGrid" '
                    'version="1.0" byte_order="LittleEndian" '
                    'header_type="UInt64">\n<PUnstructuredGrid GhostLevel="0">'
                    '\n<PPointData Scalars="temperature">\n'
                    '<PDataArray type="Float64" Name="temperature"/>\n'
                    '</PPointData>\n<PCellData>\n')
            for nm, t in (("powder", "Float64"), ("liquid", "Float64"),
                          ("solid", "Float64"), ("subdomain", "Float32")):
                f.write('<PDataArray type="%s" Name="%s"/>\n' % (t, nm))
            f.write('</PCellData>\n<PPoints>\n<PDataArray type="Float64" '
                    'NumberOfComponents="3"/>\n</PPoints>\n'
                    '<Piece Source="%s"/>\n</PUnstructuredGrid>\n'
                    '</VTKFile>\n' % name)
        self.times_filenames.append((float(time), pvtu_path))
        return pvtu_path

    def write_pvd(self) -> str:
        """write_pvd (source/PostProcessor.cc:83-91): the time of every
        .pvtu written so far."""
        name = self.filename_prefix + ".pvd"
        path = os.path.join(self.output_dir, name) if self.output_dir else name
        with open(path, "w") as f:
            f.write('<?xml version="1.0"?>\n<VTKFile type="Collection" '
                    'version="0.1" byte_order="LittleEndian">\n'
                    '<Collection>\n')
            for t, fn in self.times_filenames:
                f.write('<DataSet timestep="%r" group="" part="0" '
                        'file="%s"/>\n' % (t, fn))
            f.write('</Collection>\n</VTKFile>\n')
        return path


def read_vtu(path: str) -> Dict[str, np.ndarray]:
    """Arrays of a .vtu written by PostProcessor, by name ("points" for the
    coordinates), decoded."""
    import xml.etree.ElementTree as ET
    types = {"Float64": np.float64, "Float32": np.float32,
             "Int64": np.int64, "UInt8": np.uint8}
    out = {}
    root = ET.parse(path).getroot()
    for da in root.iter("DataArray"):
        raw = base64.b64decode(da.text.strip())
        nbytes = int(np.frombuffer(raw[:8], np.uint64)[0])
        a = np.frombuffer(raw[8:8 + nbytes], types[da.get("type")])
        nc = int(da.get("NumberOfComponents", "1"))
        out[da.get("Name", "points")] = a.reshape(-1, nc) if nc > 1 else a
    return out
