"""Problem description shared by the C-ABI (include/adm.h) and its callers.

This module mirrors what adamantine's constructors gather from the input
database before the thermal operator exists:

* ``Geometry`` box mesh and initial material state
  (source/Geometry.cc:186-260, 318-352),
* ``Boundary`` types per boundary id (source/Boundary.cc:12-109),
* ``MaterialProperty::fill_properties`` tables / polynomials
  (source/MaterialProperty.templates.hh:796-1119),
* heat sources and scan paths (source/ThermalPhysics.templates.hh:376-411,
  source/ScanPath.cc:58-195, source/BeamHeatSourceProperties.hh:55-66),
* the active-cell rule of the ThermalPhysics constructor
  (source/ThermalPhysics.templates.hh:472-493).

The ctypes structures have the exact layout of the C structures in adm.h.
"""
from __future__ import annotations

import ctypes as C
import math
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

MAX_MATERIALS = 4
TABLE_SIZE = 12
N_STATE_PROPS = 8
N_PROPS = 5
MAX_STATES = 3
MAX_POLY = 5
MAX_BEAMS = 4
MAX_BOUNDARY_IDS = 8

# StateProperty / Property order (source/types.hh:40-96)
STATE_PROPERTY_NAMES = [
    "density", "specific_heat", "thermal_conductivity_x",
    "thermal_conductivity_y", "thermal_conductivity_z", "emissivity",
    "radiation_heat_transfer_coef", "convection_heat_transfer_coef"]
PROPERTY_NAMES = ["liquidus", "solidus", "latent_heat",
                  "radiation_temperature_infty", "convection_temperature_infty"]
STATE_NAMES = ["solid", "liquid", "powder"]

BC_ADIABATIC, BC_RADIATIVE, BC_CONVECTIVE = 0x1, 0x2, 0x4
_BC_NAMES = {"adiabatic": BC_ADIABATIC, "radiative": BC_RADIATIVE,
             "convective": BC_CONVECTIVE, "clamped": 0x8, "traction_free": 0x10}

SOURCE_GOLDAK, SOURCE_ELECTRON_BEAM, SOURCE_CUBE = 0, 1, 2
FORWARD_EULER, RK_THIRD_ORDER, RK_FOURTH_ORDER = 0, 1, 2
_METHODS = {"forward_euler": FORWARD_EULER, "rk_third_order": RK_THIRD_ORDER,
            "rk_fourth_order": RK_FOURTH_ORDER}

# source/types.hh:29-34
UNIT_SCALING = {"millimeter": 1e-3, "centimeter": 1e-2, "inch": 2.54e-2,
                "meter": 1., "milliwatt": 1e-3, "watt": 1.,
                "millimeter/second": 1e-3, "centimer/second": 1e-2,
                "meter/second": 1.}


class MaterialDesc(C.Structure):
    _fields_ = [
        ("n_materials", C.c_int32), ("n_states", C.c_int32),
        ("use_table", C.c_int32), ("p_order", C.c_int32),
        ("single_material_layout", C.c_int32), ("pad_", C.c_int32),
        ("table", C.c_double * (MAX_MATERIALS * N_STATE_PROPS * MAX_STATES
                                * TABLE_SIZE * 2)),
        ("poly", C.c_double * (MAX_MATERIALS * N_STATE_PROPS * MAX_STATES
                               * MAX_POLY)),
        ("prop", C.c_double * (MAX_MATERIALS * N_PROPS)),
    ]


class MeshDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("fe_degree", C.c_int32),
        ("quadrature", C.c_int32), ("n_cells", C.c_int32 * 3),
        ("origin", C.c_double * 3), ("extent", C.c_double * 3),
    ]


class Segment(C.Structure):
    _fields_ = [("end_time", C.c_double), ("power_modifier", C.c_double),
                ("end_point", C.c_double * 3),
                ("end_rotation", C.c_double * 4)]


class BeamDesc(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("n_segments", C.c_int32),
        ("absorption_efficiency", C.c_double), ("max_power", C.c_double),
        ("radius", C.c_double), ("depth", C.c_double),
        ("start_time", C.c_double), ("end_time", C.c_double),
        ("value", C.c_double),
        ("min_point", C.c_double * 3), ("max_point", C.c_double * 3),
        ("segments", C.POINTER(Segment)),
    ]


class BeamState(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("on", C.c_int32),
        ("center", C.c_double * 3), ("alpha", C.c_double),
        ("depth", C.c_double), ("inv_depth", C.c_double),
        ("inv_depth_squared", C.c_double), ("radius_squared", C.c_double),
        ("inv_radius_squared", C.c_double), ("cube_value", C.c_double),
        ("cube_min", C.c_double * 3), ("cube_max", C.c_double * 3),
        ("rotated", C.c_int32), ("pad_", C.c_int32),
        ("rotation", C.c_double * 9),
    ]


class ProblemDesc(C.Structure):
    _fields_ = [
        ("mesh", MeshDesc), ("material", MaterialDesc),
        ("n_boundary_ids", C.c_int32),
        ("boundary_type", C.c_int32 * MAX_BOUNDARY_IDS),
        ("n_beams", C.c_int32), ("pad_", C.c_int32),
        ("beams", BeamDesc * MAX_BEAMS),
        ("cell_active", C.POINTER(C.c_uint8)),
        ("cell_material", C.POINTER(C.c_uint8)),
        ("cell_state", C.POINTER(C.c_double)),
        ("deposition_cos", C.POINTER(C.c_double)),
        ("deposition_sin", C.POINTER(C.c_double)),
    ]


# --------------------------------------------------------------------------
# Material model
# --------------------------------------------------------------------------

@dataclass
class Material:
    """Tables/polynomials of one or more materials, already in the padded form
    of MaterialProperty::fill_properties."""
    n_materials: int = 1
    n_states: int = 1
    use_table: bool = False
    p_order: int = 0
    single_material_layout: bool = True
    table: np.ndarray = field(default_factory=lambda: np.zeros(
        (MAX_MATERIALS, N_STATE_PROPS, MAX_STATES, TABLE_SIZE, 2)))
    poly: np.ndarray = field(default_factory=lambda: np.zeros(
        (MAX_MATERIALS, N_STATE_PROPS, MAX_STATES, MAX_POLY)))
    prop: np.ndarray = field(default_factory=lambda: np.full(
        (MAX_MATERIALS, N_PROPS), np.finfo(np.float64).max))

    def set_table(self, material: int, prop: str, state: str,
                  pairs: Sequence[Sequence[float]]):
        """Set a table property (the ``T,v|T,v|...`` input form); the rest of
        the 12 entries repeats the last pair (MaterialProperty.templates.hh:
        975-1004)."""
        if len(pairs) > TABLE_SIZE:
            raise ValueError("Too many coefficients, increase the table size")
        p = STATE_PROPERTY_NAMES.index(prop)
        s = STATE_NAMES.index(state)
        for i in range(TABLE_SIZE):
            t, v = pairs[min(i, len(pairs) - 1)]
            self.table[material, p, s, i, 0] = t
            self.table[material, p, s, i, 1] = v

    def set_poly(self, material: int, prop: str, state: str,
                 coefs: Sequence[float]):
        if len(coefs) > self.p_order + 1:
            raise ValueError(
                "Too many coefficients, increase the polynomial order")
        p = STATE_PROPERTY_NAMES.index(prop)
        s = STATE_NAMES.index(state)
        self.poly[material, p, s, :] = 0.0
        self.poly[material, p, s, :len(coefs)] = coefs

    def set_property(self, material: int, name: str, value: float):
        self.prop[material, PROPERTY_NAMES.index(name)] = value

    def scale_conductivity(self, s: float):
        """Multiply every thermal conductivity (table values and polynomial
        coefficients, all materials and states) by s."""
        for name in ("thermal_conductivity_x", "thermal_conductivity_y",
                     "thermal_conductivity_z"):
            k = STATE_PROPERTY_NAMES.index(name)
            self.table[:, k, :, :, 1] *= s
            self.poly[:, k, :, :] *= s

    def to_c(self) -> MaterialDesc:
        d = MaterialDesc()
        d.n_materials = self.n_materials
        d.n_states = self.n_states
        d.use_table = int(self.use_table)
        d.p_order = self.p_order
        d.single_material_layout = int(self.single_material_layout)
        C.memmove(d.table, np.ascontiguousarray(self.table, np.float64).ctypes.data,
                  self.table.nbytes)
        C.memmove(d.poly, np.ascontiguousarray(self.poly, np.float64).ctypes.data,
                  self.poly.nbytes)
        C.memmove(d.prop, np.ascontiguousarray(self.prop, np.float64).ctypes.data,
                  self.prop.nbytes)
        return d


# --------------------------------------------------------------------------
# Heat sources
# --------------------------------------------------------------------------

@dataclass
class Beam:
    type: int = SOURCE_GOLDAK
    absorption_efficiency: float = 0.0
    max_power: float = 0.0
    radius: float = 0.0          # diameter / 2
    depth: float = 0.0
    # (end_time, pmod, (x,y,z)), plus the unit quaternion (r,i,j,k) on a
    # five-axis path
    segments: List[tuple] = field(default_factory=list)
    # the file ends with SCAN_PATH_END (ScanPath::is_finished)
    scan_path_end: bool = False
    # cube
    start_time: float = 0.0
    end_time: float = 0.0
    value: float = 0.0
    min_point: tuple = (0.0, 0.0, 0.0)
    max_point: tuple = (0.0, 0.0, 0.0)


def segments_from_path(points: Sequence[dict]) -> List[tuple]:
    """Build segments the way ScanPath::load_segment_scan_path does
    (source/ScanPath.cc:58-141). Each entry: {"mode": "point"|"line",
    "xyz": (x,y,z), "pmod": float, "param": dwell time or velocity}."""
    segs: List[tuple] = []
    for k, e in enumerate(points):
        xyz = tuple(float(v) for v in e["xyz"])
        pmod = float(e["pmod"])
        if e["mode"] == "point":
            end_time = (segs[-1][0] if segs else 0.0) + float(e["param"])
        else:
            if not segs:
                raise ValueError(
                    "Error: Scan paths must begin with a 'point' segment.")
            prev = segs[-1][2]
            length = math.sqrt(sum((a - b) ** 2 for a, b in zip(xyz, prev)))
            end_time = segs[-1][0] + abs(length / float(e["param"]))
        segs.append((end_time, pmod, xyz))
    return segs


def read_scan_path(path: str, file_format: str = "segment",
                   distance_scaling: float = 1.0,
                   velocity_scaling: float = 1.0) -> List[tuple]:
    """ScanPath::load_segment_scan_path / load_event_series_scan_path
    (source/ScanPath.cc:58-195)."""
    with open(path) as f:
        lines = f.read().splitlines()
    if file_format == "segment":
        n_seg = int(lines[1].strip())
        entries = []
        for line in lines[3:]:
            if len(entries) >= n_seg:
                break
            if "SCAN_PATH_END" in line:
                break
            tok = line.split()
            if tok[0] == "0":
                mode = "line"
                param = float(tok[5]) * velocity_scaling
            elif tok[0] == "1":
                mode = "point"
                param = float(tok[5])
            else:
                raise ValueError("Mode type in scan path file not recognized.")
            entries.append({"mode": mode, "pmod": float(tok[4]), "param": param,
                            "xyz": tuple(float(tok[i]) * distance_scaling
                                         for i in (1, 2, 3))})
        return segments_from_path(entries)
    if file_format == "event_series":
        # "time x y z [r i j k] power", space or comma separated; a line's
        # power applies to the next segment (source/ScanPath.cc:143-195)
        segs = []
        last_power = 0.0
        for line in lines:
            if line == "":
                continue
            if "SCAN_PATH_END" in line:
                break
            tok = [t for t in re.split(r"[ ,]+", line.strip()) if t]
            xyz = tuple(float(tok[i]) * distance_scaling for i in (1, 2, 3))
            seg = (float(tok[0]), last_power, xyz)
            if len(tok) == 9:
                q = tuple(float(tok[i]) for i in (4, 5, 6, 7))
                if abs(sum(v * v for v in q) - 1.0) >= 1e-10:
                    raise ValueError("The quaternion is not normed.")
                seg = seg + (q,)
            segs.append(seg)
            last_power = float(tok[-1])
        if any(len(s) > 3 for s in segs) and not all(len(s) > 3 for s in segs):
            raise ValueError("Five-axis scan path %s has lines without a "
                             "rotation." % path)
        return segs
    raise ValueError("Error: Format of scan path file not recognized.")


def scan_path_finished(path: str, file_format: str = "segment") -> bool:
    """ScanPath::is_finished (source/ScanPath.cc:58-195): SCAN_PATH_END was
    read. A segment file stops reading after its declared segment count, so
    there the marker counts only if it comes before that many segments."""
    with open(path) as f:
        lines = f.read().splitlines()
    if file_format == "segment":
        n_seg = int(lines[1].strip())
        lines = lines[3:3 + n_seg]
    return any("SCAN_PATH_END" in line for line in lines)


# --------------------------------------------------------------------------
# Problem
# --------------------------------------------------------------------------

@dataclass
class Problem:
    """Structured-box heat problem (what ThermalPhysics::setup sees)."""
    n_cells: tuple = (4, 4, 4)
    extent: tuple = (1.0, 1.0, 1.0)
    origin: tuple = (0.0, 0.0, 0.0)
    fe_degree: int = 1
    quadrature: int = 0          # 0 Gauss, 1 Gauss-Lobatto
    dim: int = 3
    material: Material = field(default_factory=Material)
    boundary_types: List[int] = field(default_factory=lambda: [BC_ADIABATIC] * 6)
    beams: List[Beam] = field(default_factory=list)
    cell_active: Optional[np.ndarray] = None     # uint8 per cell
    cell_material: Optional[np.ndarray] = None   # uint8 per cell
    cell_state: Optional[np.ndarray] = None      # [n_cells, n_states]
    deposition_cos: Optional[np.ndarray] = None
    deposition_sin: Optional[np.ndarray] = None
    method: int = FORWARD_EULER
    initial_temperature: float = 300.0

    # ---- sizes -----------------------------------------------------------
    def axes(self):
        return (0, 1, 2) if self.dim == 3 else (0, 2)

    @property
    def n_cells_total(self) -> int:
        return int(np.prod([self.n_cells[a] for a in self.axes()]))

    @property
    def n_nodes(self) -> tuple:
        return tuple(self.fe_degree * self.n_cells[a] + 1 for a in self.axes())

    @property
    def n_dofs(self) -> int:
        return int(np.prod(self.n_nodes))

    def cell_centers(self) -> np.ndarray:
        """Cell centres, shape [n_cells, dim], cell order x fastest."""
        ax = self.axes()
        h = [self.extent[a] / self.n_cells[a] for a in ax]
        grids = [self.origin[a] + (np.arange(self.n_cells[a]) + 0.5) * hh
                 for a, hh in zip(ax, h)]
        mesh = np.meshgrid(*grids, indexing="ij")
        return np.stack([m.transpose(tuple(range(len(ax)))[::-1]).ravel()
                         for m in mesh], axis=1)

    def node_coordinates(self) -> np.ndarray:
        """Support-point coordinates of every node, shape [n_dofs, dim]."""
        from .basis import gauss_lobatto
        ax = self.axes()
        xs, _ = gauss_lobatto(self.fe_degree + 1)
        coords = []
        for a in ax:
            h = self.extent[a] / self.n_cells[a]
            n = self.fe_degree * self.n_cells[a] + 1
            idx = np.arange(n)
            c = np.minimum(idx // self.fe_degree, self.n_cells[a] - 1)
            loc = idx - c * self.fe_degree
            coords.append(self.origin[a] + (c + xs[loc]) * h)
        mesh = np.meshgrid(*coords, indexing="ij")
        return np.stack([m.transpose(tuple(range(len(ax)))[::-1]).ravel()
                         for m in mesh], axis=1)

    # ---- geometry helpers mirroring the reference ------------------------
    def set_material_height(self, material_height: float):
        """Active iff the cell centre z < material_height
        (source/ThermalPhysics.templates.hh:472-493)."""
        z = self.cell_centers()[:, -1]
        self.cell_active = (z < material_height).astype(np.uint8)

    def set_powder_layer(self, material_height: float, powder_layer: float):
        """Geometry::assign_material_state with use_powder
        (source/Geometry.cc:318-352): powder iff centre z >= height - layer."""
        z = self.cell_centers()[:, -1]
        ns = self.material.n_states
        st = np.zeros((self.n_cells_total, ns))
        powder = z >= material_height - powder_layer
        st[:, 0] = 1.0
        if ns == 3:
            st[powder, 0] = 0.0
            st[powder, 2] = 1.0
        self.cell_state = st

    # ---- C descriptor ----------------------------------------------------
    def to_c(self):
        """Return (ProblemDesc, keepalive) -- keepalive holds the buffers the
        descriptor points to."""
        d = ProblemDesc()
        d.mesh.dim = self.dim
        d.mesh.fe_degree = self.fe_degree
        d.mesh.quadrature = self.quadrature
        for i in range(3):
            d.mesh.n_cells[i] = int(self.n_cells[i])
            d.mesh.origin[i] = float(self.origin[i])
            d.mesh.extent[i] = float(self.extent[i])
        d.material = self.material.to_c()
        if len(self.boundary_types) > MAX_BOUNDARY_IDS:
            raise ValueError("too many boundary ids")
        d.n_boundary_ids = len(self.boundary_types)
        for i, t in enumerate(self.boundary_types):
            d.boundary_type[i] = int(t)
        if len(self.beams) > MAX_BEAMS:
            raise ValueError("too many heat sources")
        d.n_beams = len(self.beams)
        keep = []
        for i, b in enumerate(self.beams):
            bd = d.beams[i]
            bd.type = b.type
            bd.absorption_efficiency = b.absorption_efficiency
            bd.max_power = b.max_power
            bd.radius = b.radius
            bd.depth = b.depth
            bd.start_time = b.start_time
            bd.end_time = b.end_time
            bd.value = b.value
            for k in range(3):
                bd.min_point[k] = b.min_point[k]
                bd.max_point[k] = b.max_point[k]
            segs = (Segment * max(1, len(b.segments)))()
            for k, seg in enumerate(b.segments):
                segs[k].end_time = seg[0]
                segs[k].power_modifier = seg[1]
                for j in range(3):
                    segs[k].end_point[j] = seg[2][j]
                if len(seg) > 3:
                    for j in range(4):
                        segs[k].end_rotation[j] = seg[3][j]
            bd.n_segments = len(b.segments)
            bd.segments = C.cast(segs, C.POINTER(Segment))
            keep.append(segs)
        nc = self.n_cells_total

        def arr(a, dtype, ctype, shape):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dtype)
            if a.size != int(np.prod(shape)):
                raise ValueError("per-cell array has the wrong size")
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(ctype))

        d.cell_active = arr(self.cell_active, np.uint8, C.c_uint8, (nc,))
        d.cell_material = arr(self.cell_material, np.uint8, C.c_uint8, (nc,))
        d.cell_state = arr(self.cell_state, np.float64, C.c_double,
                           (nc, self.material.n_states))
        d.deposition_cos = arr(self.deposition_cos, np.float64, C.c_double, (nc,))
        d.deposition_sin = arr(self.deposition_sin, np.float64, C.c_double, (nc,))
        return d, keep


# --------------------------------------------------------------------------
# Input files (boost property_tree INFO format) -> Problem
# --------------------------------------------------------------------------

def read_info(text_or_path: str) -> Dict:
    """Minimal reader of the boost::property_tree INFO format used by
    adamantine inputs (application/adamantine.cc:312-320). Repeated keys keep
    the first value, as ptree::get does."""
    if "\n" not in text_or_path and not text_or_path.lstrip().startswith("{"):
        with open(text_or_path) as f:
            text = f.read()
    else:
        text = text_or_path
    tokens = []
    for raw in text.splitlines():
        line = raw.split(";")[0].strip()
        if not line:
            continue
        if line == "{" or line == "}":
            tokens.append(line)
            continue
        parts = line.split(None, 1)
        key = parts[0]
        val = parts[1].strip() if len(parts) > 1 else None
        if val is not None and val.startswith('"') and val.endswith('"'):
            val = val[1:-1]
        tokens.append((key, val))
    root: Dict = {}
    stack = [root]
    last_key = None
    for tok in tokens:
        if tok == "{":
            child: Dict = {}
            cur = stack[-1]
            if last_key in cur and isinstance(cur[last_key], dict):
                child = cur[last_key]
            else:
                cur.setdefault(last_key, child)
                if not isinstance(cur[last_key], dict):
                    cur[last_key] = child
            stack.append(child)
        elif tok == "}":
            stack.pop()
        else:
            key, val = tok
            cur = stack[-1]
            if key not in cur:
                cur[key] = val if val is not None else {}
            last_key = key
    return root


def _get(db: Dict, path: str, default=None):
    cur = db
    for k in path.split("."):
        if not isinstance(cur, dict) or k not in cur:
            return default
        cur = cur[k]
    return cur


def material_from_database(db: Dict, n_states: int, p_order: int) -> Material:
    """MaterialProperty::fill_properties (source/MaterialProperty.templates.hh:
    796-1119) for the thermal state properties."""
    m = Material()
    m.use_table = _get(db, "property_format") == "table"
    m.n_materials = int(_get(db, "n_materials"))
    m.n_states = n_states
    m.p_order = p_order
    m.single_material_layout = m.n_materials == 1
    for mid in range(m.n_materials):
        mdb = _get(db, "material_%d" % mid)
        for s, sname in enumerate(STATE_NAMES[:n_states]):
            sdb = mdb.get(sname)
            if not isinstance(sdb, dict):
                continue
            for pname in STATE_PROPERTY_NAMES:
                if pname not in sdb:
                    continue
                txt = re.sub(r"\s+", "", str(sdb[pname]))
                if m.use_table:
                    pairs = [tuple(float(x) for x in e.split(","))
                             for e in txt.split("|")]
                    m.set_table(mid, pname, sname, pairs)
                else:
                    m.set_poly(mid, pname, sname,
                               [float(x) for x in txt.split(",")])
        for k, pname in enumerate(PROPERTY_NAMES):
            if pname in mdb:
                m.prop[mid, k] = float(mdb[pname])
    return m


def parse_boundary_line(s: str) -> int:
    """Boundary::parse_boundary_line (source/Boundary.cc:87-109)."""
    t = 0
    for part in s.split(","):
        t |= _BC_NAMES[part.strip()]
    return t


def boundary_from_database(db: Dict, n_ids: int) -> List[int]:
    """Boundary::Boundary (source/Boundary.cc:12-70)."""
    types = [0] * n_ids
    default = _get(db, "type")
    if default is not None:
        types = [parse_boundary_line(default)] * n_ids
    for i in range(n_ids):
        s = _get(db, "boundary_%d.type" % i)
        if s is not None:
            types[i] = parse_boundary_line(s)
    s = _get(db, "printed_surface.type")
    if s is not None:
        types[-1] = parse_boundary_line(s)
    return types


def beams_from_database(db: Dict, base_dir: str = ".",
                        units: Optional[Dict] = None) -> List[Beam]:
    """Heat sources of the ThermalPhysics constructor
    (source/ThermalPhysics.templates.hh:376-411)."""
    import os
    dim_s = UNIT_SCALING[_get(units or {}, "heat_source.dimension", "meter")]
    pow_s = UNIT_SCALING[_get(units or {}, "heat_source.power", "watt")]
    sp_s = UNIT_SCALING[_get(units or {}, "heat_source.scan_path", "meter")]
    vel_s = UNIT_SCALING[_get(units or {}, "heat_source.velocity",
                              "meter/second")]
    beams = []
    for i in range(int(_get(db, "n_beams"))):
        b = _get(db, "beam_%d" % i)
        t = b["type"]
        if t == "cube":
            beams.append(Beam(
                type=SOURCE_CUBE, start_time=float(b["start_time"]),
                end_time=float(b["end_time"]), value=float(b["value"]) * pow_s,
                min_point=tuple(float(b.get("min_" + a, 0.0)) * dim_s
                                for a in "xyz"),
                max_point=tuple(float(b.get("max_" + a, 0.0)) * dim_s
                                for a in "xyz")))
            continue
        kind = {"goldak": SOURCE_GOLDAK,
                "electron_beam": SOURCE_ELECTRON_BEAM}.get(t)
        if kind is None:
            raise ValueError("Beam type '%s' not recognized." % t)
        path = os.path.join(base_dir, b["scan_path_file"])
        fmt = b["scan_path_file_format"]
        segs = read_scan_path(path, fmt, sp_s, vel_s)
        beams.append(Beam(
            type=kind, absorption_efficiency=float(b["absorption_efficiency"]),
            max_power=float(b["max_power"]) * pow_s,
            radius=float(b["diameter"]) * dim_s / 2.0,
            depth=float(b["depth"]) * dim_s, segments=segs,
            scan_path_end=scan_path_finished(path, fmt)))
    return beams
