"""Ensemble of thermal simulations stepped together on one GPU.

The reference runs an ensemble as one ``ThermalPhysics`` per member, built
from a per-member copy of the input database whose ``ensemble.<key>_stddev``
keys were drawn from a normal distribution
(source/ensemble_management.cc:166-265), and steps the members one after the
other (application/adamantine.hh:2137-2142). Here the members share one
mesh and FE space and are stepped by single kernel launches over
(cells, members): member j's nodes are entries [j*n_dofs, (j+1)*n_dofs) of
every ensemble vector.

What a member may change (``Member``): each beam's absorption efficiency
and max power, a factor on every thermal conductivity, the active cells and
the initial cell state.
"""
from __future__ import annotations

import copy
import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .problem import (MAX_BEAMS, STATE_PROPERTY_NAMES, Problem, ProblemDesc)
from .thermal import (LAUNCH_CONFIGS, LAUNCH_LINE_COLORED, N_LAUNCH_CONFIGS,
                      _dev_ptr, _stream_ptr, _vp, check, load_library)

_dp = C.POINTER(C.c_double)


class MemberDesc(C.Structure):
    _fields_ = [
        ("absorption_efficiency", C.c_double * MAX_BEAMS),
        ("max_power", C.c_double * MAX_BEAMS),
        ("conductivity_scale", C.c_double),
        ("cell_active", C.POINTER(C.c_uint8)),
        ("cell_state", C.POINTER(C.c_double)),
    ]


@dataclass
class Member:
    """One member's departures from the base problem (None: unchanged)."""
    absorption_efficiency: Optional[Sequence[float]] = None
    max_power: Optional[Sequence[float]] = None
    conductivity_scale: float = 1.0
    cell_active: Optional[np.ndarray] = None
    cell_state: Optional[np.ndarray] = None


def _bind(L):
    if getattr(L, "_ens_bound", False):
        return L
    L.adm_normal_random_vector.argtypes = [C.c_uint32, C.c_uint32, C.c_double,
                                           C.c_double, _dp]
    L.adm_ensemble_create.argtypes = [C.POINTER(ProblemDesc), C.c_int32,
                                      C.POINTER(MemberDesc), C.POINTER(_vp)]
    L.adm_ensemble_destroy.argtypes = [_vp]
    L.adm_ensemble_n_dofs.argtypes = [_vp, C.POINTER(C.c_int64)]
    L.adm_ensemble_compute_inverse_mass.argtypes = [_vp, _vp, _vp]
    L.adm_ensemble_set_time.argtypes = [_vp, C.c_double]
    L.adm_ensemble_vmult.argtypes = [_vp, _vp, _vp, C.c_int, _vp]
    L.adm_ensemble_set_launch_config.argtypes = [_vp, C.c_int32, C.c_int32]
    L.adm_ensemble_autotune.argtypes = [_vp, _vp, _vp, C.c_int32, _vp,
                                        C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int32), _dp]
    L.adm_ensemble_evolve.argtypes = [_vp, C.c_int, C.c_double, C.c_double,
                                      C.c_int, _vp, _dp, _vp]
    L.adm_ensemble_evolve_graph.argtypes = [_vp, C.c_int, C.c_double,
                                            C.c_double, C.c_int, _vp, _dp,
                                            _vp]
    L.adm_ensemble_get_qp_state.argtypes = [_vp, C.c_int32, _dp, _dp]
    L.adm_ensemble_set_qp_state.argtypes = [_vp, C.c_int32, _dp, _dp]
    L.adm_ensemble_get_face_state.argtypes = [_vp, C.c_int32, _dp]
    L.adm_ensemble_set_face_state.argtypes = [_vp, C.c_int32, _dp]
    _ip = C.POINTER(C.c_int64)
    L.adm_da_create.argtypes = [C.POINTER(_vp)]
    L.adm_da_destroy.argtypes = [_vp]
    L.adm_da_fill_noise.argtypes = [_vp, C.c_int32, C.c_int32, _dp, _dp]
    L.adm_ensemble_nearest_dofs.argtypes = [_vp, C.c_int32, C.c_int64, _dp,
                                            _ip]
    L.adm_ensemble_analysis.argtypes = [_vp, _vp, C.c_int32, _dp, C.c_int32,
                                        _ip, _dp, _dp, _dp, C.c_int32,
                                        C.c_double, _vp]
    L.adm_ensemble_transform.argtypes = [_vp, _vp, _dp, _vp]
    _ip32 = C.POINTER(C.c_int32)
    L.adm_ensemble_cells_to_activate.argtypes = [_vp, C.c_int32, C.c_int32,
                                                 _dp, _ip32]
    L.adm_ensemble_activate.argtypes = [_vp, C.c_int32, C.c_int32, _dp, _dp,
                                        _dp, C.c_double, _vp, _ip32, _vp]
    L.adm_ensemble_add_material.argtypes = [_vp, C.c_int32, _ip32, _dp, _dp,
                                            C.c_double, _vp, _vp]
    L.adm_ensemble_set_beam_gain.argtypes = [_vp, C.c_int32, C.c_int32,
                                             C.c_double, C.c_double]
    L._ens_bound = True
    return L


def _lib():
    return _bind(load_library())


# ---------------------------------------------------------------------------
# draws (source/ensemble_management.cc)
# ---------------------------------------------------------------------------

def normal_random_vector(length: int, n_rejected_draws: int, mean: float,
                         stddev: float) -> np.ndarray:
    """get_normal_random_vector (source/ensemble_management.cc:18-52):
    default-seeded std::mt19937, n_rejected_draws draws discarded, negative
    draws redrawn. Runs on the host (no GPU needed)."""
    out = np.zeros(int(length))
    check(_lib().adm_normal_random_vector(
        int(length), int(n_rejected_draws), float(mean), float(stddev),
        out.ctypes.data_as(_dp)))
    return out


def draw_member_values(keys: Sequence[Tuple[str, float, float]],
                       global_ensemble_size: int, first_local_member: int = 0,
                       local_ensemble_size: Optional[int] = None
                       ) -> Dict[str, np.ndarray]:
    """create_database_ensemble (source/ensemble_management.cc:166-265) for
    the (key, mean, stddev) triples in input order: the k-th key discards
    first_local_member + k * global_ensemble_size draws."""
    n_local = (global_ensemble_size if local_ensemble_size is None
               else local_ensemble_size)
    out = {}
    n_rejected = first_local_member
    for key, mean, stddev in keys:
        out[key] = normal_random_vector(n_local, n_rejected, mean, stddev)
        n_rejected += global_ensemble_size
    return out


# ---------------------------------------------------------------------------
# one member as a stand-alone problem
# ---------------------------------------------------------------------------

def member_problem(base: Problem, m: Member) -> Problem:
    """The stand-alone Problem of one member: what the reference builds from
    the member's database. Used to run a member alone (the serial member loop
    of run_ensemble) and by the tests."""
    p = copy.deepcopy(base)
    for b, beam in enumerate(p.beams):
        if m.absorption_efficiency is not None:
            beam.absorption_efficiency = float(m.absorption_efficiency[b])
        if m.max_power is not None:
            beam.max_power = float(m.max_power[b])
    if m.conductivity_scale != 1.0:
        p.material.scale_conductivity(float(m.conductivity_scale))
    if m.cell_active is not None:
        p.cell_active = np.asarray(m.cell_active, dtype=np.uint8).copy()
    if m.cell_state is not None:
        p.cell_state = np.asarray(m.cell_state, dtype=np.float64).copy()
    return p


# ---------------------------------------------------------------------------
# device ensemble
# ---------------------------------------------------------------------------

class EnsembleDevice:
    """All members' ThermalOperator + ThermalPhysics (explicit steppers) in
    one object; every call acts on every member."""

    def __init__(self, base: Problem, members: Sequence[Member],
                 method: Optional[int] = None):
        L = _lib()
        self.base = base
        self.members = list(members)
        self.method = base.method if method is None else method
        M = len(self.members)
        self._desc, keep = base.to_c()
        md = (MemberDesc * max(1, M))()
        nc = base.n_cells_total
        ns = base.material.n_states
        for j, m in enumerate(self.members):
            d = md[j]
            for b in range(MAX_BEAMS):
                if b < len(base.beams):
                    eta = (base.beams[b].absorption_efficiency
                           if m.absorption_efficiency is None
                           else m.absorption_efficiency[b])
                    pw = (base.beams[b].max_power if m.max_power is None
                          else m.max_power[b])
                    d.absorption_efficiency[b] = float(eta)
                    d.max_power[b] = float(pw)
            d.conductivity_scale = float(m.conductivity_scale)
            if m.cell_active is not None:
                a = np.ascontiguousarray(m.cell_active, dtype=np.uint8)
                if a.size != nc:
                    raise ValueError("cell_active has the wrong size")
                keep.append(a)
                d.cell_active = a.ctypes.data_as(C.POINTER(C.c_uint8))
            if m.cell_state is not None:
                a = np.ascontiguousarray(m.cell_state, dtype=np.float64)
                if a.size != nc * ns:
                    raise ValueError("cell_state has the wrong size")
                keep.append(a)
                d.cell_state = a.ctypes.data_as(_dp)
        h = _vp()
        check(L.adm_ensemble_create(C.byref(self._desc), M, md, C.byref(h)))
        self._h = h
        self._keep = keep
        n = C.c_int64()
        check(L.adm_ensemble_n_dofs(h, C.byref(n)))
        self.n_members = M
        self.n_dofs = n.value
        self.n_cells = nc
        self.n_q = (base.fe_degree + 1) ** base.dim
        self.n_face_q = (base.fe_degree + 1) ** (base.dim - 1)
        self._inverse_mass = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib().adm_ensemble_destroy(h)
            self._h = None

    def initialize_dof_vector(self, value: float = 0.0, device=None):
        """[n_members, n_dofs] float64 tensor (member-major storage)."""
        import torch
        return torch.full((self.n_members, self.n_dofs), float(value),
                          dtype=torch.float64, device=device or "cuda")

    def compute_inverse_mass_matrix(self, stream=None):
        self._inverse_mass = self.initialize_dof_vector()
        check(_lib().adm_ensemble_compute_inverse_mass(
            self._h, _dev_ptr(self._inverse_mass), _stream_ptr(stream)))

    def get_inverse_mass_matrix(self):
        if self._inverse_mass is None:
            self.compute_inverse_mass_matrix()
        return self._inverse_mass

    def set_time(self, t: float):
        check(_lib().adm_ensemble_set_time(self._h, float(t)))

    def _check_vec(self, v):
        if tuple(v.shape) != (self.n_members, self.n_dofs) and \
                v.numel() != self.n_members * self.n_dofs:
            raise ValueError("ensemble vectors hold n_members * n_dofs values")

    def vmult(self, dst, src, stream=None):
        self._check_vec(dst)
        self._check_vec(src)
        check(_lib().adm_ensemble_vmult(self._h, _dev_ptr(src), _dev_ptr(dst),
                                        0, _stream_ptr(stream)))

    def vmult_add(self, dst, src, stream=None):
        self._check_vec(dst)
        self._check_vec(src)
        check(_lib().adm_ensemble_vmult(self._h, _dev_ptr(src), _dev_ptr(dst),
                                        1, _stream_ptr(stream)))

    def set_launch_config(self, mode: int, waves_per_block: int):
        """mode: LAUNCH_LINE_ATOMIC (0), LAUNCH_LINE_COLORED (1) or
        LAUNCH_PLANE (2, FE degrees 1 and 2); 2..4 wavefronts per block."""
        check(_lib().adm_ensemble_set_launch_config(
            self._h, int(mode), int(waves_per_block)))

    def autotune(self, src, reps: int = 5, stream=None) -> dict:
        """ThermalOperatorDevice.autotune for the batched launches."""
        self._check_vec(src)
        dst = self.initialize_dof_vector(0.0)
        col, w = C.c_int32(), C.c_int32()
        ms = np.zeros(N_LAUNCH_CONFIGS)
        check(_lib().adm_ensemble_autotune(
            self._h, _dev_ptr(src), _dev_ptr(dst), int(reps),
            _stream_ptr(stream), C.byref(col), C.byref(w),
            ms.ctypes.data_as(_dp)))
        configs = LAUNCH_CONFIGS
        return {"mode": int(col.value),
                "use_coloring": col.value == LAUNCH_LINE_COLORED,
                "waves_per_block": int(w.value),
                "ms": dict(zip(configs, ms.tolist()))}

    def evolve(self, t: float, delta_t: float, n_steps: int, solution,
               use_graph: bool = False, stream=None) -> float:
        """n_steps batched steps of every member; with use_graph, as
        replays of one captured step (adm_ensemble_evolve_graph)."""
        self._check_vec(solution)
        t_end = C.c_double()
        fn = (_lib().adm_ensemble_evolve_graph if use_graph
              else _lib().adm_ensemble_evolve)
        check(fn(self._h, int(self.method), float(t), float(delta_t),
                 int(n_steps), _dev_ptr(solution), C.byref(t_end),
                 _stream_ptr(stream)))
        return t_end.value

    def evolve_one_time_step(self, t, delta_t, solution, stream=None):
        return self.evolve(t, delta_t, 1, solution, stream=stream)

    def elements_to_activate(self, member: int, boxes) -> np.ndarray:
        """Per cell of `member`: the last deposition box meeting an inactive
        cell, else -1 (source/material_deposition.cc:376-435)."""
        b = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 6)
        out = np.zeros(self.n_cells, dtype=np.int32)
        check(_lib().adm_ensemble_cells_to_activate(
            self._h, int(member), b.shape[0], b.ctypes.data_as(_dp),
            out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def add_material(self, member: int, box_of_cell, box_cos, box_sin,
                     new_material_temperature: float, solution, stream=None):
        """Activation for one member (add_material_start/end,
        source/ThermalPhysics.templates.hh:595-825); `solution` is the whole
        ensemble vector."""
        bc = np.ascontiguousarray(box_of_cell, dtype=np.int32)
        c = np.ascontiguousarray(box_cos, dtype=np.float64)
        s_ = np.ascontiguousarray(box_sin, dtype=np.float64)
        if solution is not None:
            self._check_vec(solution)
        check(_lib().adm_ensemble_add_material(
            self._h, int(member), bc.ctypes.data_as(C.POINTER(C.c_int32)),
            c.ctypes.data_as(_dp), s_.ctypes.data_as(_dp),
            float(new_material_temperature),
            None if solution is None else _dev_ptr(solution),
            _stream_ptr(stream)))
        self._inverse_mass = None

    def activate(self, member: int, boxes, box_cos, box_sin,
                 new_material_temperature: float, solution,
                 stream=None) -> np.ndarray:
        """elements_to_activate + add_material for one member in one call,
        on the device; returns the box that activated each cell (or -1)."""
        b = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 6)
        c = np.ascontiguousarray(box_cos, dtype=np.float64)
        s_ = np.ascontiguousarray(box_sin, dtype=np.float64)
        if c.size != b.shape[0] or s_.size != b.shape[0]:
            raise ValueError("one cos and one sin per box")
        if solution is not None:
            self._check_vec(solution)
        out = np.zeros(self.n_cells, dtype=np.int32)
        check(_lib().adm_ensemble_activate(
            self._h, int(member), b.shape[0], b.ctypes.data_as(_dp),
            c.ctypes.data_as(_dp), s_.ctypes.data_as(_dp),
            float(new_material_temperature),
            None if solution is None else _dev_ptr(solution),
            out.ctypes.data_as(C.POINTER(C.c_int32)), _stream_ptr(stream)))
        self._inverse_mass = None
        return out

    def set_beam_parameters(self, member: int, beam: int,
                            absorption_efficiency: float, max_power: float):
        """update_physics_parameters (source/ThermalPhysics.templates.hh:
        827-852) of one member: absorption efficiency and max power of a
        beam (radius, depth and path are shared by the members)."""
        check(_lib().adm_ensemble_set_beam_gain(
            self._h, int(member), int(beam), float(absorption_efficiency),
            float(max_power)))

    def qp_state(self, member: int):
        liq = np.zeros((self.n_cells, self.n_q))
        pw = np.zeros((self.n_cells, self.n_q))
        check(_lib().adm_ensemble_get_qp_state(
            self._h, int(member), liq.ctypes.data_as(_dp),
            pw.ctypes.data_as(_dp)))
        return liq, pw

    def set_qp_state(self, member: int, liquid, powder):
        lq = np.ascontiguousarray(liquid, dtype=np.float64)
        pw = np.ascontiguousarray(powder, dtype=np.float64)
        if lq.size != self.n_cells * self.n_q or pw.size != lq.size:
            raise ValueError("quadrature-point state of the wrong size")
        check(_lib().adm_ensemble_set_qp_state(
            self._h, int(member), lq.ctypes.data_as(_dp),
            pw.ctypes.data_as(_dp)))

    def face_state(self, member: int):
        fp = np.zeros((self.n_cells, 2 * self.base.dim, self.n_face_q))
        check(_lib().adm_ensemble_get_face_state(
            self._h, int(member), fp.ctypes.data_as(_dp)))
        return fp

    def set_face_state(self, member: int, face_powder):
        fp = np.ascontiguousarray(face_powder, dtype=np.float64)
        if fp.size != self.n_cells * 2 * self.base.dim * self.n_face_q:
            raise ValueError("face state of the wrong size")
        check(_lib().adm_ensemble_set_face_state(
            self._h, int(member), fp.ctypes.data_as(_dp)))


# ---------------------------------------------------------------------------
# data assimilation (source/DataAssimilator.cc)
# ---------------------------------------------------------------------------

LOCALIZATION = {"none": 0, "step_function": 1, "gaspari_cohn": 2}


class DataAssimilator:
    """Counterpart of adamantine's DataAssimilator (source/DataAssimilator.hh)
    for a device ensemble: the observation-noise stream, the observation
    operator and update_ensemble. ``localization_cutoff_function`` and
    ``localization_cutoff_distance`` are the reference's input keys."""

    def __init__(self, localization_cutoff_function: str = "none",
                 localization_cutoff_distance: float = np.finfo(float).max):
        key = localization_cutoff_function.lower()
        if key not in LOCALIZATION:
            raise ValueError("Unknown localization cutoff function. Valid "
                             "options are 'gaspari_cohn', 'step_function', "
                             "and 'none'.")
        self.localization = LOCALIZATION[key]
        self.cutoff = float(localization_cutoff_distance)
        h = _vp()
        check(_lib().adm_da_create(C.byref(h)))
        self._h = h
        self.expt_to_dof = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib().adm_da_destroy(h)
            self._h = None

    def fill_noise_vectors(self, count: int, R) -> np.ndarray:
        """fill_noise_vector (:630-670) for `count` members in order."""
        R = np.ascontiguousarray(R, dtype=np.float64)
        m = R.shape[0]
        out = np.zeros((int(count), m))
        check(_lib().adm_da_fill_noise(self._h, int(count), m,
                                       R.ctypes.data_as(_dp),
                                       out.ctypes.data_as(_dp)))
        return out

    def update_dof_mapping(self, ensemble: "EnsembleDevice", points,
                           member: int = 0) -> np.ndarray:
        """get_expt_to_dof_mapping (source/experimental_data_utils.cc:
        70-110) + update_dof_mapping: nearest active DoF of every point."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(pts.shape[0], dtype=np.int64)
        check(_lib().adm_ensemble_nearest_dofs(
            ensemble._h, int(member), pts.shape[0], pts.ctypes.data_as(_dp),
            out.ctypes.data_as(C.POINTER(C.c_int64))))
        self.expt_to_dof = out
        return out

    def update_ensemble(self, ensemble: "EnsembleDevice", solution, expt_data,
                        R, params=None, noise=True, stream=None):
        """update_ensemble (:95-189) on the device ensemble `solution`
        ([n_members, n_dofs] tensor, in place) and the host parameter block
        `params` ([n_members, n_params] float64 array, in place). noise:
        True draws the perturbations from this object's stream, an array
        gives them, False/None uses none."""
        if self.expt_to_dof is None:
            raise RuntimeError("update_dof_mapping first")
        obs = self.expt_to_dof
        y = np.ascontiguousarray(expt_data, dtype=np.float64)
        R = np.ascontiguousarray(R, dtype=np.float64)
        m = obs.size
        if y.size != m or R.shape != (m, m):
            raise ValueError("Unexpected experiment vector size.")
        M = ensemble.n_members
        if noise is True:
            nz = self.fill_noise_vectors(M, R)
        elif noise is None or noise is False:
            nz = None
        else:
            nz = np.ascontiguousarray(noise, dtype=np.float64)
        npar = 0
        pp = None
        if params is not None:
            if not (isinstance(params, np.ndarray) and params.dtype == np.float64
                    and params.flags.c_contiguous and params.shape[0] == M):
                raise ValueError("params: C-contiguous float64 [n_members, k]")
            npar = params.shape[1]
            pp = params.ctypes.data_as(_dp)
        ensemble._check_vec(solution)
        check(_lib().adm_ensemble_analysis(
            ensemble._h, _dev_ptr(solution), npar, pp, m,
            obs.ctypes.data_as(C.POINTER(C.c_int64)), y.ctypes.data_as(_dp),
            R.ctypes.data_as(_dp), None if nz is None else nz.ctypes.data_as(_dp),
            self.localization, self.cutoff, _stream_ptr(stream)))


def ensemble_transform(ensemble: EnsembleDevice, solution, Tm, stream=None):
    """solution <- solution^T-combination: member j becomes
    sum_l Tm[l, j] * member l (FP64 MFMA kernel)."""
    Tm = np.ascontiguousarray(Tm, dtype=np.float64)
    if Tm.shape != (ensemble.n_members, ensemble.n_members):
        raise ValueError("Tm must be n_members x n_members")
    ensemble._check_vec(solution)
    check(_lib().adm_ensemble_transform(ensemble._h, _dev_ptr(solution),
                                        Tm.ctypes.data_as(_dp),
                                        _stream_ptr(stream)))
