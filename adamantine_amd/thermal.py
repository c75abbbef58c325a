"""Host-side mirror of adamantine's thermal operator / physics API over the
C-ABI of libadm.so (include/adm.h).

Class and method names follow the reference (source/ThermalOperatorBase.hh,
source/ThermalPhysicsInterface.hh) so that code written against
``ThermalOperatorDevice`` / ``ThermalPhysics<..., MemorySpace::Default>``
maps one-to-one. Device vectors are torch float64 tensors on the current HIP
device; torch only provides memory and streams, every operator call runs the
hand-written HIP kernels of libadm.so.

The library must be built (``python -m adamantine_amd.build``); there is no
fallback: a missing or unloadable library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from .problem import BeamState, Problem, ProblemDesc

# cell-kernel launch modes of adm_set_launch_config (include/adm.h)
LAUNCH_LINE_ATOMIC = 0
LAUNCH_LINE_COLORED = 1
LAUNCH_PLANE = 2
# the autotuner's configurations, in the order of its timings
LAUNCH_CONFIGS = [(m, w) for m in (LAUNCH_LINE_COLORED, LAUNCH_LINE_ATOMIC,
                                   LAUNCH_PLANE) for w in (2, 3, 4)]
N_LAUNCH_CONFIGS = len(LAUNCH_CONFIGS)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libadm.so")
_lib = None

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class AdmError(RuntimeError):
    """Error reported by the C-ABI (the reference's ASSERT_THROW)."""


def load_library():
    """Load libadm.so. Raises if it is missing: the HIP path is the product,
    there is no CPU fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AdmError("libadm.so not built: run python -m adamantine_amd.build")
    # torch's HIP runtime must be the process's one: loaded first, it also
    # satisfies libadm.so's libamdhip64.so.7 dependency (same soname);
    # loaded the other way round the process would hold two HIP runtimes
    import torch  # noqa: F401
    L = C.CDLL(LIB_PATH)
    L.adm_last_error.restype = C.c_char_p
    L.adm_create.argtypes = [C.POINTER(ProblemDesc), C.POINTER(_vp)]
    L.adm_destroy.argtypes = [_vp]
    L.adm_n_dofs.argtypes = [_vp, C.POINTER(C.c_int64)]
    L.adm_sizes.argtypes = [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                            C.POINTER(C.c_int32)]
    L.adm_compute_inverse_mass.argtypes = [_vp, _vp, _vp]
    L.adm_set_time.argtypes = [_vp, C.c_double]
    L.adm_vmult.argtypes = [_vp, _vp, _vp, C.c_int, _vp]
    L.adm_evolve.argtypes = [_vp, C.c_int, C.c_double, C.c_double, C.c_int,
                             _vp, C.c_int, _dp, _vp]
    L.adm_set_cell_state.argtypes = [_vp, _dp]
    L.adm_get_cell_state.argtypes = [_vp, _dp]
    L.adm_get_qp_state.argtypes = [_vp, _dp, _dp]
    L.adm_set_qp_state.argtypes = [_vp, _dp, _dp]
    L.adm_vmult_noalloc.argtypes = [_vp, _vp, _vp, C.c_int, _vp]
    L.adm_n_nodes.argtypes = [_vp, C.POINTER(C.c_int64)]
    L.adm_set_dof_map.argtypes = [_vp, C.c_int64, C.POINTER(C.c_int64)]
    L.adm_dofs_to_nodes.argtypes = [_vp, _vp, _vp, _vp]
    L.adm_nodes_to_dofs.argtypes = [_vp, _vp, _vp, C.c_int, _vp]
    L.adm_get_face_state.argtypes = [_vp, _dp]
    L.adm_set_face_state.argtypes = [_vp, _dp]
    L.adm_set_deposition.argtypes = [_vp, _dp, _dp]
    L.adm_set_launch_config.argtypes = [_vp, C.c_int32, C.c_int32]
    L.adm_autotune.argtypes = [_vp, _vp, _vp, C.c_int32, _vp,
                               C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                               _dp]
    L.adm_get_cell_data.argtypes = [_vp, C.POINTER(C.c_uint8), _dp, _dp]
    L.adm_set_beam_parameters.argtypes = [_vp, C.c_int, C.c_double, C.c_double,
                                          C.c_double, C.c_double]
    L.adm_beam_state_at.argtypes = [_vp, C.c_int, C.c_double,
                                    C.POINTER(BeamState)]
    _ip32 = C.POINTER(C.c_int32)
    L.adm_cells_to_activate.argtypes = [_vp, C.c_int32, _dp, _ip32]
    L.adm_add_material.argtypes = [_vp, _ip32, _dp, _dp, C.c_double, _vp, _vp]
    L.adm_activate.argtypes = [_vp, C.c_int32, _dp, _dp, _dp, C.c_double, _vp,
                               _ip32, _vp]
    L.adm_adjoint_gradient.argtypes = [
        _vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, _vp, C.c_int32,
        C.POINTER(C.c_int64), _dp, _dp, _vp, _dp, _dp, _vp, _vp]
    L.adm_operator_vjp.argtypes = [_vp, C.c_double, _vp, _vp, _vp, _dp, _vp]
    _lib = L
    return L


def check(rc: int):
    if rc != 0:
        raise AdmError(load_library().adm_last_error().decode())


def _stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _dev_ptr(t, n: Optional[int] = None):
    """Device address of a vector handed to the library. With n, the vector
    must hold exactly n values: the kernels index it up to n - 1, so a
    shorter one is refused here, before any launch."""
    import torch
    if n is not None and (not isinstance(t, torch.Tensor) or t.numel() != n):
        raise AdmError("device vector of %s values where %d are needed" %
                       (t.numel() if isinstance(t, torch.Tensor) else "?", n))
    if not (isinstance(t, torch.Tensor) and t.is_cuda and
            t.dtype == torch.float64 and t.is_contiguous()):
        raise AdmError("device vectors must be contiguous float64 HIP tensors")
    return C.c_void_p(t.data_ptr())


def _host(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class ThermalOperatorDevice:
    """Drop-in counterpart of ThermalOperatorDevice
    (source/ThermalOperatorDevice.hh) with the host operator's semantics."""

    def __init__(self, problem: Problem):
        L = load_library()
        self.problem = problem
        self._desc, self._keep = problem.to_c()
        h = _vp()
        check(L.adm_create(C.byref(self._desc), C.byref(h)))
        self._h = h
        n = C.c_int64()
        check(L.adm_n_dofs(h, C.byref(n)))
        self._n_dofs = n.value
        nc, nq, nfq = C.c_int64(), C.c_int32(), C.c_int32()
        check(L.adm_sizes(h, C.byref(nc), C.byref(nq), C.byref(nfq)))
        self.n_cells, self.n_q, self.n_face_q = nc.value, nq.value, nfq.value
        self._inverse_mass = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load_library().adm_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    # ThermalOperatorBase ---------------------------------------------------
    def m(self) -> int:
        return self._n_dofs

    def n(self) -> int:
        return self._n_dofs

    def n_nodes(self) -> int:
        """Nodes of the structured box numbering."""
        n = C.c_int64()
        check(load_library().adm_n_nodes(self._h, C.byref(n)))
        return n.value

    def set_dof_map(self, dof_to_node=None):
        """Vectors in the caller's DoF order (DoF i = box node
        dof_to_node[i]) for vmult, the inverse mass and evolve; None goes
        back to the box-node numbering. m() and n() follow."""
        if dof_to_node is None:
            check(load_library().adm_set_dof_map(self._h, -1, None))
        else:
            mp = np.ascontiguousarray(dof_to_node, dtype=np.int64)
            check(load_library().adm_set_dof_map(
                self._h, mp.size, mp.ctypes.data_as(C.POINTER(C.c_int64))))
        n = C.c_int64()
        check(load_library().adm_n_dofs(self._h, C.byref(n)))
        self._n_dofs = n.value
        self._inverse_mass = None

    def dofs_to_nodes(self, ext, node, stream=None):
        check(load_library().adm_dofs_to_nodes(
            self._h, _dev_ptr(ext, self._n_dofs),
            _dev_ptr(node, self.n_nodes()), _stream_ptr(stream)))

    def nodes_to_dofs(self, node, ext, add: bool = False, stream=None):
        check(load_library().adm_nodes_to_dofs(
            self._h, _dev_ptr(node, self.n_nodes()),
            _dev_ptr(ext, self._n_dofs), int(add), _stream_ptr(stream)))

    def initialize_dof_vector(self, value: float = 0.0, device=None):
        import torch
        return torch.full((self._n_dofs,), float(value), dtype=torch.float64,
                          device=device or "cuda")

    def vmult(self, dst, src, stream=None):
        check(load_library().adm_vmult(self._h, _dev_ptr(src, self._n_dofs),
                                       _dev_ptr(dst, self._n_dofs), 0,
                                       _stream_ptr(stream)))

    def vmult_add(self, dst, src, stream=None):
        check(load_library().adm_vmult(self._h, _dev_ptr(src, self._n_dofs),
                                       _dev_ptr(dst, self._n_dofs), 1,
                                       _stream_ptr(stream)))

    def compute_inverse_mass_matrix(self, stream=None):
        self._inverse_mass = self.initialize_dof_vector()
        check(load_library().adm_compute_inverse_mass(
            self._h, _dev_ptr(self._inverse_mass), _stream_ptr(stream)))

    def get_inverse_mass_matrix(self):
        if self._inverse_mass is None:
            self.compute_inverse_mass_matrix()
        return self._inverse_mass

    def set_time(self, t: float):
        check(load_library().adm_set_time(self._h, float(t)))

    def get_state_from_material_properties(self, cell_state):
        check(load_library().adm_set_cell_state(
            self._h, _host(cell_state).ctypes.data_as(_dp)))

    def set_state_to_material_properties(self):
        ns = self.problem.material.n_states
        out = np.zeros((self.n_cells, ns))
        check(load_library().adm_get_cell_state(self._h, out.ctypes.data_as(_dp)))
        return out

    def set_material_deposition_orientation(self, deposition_cos,
                                            deposition_sin):
        c = _host(deposition_cos)
        s = _host(deposition_sin)
        check(load_library().adm_set_deposition(self._h, c.ctypes.data_as(_dp),
                                                s.ctypes.data_as(_dp)))

    # quadrature-point state --------------------------------------------------
    def set_launch_config(self, mode: int, waves_per_block: int):
        """mode: LAUNCH_LINE_ATOMIC (0), LAUNCH_LINE_COLORED (1) or
        LAUNCH_PLANE (2, FE degrees 1 and 2); 2..4 wavefronts per block."""
        check(load_library().adm_set_launch_config(
            self._h, int(mode), int(waves_per_block)))

    def autotune(self, src, reps: int = 20, stream=None) -> dict:
        """Time vmult(src) under every launch configuration, keep the
        fastest; the phase state is left as it was."""
        dst = self.initialize_dof_vector()
        col, w = C.c_int32(), C.c_int32()
        ms = np.zeros(N_LAUNCH_CONFIGS)
        check(load_library().adm_autotune(
            self._h, _dev_ptr(src, self._n_dofs),
            _dev_ptr(dst, self._n_dofs), int(reps),
            _stream_ptr(stream), C.byref(col), C.byref(w),
            ms.ctypes.data_as(_dp)))
        configs = LAUNCH_CONFIGS
        return {"mode": int(col.value),
                "use_coloring": col.value == LAUNCH_LINE_COLORED,
                "waves_per_block": int(w.value),
                "ms": dict(zip(configs, ms.tolist()))}

    def cell_data(self):
        """(active flags, deposition cos, deposition sin) per cell as they
        stand after the activations so far."""
        active = np.zeros(self.n_cells, np.uint8)
        cos_, sin_ = np.zeros(self.n_cells), np.zeros(self.n_cells)
        check(load_library().adm_get_cell_data(
            self._h, active.ctypes.data_as(C.POINTER(C.c_uint8)),
            cos_.ctypes.data_as(_dp), sin_.ctypes.data_as(_dp)))
        return active, cos_, sin_

    def qp_state(self):
        liq = np.zeros((self.n_cells, self.n_q))
        pw = np.zeros((self.n_cells, self.n_q))
        check(load_library().adm_get_qp_state(self._h, liq.ctypes.data_as(_dp),
                                              pw.ctypes.data_as(_dp)))
        return liq, pw

    def set_qp_state(self, liquid=None, powder=None):
        lq = None if liquid is None else _host(liquid)
        pw = None if powder is None else _host(powder)
        check(load_library().adm_set_qp_state(
            self._h, None if lq is None else lq.ctypes.data_as(_dp),
            None if pw is None else pw.ctypes.data_as(_dp)))

    def face_state(self):
        """Powder ratio at the points of the boundary faces with a
        boundary term, [n_cells][2 dim][n_face_q] (zero elsewhere)."""
        fp = np.zeros((self.n_cells, 2 * self.problem.dim, self.n_face_q))
        check(load_library().adm_get_face_state(self._h,
                                                fp.ctypes.data_as(_dp)))
        return fp

    def set_face_state(self, face_powder):
        fp = _host(face_powder)
        need = self.n_cells * 2 * self.problem.dim * self.n_face_q
        if fp.size != need:
            raise AdmError("face state of %d values where %d are needed" %
                           (fp.size, need))
        check(load_library().adm_set_face_state(self._h,
                                                fp.ctypes.data_as(_dp)))

    def beam_state_at(self, beam: int, t: float) -> BeamState:
        s = BeamState()
        check(load_library().adm_beam_state_at(self._h, beam, float(t),
                                               C.byref(s)))
        return s

    def set_beam_parameters(self, beam, absorption_efficiency, max_power,
                            radius, depth):
        check(load_library().adm_set_beam_parameters(
            self._h, int(beam), float(absorption_efficiency), float(max_power),
            float(radius), float(depth)))


class ThermalPhysicsDevice:
    """Counterpart of ThermalPhysics<dim, ..., MemorySpace::Default>
    (source/ThermalPhysics.hh) for the explicit steppers."""

    def __init__(self, problem: Problem, method: Optional[int] = None):
        self.problem = problem
        self.method = problem.method if method is None else method
        self.operator = ThermalOperatorDevice(problem)

    def setup(self):
        """setup(): DoFs, deposition orientation, inverse mass, state
        (source/ThermalPhysics.templates.hh:496-506)."""
        self.operator.compute_inverse_mass_matrix()

    def initialize_dof_vector(self, value: float = 0.0):
        return self.operator.initialize_dof_vector(value)

    def evolve_one_time_step(self, t: float, delta_t: float, solution,
                             stream=None) -> float:
        return self.evolve(t, delta_t, 1, solution, stream=stream)

    def evolve(self, t: float, delta_t: float, n_steps: int, solution,
               use_graph: bool = False, stream=None) -> float:
        t_end = C.c_double()
        check(load_library().adm_evolve(
            self.operator.handle, int(self.method), float(t), float(delta_t),
            int(n_steps), _dev_ptr(solution, self.operator.m()),
            int(use_graph), C.byref(t_end),
            _stream_ptr(stream)))
        return t_end.value

    def update_physics_parameters(self, beam: int, absorption_efficiency,
                                  max_power, radius, depth):
        """update_physics_parameters (source/ThermalPhysics.templates.hh:
        827-852) for one beam."""
        self.operator.set_beam_parameters(beam, absorption_efficiency,
                                          max_power, radius, depth)

    def get_state_from_material_properties(self, cell_state):
        self.operator.get_state_from_material_properties(cell_state)

    def elements_to_activate(self, boxes) -> np.ndarray:
        """get_elements_to_activate (source/material_deposition.cc:376-435):
        per cell, the last deposition box (rows of [min x, y, z, max x, y,
        z]) meeting an inactive cell, else -1."""
        b = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 6)
        out = np.zeros(self.operator.n_cells, dtype=np.int32)
        check(load_library().adm_cells_to_activate(
            self.operator.handle, b.shape[0], b.ctypes.data_as(_dp),
            out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def add_material(self, box_of_cell, box_cos, box_sin,
                     new_material_temperature: float, solution,
                     stream=None):
        """add_material_start + add_material_end
        (source/ThermalPhysics.templates.hh:595-825)."""
        bc = np.ascontiguousarray(box_of_cell, dtype=np.int32)
        c = _host(box_cos)
        s_ = _host(box_sin)
        check(load_library().adm_add_material(
            self.operator.handle, bc.ctypes.data_as(C.POINTER(C.c_int32)),
            c.ctypes.data_as(_dp), s_.ctypes.data_as(_dp),
            float(new_material_temperature),
            None if solution is None else _dev_ptr(solution,
                                                   self.operator.m()),
            _stream_ptr(stream)))
        # compute_inverse_mass_matrix (source/ThermalPhysics.templates.hh:741)
        self.operator._inverse_mass = None

    def activate(self, boxes, box_cos, box_sin,
                 new_material_temperature: float, solution,
                 stream=None) -> np.ndarray:
        """elements_to_activate + add_material in one call, on the device
        (adm_activate); returns the box that activated each cell (or -1)."""
        b = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 6)
        c = _host(box_cos)
        s_ = _host(box_sin)
        if c.size != b.shape[0] or s_.size != b.shape[0]:
            raise ValueError("one cos and one sin per box")
        out = np.zeros(self.operator.n_cells, dtype=np.int32)
        check(load_library().adm_activate(
            self.operator.handle, b.shape[0], b.ctypes.data_as(_dp),
            c.ctypes.data_as(_dp), s_.ctypes.data_as(_dp),
            float(new_material_temperature),
            None if solution is None else _dev_ptr(solution,
                                                   self.operator.m()),
            out.ctypes.data_as(C.POINTER(C.c_int32)), _stream_ptr(stream)))
        self.operator._inverse_mass = None
        return out

    def set_state_to_material_properties(self):
        return self.operator.set_state_to_material_properties()

    def save_checkpoint(self, filename: str, solution):
        """save_checkpoint (source/ThermalPhysics.templates.hh:1087-1185):
        per cell the state ratios, the deposition cos/sin (10 on inactive
        cells), has_melted and the active flag, plus the temperature, into
        <filename>.npz. The mesh is the input file's box, so it is not
        saved. The cell state is the average of the quadrature-point state
        at this moment; the quadrature-point liquid and powder ratios and
        the boundary faces' powder are saved too, so that a restart goes on
        from the same point history (the reference keeps only the cell
        averages)."""
        active, cos_, sin_ = self.operator.cell_data()
        state = self.set_state_to_material_properties()
        lq, pw = self.operator.qp_state()
        on = active.astype(bool)
        T = solution.detach().cpu().numpy() if hasattr(solution, "cpu") \
            else np.asarray(solution, np.float64)
        np.savez(filename + ".npz", temperature=T,
                 cell_state=state, active=active,
                 deposition_cos=np.where(on, cos_, 10.0),
                 deposition_sin=np.where(on, sin_, 10.0),
                 has_melted=np.zeros(self.operator.n_cells, np.uint8),
                 liquid_qp=lq, powder_qp=pw,
                 face_powder=self.operator.face_state())

    def load_checkpoint(self, filename: str):
        """load_checkpoint (source/ThermalPhysics.templates.hh:974-1085):
        rebuilds the operator with the saved active cells, orientations and
        cell states, redoes setup(), puts back the saved quadrature-point
        and face-point state (a file without it: every point takes its
        cell's state, as in the reference), and returns the saved
        temperature on the device."""
        import dataclasses
        import torch
        with np.load(filename + ".npz", allow_pickle=False) as f:
            data = {k: f[k] for k in f.files}
        n = self.operator.n_cells
        if data["active"].shape != (n,) or data["temperature"].shape != (
                self.operator.m(),):
            raise ValueError("checkpoint %s does not match the problem" %
                             filename)
        on = data["active"].astype(bool)
        p = dataclasses.replace(
            self.problem, cell_active=data["active"].astype(np.uint8),
            cell_state=data["cell_state"],
            deposition_cos=np.where(on, data["deposition_cos"], 1.0),
            deposition_sin=np.where(on, data["deposition_sin"], 0.0))
        self.problem = p
        self.operator = ThermalOperatorDevice(p)
        self.setup()
        if "powder_qp" in data:
            # the saved point history, not the cell averages
            self.operator.set_qp_state(data["liquid_qp"], data["powder_qp"])
            self.operator.set_face_state(data["face_powder"])
        return torch.tensor(data["temperature"], dtype=torch.float64,
                            device="cuda")
