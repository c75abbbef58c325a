"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker of the HIP library.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from adamantine_amd.problem import (BeamDesc, BeamState, MaterialDesc, Problem,
                                    ProblemDesc)

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(ProblemDesc)]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_n_dofs.restype = C.c_int64
        L.oracle_n_dofs.argtypes = [C.c_void_p]
        L.oracle_n_cells.restype = C.c_int64
        L.oracle_n_cells.argtypes = [C.c_void_p]
        L.oracle_n_q.argtypes = [C.c_void_p]
        L.oracle_n_face_q.argtypes = [C.c_void_p]
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_inverse_mass.argtypes = [C.c_void_p, _dp]
        L.oracle_set_time.argtypes = [C.c_void_p, C.c_double]
        L.oracle_vmult.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.oracle_evaluate.argtypes = [C.c_void_p, C.c_double, _dp, _dp]
        L.oracle_evolve.restype = C.c_double
        L.oracle_evolve.argtypes = [C.c_void_p, C.c_int, C.c_double,
                                    C.c_double, C.c_int, _dp]
        L.oracle_get_qp_state.argtypes = [C.c_void_p, _dp, _dp]
        L.oracle_set_qp_state.argtypes = [C.c_void_p, _dp, _dp]
        L.oracle_get_face_powder.argtypes = [C.c_void_p, _dp]
        L.oracle_set_cell_state.argtypes = [C.c_void_p, _dp]
        L.oracle_get_cell_state.argtypes = [C.c_void_p, _dp]
        L.oracle_set_deposition.argtypes = [C.c_void_p, _dp, _dp]
        L.oracle_set_beam_parameters.argtypes = [C.c_void_p, C.c_int,
                                                 C.c_double, C.c_double,
                                                 C.c_double, C.c_double]
        L.oracle_beam_update.argtypes = [C.POINTER(BeamDesc), C.c_double,
                                         C.POINTER(C.c_int),
                                         C.POINTER(BeamState)]
        L.oracle_scan_path_position.argtypes = [C.POINTER(BeamDesc), C.c_double,
                                                C.POINTER(C.c_int), _dp]
        L.oracle_scan_path_power_modifier.restype = C.c_double
        L.oracle_scan_path_power_modifier.argtypes = [
            C.POINTER(BeamDesc), C.c_double, C.POINTER(C.c_int)]
        L.oracle_beam_value.restype = C.c_double
        L.oracle_beam_value.argtypes = [C.POINTER(BeamState), C.c_int, _dp]
        L.oracle_quaternion_mul.argtypes = [_dp, _dp]
        L.oracle_quaternion_div.argtypes = [_dp, _dp]
        L.oracle_quaternion_pow.argtypes = [_dp, C.c_double]
        L.oracle_quaternion_matrix.argtypes = [_dp, _dp]
        L.oracle_rotate.argtypes = [_dp, C.c_int, _dp, _dp]
        L.oracle_scan_path_quaternion.argtypes = [
            C.POINTER(BeamDesc), C.c_double, C.POINTER(C.c_int), _dp]
        L.oracle_material_property.restype = C.c_double
        L.oracle_material_property.argtypes = [C.POINTER(MaterialDesc), C.c_int,
                                               C.c_int, _dp, C.c_double]
        L.oracle_property_from_table.restype = C.c_double
        L.oracle_property_from_table.argtypes = [C.POINTER(MaterialDesc),
                                                 C.c_int, C.c_int, C.c_int,
                                                 C.c_double]
        L.oracle_basis_1d.argtypes = [C.c_void_p] + [_dp] * 6
        _ip32 = C.POINTER(C.c_int32)
        L.oracle_cells_to_activate.argtypes = [C.c_void_p, C.c_int, _dp, _ip32]
        L.oracle_add_material.argtypes = [C.c_void_p, _ip32, _dp, _dp,
                                          C.c_double, _dp]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


def set_threads(n: int):
    lib().oracle_set_threads(int(n))


class Oracle:
    """CPU restatement of the host ThermalOperator + ThermalPhysics."""

    def __init__(self, problem: Problem):
        self.problem = problem
        self._desc, self._keep = problem.to_c()
        self._h = lib().oracle_create(C.byref(self._desc))
        if not self._h:
            raise RuntimeError("oracle_create failed")
        L = lib()
        self.n_dofs = L.oracle_n_dofs(self._h)
        self.n_cells = L.oracle_n_cells(self._h)
        self.n_q = L.oracle_n_q(self._h)
        self.n_face_q = L.oracle_n_face_q(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_destroy(self._h)
            self._h = None

    def inverse_mass(self):
        out = np.zeros(self.n_dofs)
        lib().oracle_inverse_mass(self._h, _ptr(out))
        return out

    def set_time(self, t):
        lib().oracle_set_time(self._h, float(t))

    def vmult(self, src, dst=None, add=False):
        src = np.ascontiguousarray(src, np.float64)
        if dst is None:
            dst = np.zeros(self.n_dofs)
        lib().oracle_vmult(self._h, _ptr(src), _ptr(dst), int(add))
        return dst

    def evaluate(self, t, y):
        y = np.ascontiguousarray(y, np.float64)
        out = np.zeros(self.n_dofs)
        lib().oracle_evaluate(self._h, float(t), _ptr(y), _ptr(out))
        return out

    def evolve(self, method, t, dt, n_steps, T):
        T = np.ascontiguousarray(T, np.float64)
        t_end = lib().oracle_evolve(self._h, int(method), float(t), float(dt),
                                    int(n_steps), _ptr(T))
        return T, t_end

    def qp_state(self):
        liq = np.zeros(self.n_cells * self.n_q)
        pow_ = np.zeros(self.n_cells * self.n_q)
        lib().oracle_get_qp_state(self._h, _ptr(liq), _ptr(pow_))
        return liq.reshape(self.n_cells, self.n_q), pow_.reshape(
            self.n_cells, self.n_q)

    def set_qp_state(self, liquid, powder):
        liquid = np.ascontiguousarray(liquid, np.float64).ravel()
        powder = np.ascontiguousarray(powder, np.float64).ravel()
        lib().oracle_set_qp_state(self._h, _ptr(liquid), _ptr(powder))

    def face_powder(self):
        dim = self.problem.dim
        out = np.zeros(self.n_cells * 2 * dim * self.n_face_q)
        lib().oracle_get_face_powder(self._h, _ptr(out))
        return out.reshape(self.n_cells, 2 * dim, self.n_face_q)

    def set_cell_state(self, cs):
        cs = np.ascontiguousarray(cs, np.float64)
        lib().oracle_set_cell_state(self._h, _ptr(cs))

    def cell_state(self):
        ns = self.problem.material.n_states
        out = np.zeros(self.n_cells * ns)
        lib().oracle_get_cell_state(self._h, _ptr(out))
        return out.reshape(self.n_cells, ns)

    def set_deposition(self, c, s):
        c = np.ascontiguousarray(c, np.float64)
        s = np.ascontiguousarray(s, np.float64)
        lib().oracle_set_deposition(self._h, _ptr(c), _ptr(s))

    def set_beam_parameters(self, beam, absorption, max_power, radius, depth):
        lib().oracle_set_beam_parameters(self._h, beam, absorption, max_power,
                                         radius, depth)

    def cells_to_activate(self, boxes):
        b = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 6)
        out = np.zeros(self.n_cells, dtype=np.int32)
        lib().oracle_cells_to_activate(self._h, b.shape[0], _ptr(b),
                                       out.ctypes.data_as(C.POINTER(C.c_int32)))
        return out

    def add_material(self, box_of_cell, box_cos, box_sin, new_temperature, T):
        """Activation; T (host float64) is updated in place."""
        bc = np.ascontiguousarray(box_of_cell, dtype=np.int32)
        c = np.ascontiguousarray(box_cos, dtype=np.float64)
        s_ = np.ascontiguousarray(box_sin, dtype=np.float64)
        assert T.dtype == np.float64 and T.flags.c_contiguous
        lib().oracle_add_material(self._h, bc.ctypes.data_as(C.POINTER(C.c_int32)),
                                  _ptr(c), _ptr(s_), float(new_temperature),
                                  _ptr(T))

    def basis_1d(self):
        n = self.problem.fe_degree + 1
        arrs = [np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n * n),
                np.zeros(n * n), np.zeros(n)]
        lib().oracle_basis_1d(self._h, *[_ptr(a) for a in arrs])
        xs, xq, wq, S, Dc, wgl = arrs
        return xs, xq, wq, S.reshape(n, n), Dc.reshape(n, n), wgl


def beam_state(beam_desc: BeamDesc, t: float, old_segment=None) -> BeamState:
    s = BeamState()
    seg = C.c_int(0 if old_segment is None else old_segment)
    lib().oracle_beam_update(C.byref(beam_desc), float(t), C.byref(seg),
                             C.byref(s))
    return s


def scan_path_quaternion(beam_desc: BeamDesc, t: float,
                         old_segment=None) -> np.ndarray:
    q = np.zeros(4)
    seg = C.c_int(0 if old_segment is None else old_segment)
    lib().oracle_scan_path_quaternion(C.byref(beam_desc), float(t),
                                      C.byref(seg), _ptr(q))
    return q


def quaternion_matrix(q) -> np.ndarray:
    q = np.ascontiguousarray(q, np.float64)
    R = np.zeros(9)
    lib().oracle_quaternion_matrix(_ptr(q), _ptr(R))
    return R


def rotate(R, p, inverse: bool = False) -> np.ndarray:
    R = np.ascontiguousarray(R, np.float64)
    p = np.ascontiguousarray(p, np.float64)
    out = np.zeros(3)
    lib().oracle_rotate(_ptr(R), int(inverse), _ptr(p), _ptr(out))
    return out


def quaternion_op(op: str, a, b=None, e: float = 0.0) -> np.ndarray:
    """a * b ("mul"), a / b ("div") or a^e ("pow") as Quaternion does."""
    a = np.array(a, np.float64)
    if op == "pow":
        lib().oracle_quaternion_pow(_ptr(a), float(e))
    else:
        b = np.ascontiguousarray(b, np.float64)
        getattr(lib(), "oracle_quaternion_" + op)(_ptr(a), _ptr(b))
    return a


def beam_value(state: BeamState, dim: int, point) -> float:
    pt = np.ascontiguousarray(point, np.float64)
    return lib().oracle_beam_value(C.byref(state), dim, _ptr(pt))


def material_property(mat_desc: MaterialDesc, prop: int, material_id: int,
                      ratios, T: float) -> float:
    r = np.ascontiguousarray(ratios, np.float64)
    return lib().oracle_material_property(C.byref(mat_desc), prop, material_id,
                                          _ptr(r), float(T))
