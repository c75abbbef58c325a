"""Discrete adjoint of the explicit steppers (no reference counterpart).

``adjoint_gradient`` runs the forward problem from ``T0`` exactly as
``ThermalPhysicsDevice.evolve`` does and returns the final-time misfit
J = 1/2 sum_k w_k (T[obs_dof_k] - obs_k)^2, its gradient with respect to the
initial temperature and with respect to three physics parameters:

  0  absorption efficiency of beam 0,
  1  a factor s on every thermal conductivity (derivative at s = 1),
  2  radius of beam 0.

The work is done by ``adm_adjoint_gradient`` of libadm.so: the transposed
stage sequence of forward Euler, RK3 or classic RK4, the linearised cell and face
operators as HIP kernels, the trajectory recomputed from checkpoints.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .thermal import (ThermalPhysicsDevice, _dev_ptr, _dp, _stream_ptr, check,
                      load_library)

PARAMETERS = ("absorption_efficiency", "conductivity_scale", "beam_radius")


@dataclass
class AdjointResult:
    J: float
    grad: np.ndarray      # [3], order of PARAMETERS
    T_final: object       # device tensor
    lambda0: object       # device tensor: dJ/dT0


def adjoint_gradient(physics: ThermalPhysicsDevice, t0: float, dt: float,
                     n_steps: int, T0, obs_dof, obs, obs_weight=None,
                     checkpoint_every: int = 0, method=None,
                     stream=None) -> AdjointResult:
    import torch
    op = physics.operator
    dof = np.ascontiguousarray(obs_dof, dtype=np.int64)
    y = np.ascontiguousarray(obs, dtype=np.float64)
    if dof.shape != y.shape or dof.ndim != 1:
        raise ValueError("obs_dof and obs must be 1-D of equal length")
    w = None
    if obs_weight is not None:
        w = np.ascontiguousarray(obs_weight, dtype=np.float64)
        if w.shape != y.shape:
            raise ValueError("obs_weight must match obs")
    T_final = torch.empty_like(T0)
    lam = torch.empty_like(T0)
    J = C.c_double()
    g = np.zeros(3)
    m = physics.method if method is None else method
    check(load_library().adm_adjoint_gradient(
        op.handle, int(m), float(t0), float(dt), int(n_steps),
        int(checkpoint_every), _dev_ptr(T0, op.m()), int(dof.size),
        dof.ctypes.data_as(C.POINTER(C.c_int64)), y.ctypes.data_as(_dp),
        None if w is None else w.ctypes.data_as(_dp), _dev_ptr(T_final, op.m()),
        C.byref(J), g.ctypes.data_as(_dp), _dev_ptr(lam, op.m()),
        _stream_ptr(stream)))
    return AdjointResult(J.value, g, T_final, lam)


def operator_vjp(op, t: float, T, mu, stream=None):
    """(dA/dT)^T mu at T and mu . dA/dtheta, with the beams at time t and
    the operator's current phase state (not updated). Returns (out, grad)."""
    import torch
    out = torch.empty_like(T)
    g = np.zeros(3)
    check(load_library().adm_operator_vjp(
        op.handle, float(t), _dev_ptr(T, op.m()), _dev_ptr(mu, op.m()),
        _dev_ptr(out, op.m()),
        g.ctypes.data_as(_dp), _stream_ptr(stream)))
    return out, g


class Calibration:
    """Calibration harness over theta = (absorption efficiency of beam 0,
    conductivity scale s, radius of beam 0): the misfit of a forward run of
    ``n_steps`` from ``T0`` against point observations, differentiable in
    torch through ``misfit(theta, T0)``.

    The conductivity scale multiplies every thermal-conductivity table
    value or polynomial coefficient of ``problem``; a new s rebuilds the
    device operator (the adjoint gives the derivative around the operator
    as built, so dJ/ds = (dJ/ds')|_{s'=1} / s). Absorption efficiency and
    radius go through update_physics_parameters without a rebuild.

    Every run starts from the phase state the operator was built with (the
    point and face powder and liquid ratios are saved at the build and put
    back before each run), so J and its gradient do not depend on the runs
    made before."""

    def __init__(self, problem, t0: float, dt: float, n_steps: int, obs_dof,
                 obs, obs_weight=None, checkpoint_every: int = 0,
                 method=None):
        self.problem = problem
        self.t0, self.dt, self.n_steps = float(t0), float(dt), int(n_steps)
        self.obs_dof, self.obs, self.obs_weight = obs_dof, obs, obs_weight
        self.checkpoint_every = int(checkpoint_every)
        self.method = problem.method if method is None else method
        self._physics = None
        self._scale = None
        self._state0 = None

    def physics_for(self, theta) -> ThermalPhysicsDevice:
        import copy
        eta, s, radius = (float(v) for v in theta)
        if self._physics is None or s != self._scale:
            p = copy.deepcopy(self.problem)
            if s != 1.0:
                p.material.scale_conductivity(s)
            self._physics = ThermalPhysicsDevice(p, self.method)
            self._physics.setup()
            self._scale = s
            op = self._physics.operator
            self._state0 = (op.qp_state(), op.face_state())
        else:
            (lq, pw), fp = self._state0
            op = self._physics.operator
            op.set_qp_state(lq, pw)
            op.set_face_state(fp)
        b = self.problem.beams[0]
        self._physics.update_physics_parameters(0, eta, b.max_power, radius,
                                                b.depth)
        return self._physics

    def run(self, theta, T0) -> AdjointResult:
        return adjoint_gradient(self.physics_for(theta), self.t0, self.dt,
                                self.n_steps, T0, self.obs_dof, self.obs,
                                self.obs_weight, self.checkpoint_every,
                                self.method)

    def misfit(self, theta, T0):
        """J(theta, T0) as a torch scalar with gradients for theta (CPU,
        float64, 3 entries) and T0 (device vector)."""
        return _Misfit.apply(theta, T0, self)


def _misfit_function():
    import torch

    class Misfit(torch.autograd.Function):
        """torch.autograd.Function over adm_adjoint_gradient: forward runs
        the misfit and its adjoint once; backward scales the stored
        gradients by the incoming gradient."""

        @staticmethod
        def forward(ctx, theta, T0, cal):
            r = cal.run(theta.detach().cpu().numpy(), T0.detach())
            g = r.grad.copy()
            g[1] /= float(theta[1])   # derivative around the built operator
            ctx.save_for_backward(torch.tensor(g, dtype=theta.dtype),
                                  r.lambda0)
            return torch.tensor(r.J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            g, lam = ctx.saved_tensors
            return (gJ * g).to(gJ.dtype), gJ.to(lam.device) * lam, None

    return Misfit


class _MisfitProxy:
    _fn = None

    @classmethod
    def apply(cls, *args):
        if cls._fn is None:
            cls._fn = _misfit_function()
        return cls._fn.apply(*args)


_Misfit = _MisfitProxy


class ActivationEvent(C.Structure):
    """adm_activation_event: before step `step` of the window, member
    `member` activates the cells its boxes meet (adm_activate)."""
    _fields_ = [("member", C.c_int32), ("step", C.c_int32),
                ("n_boxes", C.c_int32), ("boxes", _dp), ("box_cos", _dp),
                ("box_sin", _dp), ("new_temperature", C.c_double)]


@dataclass
class EnsembleAdjointResult:
    J: np.ndarray            # [n_members]
    grad: np.ndarray         # [n_members, 3], PARAMETERS order
    T_final: object          # device [n_members, n_dofs]
    lambda0: object          # device [n_members, n_dofs]: dJ_j/dT0_j


def ensemble_adjoint_gradient(ensemble, t0: float, dt: float, n_steps: int,
                              T0, obs_dof, obs, obs_weight=None,
                              events=(), checkpoint_every: int = 0,
                              method=None,
                              stream=None) -> EnsembleAdjointResult:
    """Every member's misfit J_j = 1/2 sum_k w_k (T_j[obs_dof_k] - obs[j,k])^2
    at the end of n_steps and its gradient, by one batched forward and one
    batched backward sweep (adm_ensemble_adjoint_gradient). events: (member,
    step, boxes [k x 6], box_cos [k], box_sin [k], new_temperature) tuples,
    the activations the forward run makes before the given steps."""
    import torch
    L = load_library()
    from .ensemble import _bind
    _bind(L)
    if not getattr(L, "_ens_adj_bound", False):
        L.adm_ensemble_adjoint_gradient.argtypes = [
            C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
            C.c_void_p, C.c_int32, C.POINTER(C.c_int64), _dp, _dp, C.c_int32,
            C.POINTER(ActivationEvent), C.c_void_p, _dp, _dp, C.c_void_p,
            C.c_void_p]
        L._ens_adj_bound = True
    N = ensemble.n_members
    dof = np.ascontiguousarray(obs_dof, dtype=np.int64)
    y = np.ascontiguousarray(obs, dtype=np.float64)
    if dof.ndim != 1 or y.shape != (N, dof.size):
        raise ValueError("obs_dof: [n_obs]; obs: [n_members, n_obs]")
    w = None
    if obs_weight is not None:
        w = np.ascontiguousarray(obs_weight, dtype=np.float64)
        if w.shape != dof.shape:
            raise ValueError("obs_weight must match obs_dof")
    ensemble._check_vec(T0)
    keep = []
    ev = (ActivationEvent * max(len(events), 1))()
    for i, (member, step, boxes, cs, sn, new_t) in enumerate(events):
        b = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 6)
        c = np.ascontiguousarray(cs, dtype=np.float64).reshape(-1)
        s_ = np.ascontiguousarray(sn, dtype=np.float64).reshape(-1)
        if c.size != b.shape[0] or s_.size != b.shape[0]:
            raise ValueError("one cos and one sin per box")
        keep += [b, c, s_]
        ev[i] = ActivationEvent(int(member), int(step), b.shape[0],
                                b.ctypes.data_as(_dp), c.ctypes.data_as(_dp),
                                s_.ctypes.data_as(_dp), float(new_t))
    T_final = torch.empty_like(T0)
    lam = torch.empty_like(T0)
    J = np.zeros(N)
    g = np.zeros((N, 3))
    m = ensemble.method if method is None else method
    check(L.adm_ensemble_adjoint_gradient(
        ensemble._h, int(m), float(t0), float(dt), int(n_steps),
        int(checkpoint_every), _dev_ptr(T0), dof.size,
        dof.ctypes.data_as(C.POINTER(C.c_int64)), y.ctypes.data_as(_dp),
        None if w is None else w.ctypes.data_as(_dp), len(events), ev,
        _dev_ptr(T_final), J.ctypes.data_as(_dp), g.ctypes.data_as(_dp),
        _dev_ptr(lam), _stream_ptr(stream)))
    ensemble._inverse_mass = None
    return EnsembleAdjointResult(J, g, T_final, lam)
