"""Discrete adjoint of the explicit steppers (no reference counterpart).

``adjoint_gradient`` runs the forward problem from ``T0`` exactly as
``ThermalPhysicsDevice.evolve`` does and returns the final-time misfit
J = 1/2 sum_k w_k (T[obs_dof_k] - obs_k)^2, its gradient with respect to the
initial temperature and with respect to three physics parameters:

  0  absorption efficiency of beam 0,
  1  a factor s on every thermal conductivity (derivative at s = 1),
  2  radius of beam 0.

The work is done by ``adm_adjoint_gradient`` of libadm.so: the transposed
stage sequence of forward Euler or classic RK4, the linearised cell and face
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
        int(checkpoint_every), _dev_ptr(T0), int(dof.size),
        dof.ctypes.data_as(C.POINTER(C.c_int64)), y.ctypes.data_as(_dp),
        None if w is None else w.ctypes.data_as(_dp), _dev_ptr(T_final),
        C.byref(J), g.ctypes.data_as(_dp), _dev_ptr(lam), _stream_ptr(stream)))
    return AdjointResult(J.value, g, T_final, lam)


def operator_vjp(op, t: float, T, mu, stream=None):
    """(dA/dT)^T mu at T and mu . dA/dtheta, with the beams at time t and
    the operator's current phase state (not updated). Returns (out, grad)."""
    import torch
    out = torch.empty_like(T)
    g = np.zeros(3)
    check(load_library().adm_operator_vjp(
        op.handle, float(t), _dev_ptr(T), _dev_ptr(mu), _dev_ptr(out),
        g.ctypes.data_as(_dp), _stream_ptr(stream)))
    return out, g
