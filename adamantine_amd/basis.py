"""1-D quadrature rules and Lagrange bases on [0, 1].

FE_Q(p) uses the Gauss-Lobatto points as support points; the operator uses
QGauss(p+1) (or QGaussLobatto(p+1)) and the lumped mass QGaussLobatto(p+1)
(source/ThermalPhysics.templates.hh:364-369,
source/ThermalOperator.templates.hh:127-129). Points and weights are computed
by Newton iteration on Legendre polynomials in extended precision.
"""
from __future__ import annotations

import numpy as np


def _legendre(n, x):
    p0, p1 = np.longdouble(1.0), x
    if n == 0:
        return np.longdouble(1.0), np.longdouble(0.0)
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    dp = n * (x * p1 - p0) / (x * x - 1)
    return p1, dp


def gauss(n: int):
    x = np.zeros(n)
    w = np.zeros(n)
    pi = np.longdouble(np.pi)
    for i in range(n):
        z = np.cos(pi * (np.longdouble(i) + np.longdouble(0.75)) /
                   (np.longdouble(n) + np.longdouble(0.5)))
        for _ in range(100):
            p, dp = _legendre(n, z)
            dz = p / dp
            z -= dz
            if abs(dz) < 1e-19:
                break
        p, dp = _legendre(n, z)
        ww = 2 / ((1 - z * z) * dp * dp)
        x[n - 1 - i] = float(0.5 + 0.5 * z)
        w[n - 1 - i] = float(0.5 * ww)
    return x, w


def gauss_lobatto(n: int):
    m = n - 1
    x = np.zeros(n)
    w = np.zeros(n)
    x[0], x[-1] = 0.0, 1.0
    w[0] = w[-1] = float(np.longdouble(1.0) / (m * (m + 1)))
    pi = np.longdouble(np.pi)
    for i in range(1, n - 1):
        z = -np.cos(pi * i / m)
        for _ in range(200):
            p, dp = _legendre(m, z)
            d2p = (2 * z * dp - m * (m + 1) * p) / (1 - z * z)
            dz = dp / d2p
            z -= dz
            if abs(dz) < 1e-19:
                break
        p, dp = _legendre(m, z)
        x[i] = float(0.5 + 0.5 * z)
        w[i] = float(np.longdouble(1.0) / (m * (m + 1) * p * p))
    return x, w


def lagrange_matrix(support, points):
    """M[q, i] = l_i(points[q]) for the Lagrange basis on `support`."""
    n = len(support)
    M = np.ones((len(points), n))
    for q, x in enumerate(points):
        for i in range(n):
            v = 1.0
            for j in range(n):
                if j != i:
                    v *= (x - support[j]) / (support[i] - support[j])
            M[q, i] = v
    return M


def lagrange_deriv_matrix(support, points):
    """M[q, k] = l_k'(points[q])."""
    n = len(support)
    M = np.zeros((len(points), n))
    for q, x in enumerate(points):
        for k in range(n):
            s = 0.0
            for m in range(n):
                if m == k:
                    continue
                t = 1.0 / (support[k] - support[m])
                for j in range(n):
                    if j != k and j != m:
                        t *= (x - support[j]) / (support[k] - support[j])
                s += t
            M[q, k] = s
    return M


def tables(fe_degree: int, quadrature: int = 0):
    """(support, qpoints, qweights, shape S[q,i], collocation derivative
    Dc[q,k], Gauss-Lobatto weights) for FE_Q(fe_degree)."""
    n = fe_degree + 1
    xs, wgl = gauss_lobatto(n)
    xq, wq = gauss_lobatto(n) if quadrature == 1 else gauss(n)
    S = lagrange_matrix(xs, xq)
    Dc = lagrange_deriv_matrix(xq, xq)
    return xs, xq, wq, S, Dc, wgl
