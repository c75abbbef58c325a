"""Brute-force finite-element assembly used to pin the oracle.

Independent of the oracle and of the library: the quadrature comes from
numpy's Gauss-Legendre rule, the Lagrange basis on the Gauss-Lobatto support
points, written in closed form (the end points and the roots of P_p' on
[-1, 1]: 0; +-1/sqrt(5); 0, +-sqrt(3/7); +-sqrt(1/3 -+ 2 sqrt(7)/21)), is
evaluated directly (no sum factorisation), and element matrices are
assembled into dense global matrices on the lexicographic node numbering.
"""
import itertools

import numpy as np


def _gl01(interior):
    r = np.sort(np.concatenate([[-1.0, 1.0], interior]))
    return 0.5 * (r + 1.0)


SUPPORT = {
    1: np.array([0.0, 1.0]),
    2: np.array([0.0, 0.5, 1.0]),
    3: _gl01([-1.0 / np.sqrt(5.0), 1.0 / np.sqrt(5.0)]),
    4: _gl01([-np.sqrt(3.0 / 7.0), 0.0, np.sqrt(3.0 / 7.0)]),
    5: _gl01([s * np.sqrt(1.0 / 3.0 + t * 2.0 * np.sqrt(7.0) / 21.0)
              for s in (-1.0, 1.0) for t in (-1.0, 1.0)]),
}


def gauss01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def lagrange_1d(p, x):
    """values [len(x), p+1] and derivatives of the 1-D Lagrange basis."""
    s = SUPPORT[p]
    n = p + 1
    V = np.ones((len(x), n))
    D = np.zeros((len(x), n))
    for i in range(n):
        for j in range(n):
            if j != i:
                V[:, i] *= (x - s[j]) / (s[i] - s[j])
        for m in range(n):
            if m == i:
                continue
            t = np.full(len(x), 1.0 / (s[i] - s[m]))
            for j in range(n):
                if j != i and j != m:
                    t *= (x - s[j]) / (s[i] - s[j])
            D[:, i] += t
    return V, D


def cell_nodes(nc, p, cc):
    """global node indices of cell cc (x fastest), lexicographic local order"""
    dim = len(nc)
    nn = [p * n + 1 for n in nc]
    out = []
    for loc in itertools.product(*[range(p + 1)] * dim):
        loc = loc[::-1]  # x fastest
        idx, stride = 0, 1
        for d in range(dim):
            idx += (cc[d] * p + loc[d]) * stride
            stride *= nn[d]
        out.append(idx)
    return np.array(out)


def cells(nc):
    for cc in itertools.product(*[range(n) for n in nc[::-1]]):
        yield cc[::-1]


def element_tensors(dim, p, h, nq=None):
    """shape values [q, i], reference gradients [q, i, d], JxW [q]."""
    nq = nq or p + 1
    xq, wq = gauss01(nq)
    V1, D1 = lagrange_1d(p, xq)
    n = p + 1
    qs = list(itertools.product(range(nq), repeat=dim))
    qs = [q[::-1] for q in qs]
    ls = [l[::-1] for l in itertools.product(range(n), repeat=dim)]
    N = np.zeros((len(qs), len(ls)))
    G = np.zeros((len(qs), len(ls), dim))
    W = np.zeros(len(qs))
    X = np.zeros((len(qs), dim))
    for a, q in enumerate(qs):
        W[a] = np.prod([wq[q[d]] for d in range(dim)]) * np.prod(h)
        X[a] = [xq[q[d]] * h[d] for d in range(dim)]
        for b, l in enumerate(ls):
            N[a, b] = np.prod([V1[q[d], l[d]] for d in range(dim)])
            for e in range(dim):
                G[a, b, e] = np.prod([(D1 if d == e else V1)[q[d], l[d]]
                                      for d in range(dim)]) / h[e]
    return N, G, W, X


def stiffness(nc, extent, p, kdiag, nq=None):
    """Global matrix of  int K grad(phi_i) . grad(phi_j)  (K diagonal)."""
    dim = len(nc)
    h = [extent[d] / nc[d] for d in range(dim)]
    N, G, W, _ = element_tensors(dim, p, h, nq)
    Ke = np.einsum("qid,qjd,d,q->ij", G, G, np.asarray(kdiag, float), W)
    n = int(np.prod([p * c + 1 for c in nc]))
    A = np.zeros((n, n))
    for cc in cells(nc):
        idx = cell_nodes(nc, p, cc)
        A[np.ix_(idx, idx)] += Ke
    return A


def stiffness_tensor(nc, extent, p, Kmat, nq=None):
    """Global matrix of  int (K grad phi_j) . grad phi_i  for a full K."""
    dim = len(nc)
    h = [extent[d] / nc[d] for d in range(dim)]
    N, G, W, _ = element_tensors(dim, p, h, nq)
    Ke = np.einsum("qid,de,qje,q->ij", G, np.asarray(Kmat, float), G, W)
    n = int(np.prod([p * c + 1 for c in nc]))
    A = np.zeros((n, n))
    for cc in cells(nc):
        idx = cell_nodes(nc, p, cc)
        A[np.ix_(idx, idx)] += Ke
    return A


def boundary_mass(nc, extent, p, face):
    """Global matrix of  int_face phi_i phi_j  over the domain face `face`
    (colorized id 2*d + side)."""
    dim = len(nc)
    d, side = divmod(face, 2)
    h = [extent[k] / nc[k] for k in range(dim)]
    xq, wq = gauss01(p + 1)
    V1, _ = lagrange_1d(p, xq)
    end = np.array([1.0 if side else 0.0])
    Ve, _ = lagrange_1d(p, end)
    n = int(np.prod([p * c + 1 for c in nc]))
    A = np.zeros((n, n))
    tang = [k for k in range(dim) if k != d]
    qs = [q[::-1] for q in itertools.product(range(p + 1), repeat=dim - 1)]
    ls = [l[::-1] for l in itertools.product(range(p + 1), repeat=dim)]
    for cc in cells(nc):
        if cc[d] != (nc[d] - 1 if side else 0):
            continue
        idx = cell_nodes(nc, p, cc)
        M = np.zeros((len(ls), len(ls)))
        for q in qs:
            w = np.prod([wq[q[t]] * h[tang[t]] for t in range(dim - 1)])
            phi = np.array([
                Ve[0, l[d]] * np.prod([V1[q[t], l[tang[t]]]
                                       for t in range(dim - 1)]) for l in ls])
            M += w * np.outer(phi, phi)
        A[np.ix_(idx, idx)] += M
    return A
