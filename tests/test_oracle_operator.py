"""Oracle operator pinned by brute-force assembly, as the reference's
tests/test_thermal_operator.cc pins the matrix-free operator against
deal.II's assembled matrices (1e-12)."""
import numpy as np
import pytest

from adamantine_amd.problem import (BC_ADIABATIC, BC_CONVECTIVE, BC_RADIATIVE,
                                    SOURCE_GOLDAK, Beam, Material, Problem,
                                    segments_from_path)
from oracle.oracle import Oracle
from tests import fe_reference as fe

STATES = ("solid", "liquid", "powder")


def const_material(n_states=3, rho=1.0, cp=1.0, kx=1.0, ky=1.0, kz=1.0,
                   n_materials_layout=1, extra=None):
    m = Material(n_materials=1, n_states=n_states, use_table=False, p_order=2,
                 single_material_layout=bool(n_materials_layout))
    for s in STATES[:n_states]:
        m.set_poly(0, "density", s, [rho])
        m.set_poly(0, "specific_heat", s, [cp])
        m.set_poly(0, "thermal_conductivity_x", s, [kx])
        m.set_poly(0, "thermal_conductivity_y", s, [ky])
        m.set_poly(0, "thermal_conductivity_z", s, [kz])
        for k, v in (extra or {}).items():
            m.set_poly(0, k, s, [v])
    return m


def operator_matrix(o):
    n = o.n_dofs
    A = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        A[:, i] = o.vmult(e)
    return A


def problem_2d(m, bts=None):
    # Geometry of tests/test_thermal_operator.cc:44-55: 12 x 6, 4 x 5 cells
    return Problem(n_cells=(4, 1, 5), extent=(12.0, 1.0, 6.0), fe_degree=2,
                   dim=2, material=m, boundary_types=bts or [BC_ADIABATIC] * 4)


def test_thermal_operator_2d():
    # tests/test_thermal_operator.cc:41-146
    o = Oracle(problem_2d(const_material(kx=10.0, kz=10.0)))
    assert o.n_dofs == 99
    src = np.ones(o.n_dofs)
    dst1 = o.vmult(src)
    assert np.abs(dst1).sum() < 1e-12 * 10.0 * o.n_dofs
    dst2 = o.vmult(src, np.ones(o.n_dofs), add=True)
    l1, l2 = np.abs(dst1 + src).sum(), np.abs(dst2).sum()
    assert abs(l1 - l2) <= 1e-15 * l2


def test_spmv_laplace_2d():
    # tests/test_thermal_operator.cc:148-263: vmult(e_i) == -Laplace e_i
    o = Oracle(problem_2d(const_material(n_materials_layout=0)))
    A = operator_matrix(o)
    K = fe.stiffness((4, 5), (12.0, 6.0), 2, [1.0, 1.0])
    np.testing.assert_allclose(A, -K, rtol=0, atol=1e-12 * np.abs(K).max())


def test_spmv_anisotropic_2d():
    # tests/test_thermal_operator.cc:265-413 (k_z = 0)
    o = Oracle(problem_2d(const_material(kx=1.0, kz=0.0)))
    A = operator_matrix(o)
    K = fe.stiffness((4, 5), (12.0, 6.0), 2, [1.0, 0.0])
    np.testing.assert_allclose(A, -K, rtol=0, atol=1e-12 * np.abs(K).max())


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_spmv_rotated_anisotropic_3d(p):
    # tests/test_thermal_operator.cc:415-578: angle pi/6, k = (1, 0.8, 0.8)
    m = const_material(kx=1.0, ky=0.8, kz=0.8)
    nc = (3, 2, 2) if p <= 3 else (2, 1, 2)
    ext = (3.0, 2.0, 1.5)
    prob = Problem(n_cells=nc, extent=ext, fe_degree=p, dim=3, material=m)
    o = Oracle(prob)
    th = np.pi / 6
    o.set_deposition(np.full(o.n_cells, np.cos(th)), np.full(o.n_cells, np.sin(th)))
    A = operator_matrix(o)
    c, s = np.cos(th), np.sin(th)
    R = np.array([[c, -s], [s, c]])
    Kxy = R @ np.diag([1.0, 0.8]) @ R.T
    Kmat = np.zeros((3, 3))
    Kmat[:2, :2] = Kxy
    Kmat[2, 2] = 0.8
    K = fe.stiffness_tensor(nc, ext, p, Kmat)
    np.testing.assert_allclose(A, -K, rtol=0, atol=1e-10 * np.abs(K).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_convective_face(dim):
    # tests/test_thermal_operator.cc:772-952: h = 1, T_inf = 0, k = 0
    m = const_material(kx=0.0, ky=0.0, kz=0.0,
                       extra={"convection_heat_transfer_coef": 1.0})
    m.set_property(0, "convection_temperature_infty", 0.0)
    m.set_property(0, "radiation_temperature_infty", 0.0)
    nb = 2 * dim
    bts = [BC_CONVECTIVE] * nb
    if dim == 2:
        prob = Problem(n_cells=(4, 1, 5), extent=(12.0, 1.0, 6.0),
                       fe_degree=2, dim=2, material=m, boundary_types=bts)
        nc, ext = (4, 5), (12.0, 6.0)
    else:
        prob = Problem(n_cells=(3, 2, 2), extent=(3.0, 2.0, 1.5), fe_degree=2,
                       dim=3, material=m, boundary_types=bts)
        nc, ext = (3, 2, 2), (3.0, 2.0, 1.5)
    A = operator_matrix(Oracle(prob))
    M = sum(fe.boundary_mass(nc, ext, 2, f) for f in range(nb))
    np.testing.assert_allclose(A, -M, rtol=0, atol=1e-12 * np.abs(M).max())


def test_radiative_face_uniform():
    # radiative term -eps sigma (T + Tinf)(T^2 + Tinf^2)(T - Tinf) at uniform T
    eps, T0 = 1e5, 3.0
    m = const_material(kx=0.0, ky=0.0, kz=0.0, extra={"emissivity": eps})
    m.set_property(0, "radiation_temperature_infty", 0.0)
    bts = [BC_ADIABATIC] * 5 + [BC_RADIATIVE]
    prob = Problem(n_cells=(3, 2, 2), extent=(3.0, 2.0, 1.5), fe_degree=2,
                   dim=3, material=m, boundary_types=bts)
    o = Oracle(prob)
    got = o.vmult(np.full(o.n_dofs, T0))
    M = fe.boundary_mass((3, 2, 2), (3.0, 2.0, 1.5), 2, 5)
    want = -eps * 5.670374419e-8 * T0 ** 4 * M.sum(axis=1)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("p", [1, 2])
def test_nonlinear_cell_term_3d(p):
    """T-dependent k(T) and a Goldak source: the oracle's sum factorisation
    against direct evaluation at the quadrature points."""
    m = Material(n_materials=1, n_states=1, use_table=False, p_order=1)
    m.set_poly(0, "density", "solid", [7904.0])
    m.set_poly(0, "specific_heat", "solid", [714.0])
    for k in ("thermal_conductivity_x", "thermal_conductivity_y",
              "thermal_conductivity_z"):
        m.set_poly(0, k, "solid", [11.8, 0.0138])
    nc, ext = (3, 2, 2), (3e-4, 2e-4, 2e-4)
    segs = segments_from_path([
        {"mode": "point", "xyz": (1e-4, 1e-4, 2e-4), "pmod": 1.0, "param": 1.0}])
    beam = Beam(type=SOURCE_GOLDAK, absorption_efficiency=0.3, max_power=100.0,
                radius=8e-5, depth=6e-5, segments=segs)
    prob = Problem(n_cells=nc, extent=ext, fe_degree=p, dim=3, material=m,
                   beams=[beam])
    o = Oracle(prob)
    rng = np.random.RandomState(7)
    T = rng.uniform(300.0, 900.0, o.n_dofs)
    o.set_time(0.5)
    got = o.vmult(T)
    # direct evaluation
    from oracle import oracle as orc
    bs = orc.beam_state(prob.to_c()[0].beams[0], 0.5)
    h = [ext[d] / nc[d] for d in range(3)]
    N, G, W, X = fe.element_tensors(3, p, h)
    want = np.zeros(o.n_dofs)
    for cc in fe.cells(nc):
        idx = fe.cell_nodes(nc, p, cc)
        Tc = T[idx]
        Tq = N @ Tc
        gq = np.einsum("qid,i->qd", G, Tc)
        k = 11.8 + 0.0138 * Tq
        irc = 1.0 / (7904.0 * 714.0)
        flux = -irc * k[:, None] * gq
        pts = X + np.array([cc[d] * h[d] for d in range(3)])
        Q = np.array([orc.beam_value(bs, 3, x) for x in pts]) * irc
        want[idx] += np.einsum("qid,qd,q->i", G, flux, W) + N.T @ (Q * W)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("p,dim", [(1, 3), (2, 3), (3, 3), (4, 3), (5, 3),
                                   (2, 2)])
def test_inverse_mass(p, dim):
    """Lumped Gauss-Lobatto mass: sums to the active volume, equals the row
    sums of the Gauss-Lobatto mass matrix, and is zero off the active cells."""
    m = const_material()
    if dim == 3:
        prob = Problem(n_cells=(3, 2, 4), extent=(1.5, 1.0, 2.0), fe_degree=p,
                       dim=3, material=m)
        prob.set_material_height(1.2)  # 2 of 4 cell layers active
    else:
        prob = Problem(n_cells=(4, 1, 5), extent=(12.0, 1.0, 6.0),
                       fe_degree=p, dim=2, material=m, boundary_types=[1] * 4)
    o = Oracle(prob)
    im = o.inverse_mass()
    mass = np.where(im > 0, 1.0 / np.where(im > 0, im, 1.0), 0.0)
    if dim == 3:
        assert np.isclose(mass.sum(), 1.5 * 1.0 * 1.0, rtol=1e-13)
        z = prob.node_coordinates()[:, 2]
        assert np.all(im[z > 1.0 + 1e-12] == 0.0)
        assert np.all(im[z < 1.0 - 1e-12] > 0.0)
    else:
        assert np.isclose(mass.sum(), 72.0, rtol=1e-13)
