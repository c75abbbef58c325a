"""The two-dimensional device path against the oracle. The reference
instantiates its operators and physics for dim = 2 as well as 3
(source/instantiation.hh:121-179, source/ThermalPhysicsInst*.cc); its 2-D
meshes are (length, height) boxes whose second axis is the height, with
boundary ids 0..3 (x-, x+, bottom, top).

Bars as for 3-D: vmult within 1e-12 relative L2, steps within 1e-10, the
lumped mass and the quadrature-point and face-point phase state bit-exact;
the reference's own 2-D physics tests (thermal_2d l2 norm, energy
conservation) on the device."""
import numpy as np
import pytest

from adamantine_amd.problem import (BC_ADIABATIC, BC_CONVECTIVE, BC_RADIATIVE,
                                    FORWARD_EULER, RK_FOURTH_ORDER,
                                    RK_THIRD_ORDER, SOURCE_CUBE,
                                    SOURCE_ELECTRON_BEAM, SOURCE_GOLDAK, Beam,
                                    Material, Problem, read_scan_path,
                                    segments_from_path)
from adamantine_amd.workloads import slp_tables
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

VMULT_TOL = 1e-12
STEP_TOL = 1e-10
MM = 1e-3


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def poly_material():
    """Solid/liquid/powder polynomials of degree 2 (p_order 2)."""
    m = Material(n_materials=1, n_states=3, use_table=False, p_order=2,
                 single_material_layout=True)
    for st, rho_f, k_f in (("solid", 1.0, 1.0), ("liquid", 0.97, 1.2),
                           ("powder", 0.6, 0.01)):
        m.set_poly(0, "density", st, [7900.0 * rho_f])
        m.set_poly(0, "specific_heat", st, [600.0, 0.1, 1e-5])
        m.set_poly(0, "thermal_conductivity_x", st,
                   [12.0 * k_f, 0.012 * k_f, 1e-6 * k_f])
        m.set_poly(0, "thermal_conductivity_z", st,
                   [9.0 * k_f, 0.010 * k_f, 2e-6 * k_f])
        m.set_poly(0, "emissivity", st, [0.3])
        m.set_poly(0, "convection_heat_transfer_coef", st, [10.0])
    m.set_property(0, "solidus", 1675.0)
    m.set_property(0, "liquidus", 1708.0)
    m.set_property(0, "latent_heat", 2.6e5)
    m.set_property(0, "radiation_temperature_infty", 300.0)
    m.set_property(0, "convection_temperature_infty", 300.0)
    return m


def problem_2d(p=2, nc=(20, 12), kind="merged", source=SOURCE_GOLDAK,
               method=RK_FOURTH_ORDER, active_rows=None, powder_rows=2):
    """A (2 mm, 1 mm) 2-D box of nc cells, degree p; the top id (3)
    radiative and convective, the others adiabatic; a beam along the top of
    the active material (in 2-D the scan path's second coordinate is the
    height); the top `powder_rows` active rows hold powder."""
    from tests.test_gpu_configs import table_material
    ext = (2 * MM, 1.0, 1 * MM)
    if kind == "merged":
        m = slp_tables()
    elif kind == "poly":
        m = poly_material()
    else:
        m = table_material(kind)
    hz = ext[2] / nc[1]
    rows = nc[1] if active_rows is None else active_rows
    H = rows * hz
    path = segments_from_path([
        {"mode": "point", "xyz": (0.2 * MM, H, 0.0), "pmod": 0.0,
         "param": 1e-9},
        {"mode": "line", "xyz": (1.8 * MM, H, 0.0), "pmod": 1.0,
         "param": 1.0},
    ])
    if source == SOURCE_CUBE:
        beam = Beam(type=SOURCE_CUBE, start_time=0.0, end_time=1.0,
                    value=4e12, min_point=(0.5 * MM, H - 3 * hz, 0.0),
                    max_point=(1.1 * MM, H, 0.0))
    else:
        beam = Beam(type=source, absorption_efficiency=0.4, max_power=400.0,
                    radius=120e-6, depth=80e-6, segments=path)
    prob = Problem(n_cells=(nc[0], 1, nc[1]), extent=ext, fe_degree=p,
                   quadrature=0, dim=2, material=m,
                   boundary_types=[BC_ADIABATIC] * 3 +
                   [BC_RADIATIVE | BC_CONVECTIVE],
                   beams=[beam], method=method)
    if active_rows is not None:
        prob.set_material_height(rows * hz)
    prob.set_powder_layer(H, powder_rows * hz)
    return prob


def face_mask(prob):
    """[n_cells][4]: the faces that carry a boundary term (a domain face of
    a non-adiabatic id, or a face towards an inactive cell, whose type is
    the last id's)."""
    nx, nz = prob.n_cells[0], prob.n_cells[2]
    act = (np.ones(nx * nz, np.uint8) if prob.cell_active is None
           else prob.cell_active)
    bts = prob.boundary_types
    mask = np.zeros((nx * nz, 4), bool)
    for c in range(nx * nz):
        if not act[c]:
            continue
        cx, cz = c % nx, c // nx
        for f in range(4):
            d, side = f // 2, f % 2
            cc = [cx, cz]
            cc[d] += 1 if side else -1
            if cc[d] < 0 or cc[d] >= (nx, nz)[d]:
                t = bts[f]
            elif act[cc[0] + nx * cc[1]]:
                continue
            else:
                t = bts[-1]
            mask[c, f] = bool(t & (BC_RADIATIVE | BC_CONVECTIVE))
    return mask


def gpu_vmult(op, T, t):
    import torch
    op.set_time(t)
    dst = op.initialize_dof_vector()
    op.vmult(dst, torch.tensor(T, device="cuda"))
    torch.cuda.synchronize()
    return dst.cpu().numpy()


def check_state(op, o, prob):
    lq, pw = op.qp_state()
    olq, opw = o.qp_state()
    np.testing.assert_array_equal(pw, opw)
    np.testing.assert_array_equal(lq, olq)
    mask = face_mask(prob)
    fp, ofp = op.face_state(), o.face_powder()
    assert fp.shape == ofp.shape == (prob.n_cells_total, 4, prob.fe_degree + 1)
    np.testing.assert_array_equal(fp[mask], ofp[mask])
    assert not fp[~mask].any()
    return mask


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_inverse_mass_2d_bit_exact(p):
    from adamantine_amd.thermal import ThermalOperatorDevice
    prob = problem_2d(p=p, nc=(9, 7), active_rows=5)
    op = ThermalOperatorDevice(prob)
    got = op.get_inverse_mass_matrix().cpu().numpy()
    want = Oracle(prob).inverse_mass()
    assert (want == 0.0).sum() > 0  # nodes of inactive rows only
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_vmult_2d_degrees(p):
    """Merged tables, Goldak beam on, cells and faces melting: two
    applications (the second from the first's phase state)."""
    from adamantine_amd.thermal import ThermalOperatorDevice
    print("2-D vmult, degree", p, flush=True)
    prob = problem_2d(p=p, nc=(14, 9) if p <= 3 else (9, 6), active_rows=7
                      if p <= 3 else 5)
    op = ThermalOperatorDevice(prob)
    o = Oracle(prob)
    for k, t in enumerate((2e-4, 8e-4)):
        T = np.random.RandomState(5 + k).uniform(1200.0, 2400.0, prob.n_dofs)
        o.set_time(t)
        want = o.vmult(T)
        got = gpu_vmult(op, T, t)
        assert rel(got, want) < VMULT_TOL, (t, rel(got, want))
    mask = check_state(op, o, prob)
    assert mask.sum() > 0
    _, opw = o.qp_state()
    assert ((opw > 0.0) & (opw < 1.0)).sum() > 0


@pytest.mark.parametrize("kind", ["poly", "general", "twelve"])
@pytest.mark.parametrize("mode", ["colored", "atomic"])
def test_vmult_2d_properties_and_launch(kind, mode):
    """Each property evaluator (polynomials, general tables, 12-entry
    tables), each line launch mode, and the electron beam."""
    from adamantine_amd.thermal import (LAUNCH_LINE_ATOMIC,
                                        LAUNCH_LINE_COLORED,
                                        ThermalOperatorDevice)
    print("2-D vmult:", kind, mode, flush=True)
    prob = problem_2d(p=2, nc=(16, 10), kind=kind,
                      source=SOURCE_ELECTRON_BEAM, active_rows=8)
    op = ThermalOperatorDevice(prob)
    op.set_launch_config(LAUNCH_LINE_COLORED if mode == "colored"
                         else LAUNCH_LINE_ATOMIC, 4)
    o = Oracle(prob)
    T = np.random.RandomState(7).uniform(250.0, 2600.0, prob.n_dofs)
    for t in (3e-4, 9e-4):
        o.set_time(t)
        assert rel(gpu_vmult(op, T, t), o.vmult(T)) < VMULT_TOL
    check_state(op, o, prob)
    # the beam is on and matters
    q = problem_2d(p=2, nc=(16, 10), kind=kind, source=SOURCE_ELECTRON_BEAM,
                   active_rows=8)
    q.beams[0].max_power = 0.0
    oq, o2 = Oracle(q), Oracle(prob)
    oq.set_time(9e-4)
    o2.set_time(9e-4)
    assert rel(oq.vmult(T), o2.vmult(T)) > 1e-6


def test_plane_launch_is_refused_in_2d():
    from adamantine_amd.thermal import (LAUNCH_PLANE, AdmError,
                                        ThermalOperatorDevice)
    op = ThermalOperatorDevice(problem_2d(p=2, nc=(6, 4)))
    with pytest.raises(AdmError):
        op.set_launch_config(LAUNCH_PLANE, 4)


@pytest.mark.parametrize("method", [FORWARD_EULER, RK_THIRD_ORDER,
                                    RK_FOURTH_ORDER])
def test_evolve_2d(method):
    """Explicit steps while powder melts under the beam, with the cube
    source for forward Euler; the captured loop gives the same bits."""
    import torch
    from adamantine_amd.thermal import ThermalPhysicsDevice
    print("2-D evolve, method", method, flush=True)
    src = SOURCE_CUBE if method == FORWARD_EULER else SOURCE_GOLDAK
    prob = problem_2d(p=3, nc=(16, 10), source=src, method=method,
                      active_rows=8)
    X = prob.node_coordinates()
    H = 8 * prob.extent[2] / 10
    T0 = np.where(X[:, 1] > 0.7 * H, 1650.0, 600.0) + \
        np.random.RandomState(9).uniform(0.0, 40.0, prob.n_dofs)
    dt, n = 2e-7, 8
    want, _ = Oracle(prob).evolve(method, 1e-4, dt, n, T0.copy())
    res = []
    for graph in (False, True):
        ph = ThermalPhysicsDevice(prob)
        ph.setup()
        sol = torch.tensor(T0, device="cuda")
        t = ph.evolve(1e-4, dt, n, sol, use_graph=graph)
        torch.cuda.synchronize()
        res.append(sol.cpu().numpy())
    assert rel(res[0], want) < STEP_TOL, rel(res[0], want)
    np.testing.assert_array_equal(res[1], res[0])
    assert np.abs(want - T0).max() > 1.0


def test_thermal_2d_explicit_device(golden):
    """thermal_2d_explicit (tests/test_thermal_physics.cc:15-26,
    tests/test_thermal_physics.hh:77-155) on the device: 2-D Q2 on 12 x 6 mm
    (4 x 5 cells), an electron beam with infinite radius and depth, forward
    Euler with dt = 0.05 to t = 0.1; l2 norm 0.291705 at 1e-3."""
    import torch
    from adamantine_amd.problem import Material
    from adamantine_amd.thermal import ThermalPhysicsDevice
    segs = read_scan_path(golden("scan_path_test_thermal_physics.txt"))
    eb = Beam(type=SOURCE_ELECTRON_BEAM, absorption_efficiency=0.1,
              max_power=1e300, radius=1e100 / 2.0, depth=1e100, segments=segs)
    m = Material(n_materials=1, n_states=3, use_table=False, p_order=1,
                 single_material_layout=True)
    for s in ("solid", "liquid", "powder"):
        for key in ("density", "specific_heat", "thermal_conductivity_x",
                    "thermal_conductivity_z"):
            m.set_poly(0, key, s, [1.0])
    prob = Problem(n_cells=(4, 1, 5), extent=(12e-3, 1.0, 6e-3), fe_degree=2,
                   dim=2, material=m, boundary_types=[BC_ADIABATIC] * 4,
                   beams=[eb], method=FORWARD_EULER)
    ph = ThermalPhysicsDevice(prob)
    ph.setup()
    sol = torch.zeros(prob.n_dofs, dtype=torch.float64, device="cuda")
    t = ph.evolve(0.0, 0.05, 2, sol)
    torch.cuda.synchronize()
    T = sol.cpu().numpy()
    assert abs(t - 0.1) <= 1e-3 * 0.1
    assert abs(np.linalg.norm(T) - 0.291705) <= 1e-3 * 0.291705
    want, _ = Oracle(prob).evolve(FORWARD_EULER, 0.0, 0.05, 2,
                                  np.zeros(prob.n_dofs))
    assert rel(T, want) < STEP_TOL


def test_energy_conservation_2d_device():
    """energy conservation (tests/test_thermal_physics.hh:287-393) on the
    device: a cube source of 5 for t < 5 in a 10 x 10 adiabatic box, rho 0.5,
    c_p 4; at t = 100 the field is uniform at 10.5 (1e-9)."""
    import torch
    from adamantine_amd.thermal import ThermalPhysicsDevice
    m = Material(n_materials=1, n_states=3, use_table=False, p_order=0,
                 single_material_layout=True)
    for s in ("solid", "liquid", "powder"):
        m.set_poly(0, "density", s, [0.5])
        m.set_poly(0, "specific_heat", s, [4.0])
        m.set_poly(0, "thermal_conductivity_x", s, [2.0])
        m.set_poly(0, "thermal_conductivity_z", s, [2.0])
    cube = Beam(type=SOURCE_CUBE, start_time=0.0, end_time=5.0, value=5.0,
                min_point=(4.0, 4.0, 0.0), max_point=(6.0, 6.0, 0.0))
    prob = Problem(n_cells=(10, 1, 10), extent=(10.0, 1.0, 10.0), fe_degree=2,
                   dim=2, material=m, boundary_types=[BC_ADIABATIC] * 4,
                   beams=[cube], method=FORWARD_EULER)
    ph = ThermalPhysicsDevice(prob)
    ph.setup()
    sol = torch.full((prob.n_dofs,), 10.0, dtype=torch.float64, device="cuda")
    t = 0.0
    while t < 100.0:
        t = ph.evolve(t, 0.05, 1, sol)
    torch.cuda.synchronize()
    T = sol.cpu().numpy()
    assert abs(T.mean() - 10.5) <= 1e-9 * 10.5
    assert abs(T.max() - T.min()) <= 1e-9 * T.max()


def test_activation_2d():
    """Deposition in 2-D (boxes of (x, -, z) bounds): the cells, the new
    temperatures, the inverse mass and the carried phase state equal the
    oracle's; then steps across the new material."""
    import torch
    from adamantine_amd.thermal import ThermalPhysicsDevice
    prob = problem_2d(p=2, nc=(14, 10), method=RK_FOURTH_ORDER,
                      active_rows=6, powder_rows=2)
    hz = prob.extent[2] / 10
    X = prob.node_coordinates()
    T0 = np.where(X[:, 1] > 4.5 * hz, 1690.0, 700.0) + \
        np.random.RandomState(2).uniform(0.0, 30.0, prob.n_dofs)
    boxes = np.array([[0.3 * MM, 0.0, 6.1 * hz, 1.2 * MM, 0.0, 6.9 * hz],
                      [1.0 * MM, 0.0, 6.2 * hz, 1.9 * MM, 0.0, 7.8 * hz]])
    cs, sn = np.array([1.0, 0.0]), np.array([0.0, 1.0])
    o = Oracle(prob)
    ph = ThermalPhysicsDevice(prob)
    ph.setup()
    sol = torch.tensor(T0, device="cuda")
    T = T0.copy()
    # a few steps first, so that the carried state is not uniform
    ph.evolve(0.0, 2e-7, 3, sol)
    T, t = o.evolve(RK_FOURTH_ORDER, 0.0, 2e-7, 3, T)
    boc = ph.activate(boxes, cs, sn, 1550.0, sol)
    want_boc = o.cells_to_activate(boxes)
    np.testing.assert_array_equal(boc, want_boc)
    assert (boc >= 0).sum() > 0
    o.add_material(want_boc, cs, sn, 1550.0, T)
    torch.cuda.synchronize()
    assert rel(sol.cpu().numpy(), T) < STEP_TOL
    np.testing.assert_array_equal(
        ph.operator.get_inverse_mass_matrix().cpu().numpy(), o.inverse_mass())
    lq, pw = ph.operator.qp_state()
    olq, opw = o.qp_state()
    np.testing.assert_allclose(pw, opw, rtol=1e-15, atol=0)
    ph.evolve(t, 2e-7, 4, sol)
    T, _ = o.evolve(RK_FOURTH_ORDER, t, 2e-7, 4, T)
    torch.cuda.synchronize()
    assert rel(sol.cpu().numpy(), T) < STEP_TOL


def test_ensemble_2d():
    """A 2-D ensemble (members with their own absorption and conductivity
    scale) stepped in single launches over (cells, members) equals each
    member's oracle."""
    import torch
    from adamantine_amd.ensemble import EnsembleDevice, Member, member_problem
    prob = problem_2d(p=2, nc=(12, 8), method=RK_FOURTH_ORDER, active_rows=7)
    members = [Member(absorption_efficiency=[0.3 + 0.1 * j],
                      max_power=[prob.beams[0].max_power],
                      conductivity_scale=1.0 + 0.1 * j) for j in range(3)]
    ens = EnsembleDevice(prob, members, method=RK_FOURTH_ORDER)
    ens.compute_inverse_mass_matrix()
    T0 = np.stack([np.random.RandomState(30 + j).uniform(1500.0, 1900.0,
                                                         prob.n_dofs)
                   for j in range(3)])
    sol = torch.tensor(T0, device="cuda")
    ens.evolve(2e-4, 2e-7, 5, sol)
    torch.cuda.synchronize()
    got = sol.cpu().numpy()
    for j in range(3):
        want, _ = Oracle(member_problem(prob, members[j])).evolve(
            RK_FOURTH_ORDER, 2e-4, 2e-7, 5, T0[j].copy())
        assert rel(got[j], want) < STEP_TOL, (j, rel(got[j], want))


def test_dof_map_2d():
    """A permuted DoF numbering through the C-ABI's DoF map."""
    import torch
    from adamantine_amd.thermal import ThermalOperatorDevice
    prob = problem_2d(p=2, nc=(8, 6))
    op = ThermalOperatorDevice(prob)
    perm = np.random.RandomState(3).permutation(prob.n_dofs).astype(np.int64)
    op.set_dof_map(perm)
    assert op.m() == prob.n_dofs
    T = np.random.RandomState(4).uniform(300.0, 1800.0, prob.n_dofs)
    o = Oracle(prob)
    o.set_time(5e-4)
    want = o.vmult(T)
    op.set_time(5e-4)
    dst = op.initialize_dof_vector()
    op.vmult(dst, torch.tensor(T[perm], device="cuda"))
    torch.cuda.synchronize()
    assert rel(dst.cpu().numpy(), want[perm]) < VMULT_TOL


def test_adjoint_is_three_dimensional():
    import torch
    from adamantine_amd.adjoint import adjoint_gradient
    from adamantine_amd.thermal import AdmError, ThermalPhysicsDevice
    prob = problem_2d(p=2, nc=(6, 4))
    ph = ThermalPhysicsDevice(prob)
    ph.setup()
    T0 = torch.full((prob.n_dofs,), 300.0, dtype=torch.float64, device="cuda")
    with pytest.raises(AdmError, match="three-dimensional"):
        adjoint_gradient(ph, 0.0, 1e-7, 2, T0, np.array([0], np.int64),
                         np.array([300.0]))


def test_checkpoint_2d(tmp_path):
    """A 2-D checkpoint taken while powder melts restarts bit for bit
    (temperatures, point and face state)."""
    import torch
    from adamantine_amd.thermal import ThermalPhysicsDevice
    prob = problem_2d(p=2, nc=(12, 8), method=RK_FOURTH_ORDER, active_rows=7)
    X = prob.node_coordinates()
    T0 = np.where(X[:, 1] > 0.5 * prob.extent[2], 1690.0, 500.0) + \
        np.random.RandomState(12).uniform(0.0, 40.0, prob.n_dofs)
    a = ThermalPhysicsDevice(prob)
    a.setup()
    sol = torch.tensor(T0, device="cuda")
    t = a.evolve(0.0, 2e-7, 6, sol)
    prefix = str(tmp_path / "ck2d")
    a.save_checkpoint(prefix, sol)
    _, pw = a.operator.qp_state()
    assert ((pw > 0.0) & (pw < 1.0)).sum() > 0
    t_a = a.evolve(t, 2e-7, 6, sol)
    torch.cuda.synchronize()
    b = ThermalPhysicsDevice(prob)
    T = b.load_checkpoint(prefix)
    t_b = b.evolve(t, 2e-7, 6, T)
    torch.cuda.synchronize()
    assert t_b == t_a
    np.testing.assert_array_equal(T.cpu().numpy(), sol.cpu().numpy())
    np.testing.assert_array_equal(b.operator.qp_state()[1],
                                  a.operator.qp_state()[1])
    np.testing.assert_array_equal(b.operator.face_state(),
                                  a.operator.face_state())
