"""Parity of the HIP library (through the C-ABI) with the CPU oracle.

Bars: vmult within 1e-12 relative L2, explicit steps within 1e-10, lumped
mass bit-exact, phase state (powder/liquid ratios at quadrature points) and
the set of quadrature points whose powder changed bit-exact.
"""
import numpy as np
import pytest

from adamantine_amd.problem import (BC_ADIABATIC, BC_CONVECTIVE,
                                    FORWARD_EULER, RK_FOURTH_ORDER,
                                    RK_THIRD_ORDER, SOURCE_CUBE, Beam,
                                    Material, Problem, material_from_database,
                                    read_info)
from adamantine_amd.workloads import (plain_large, plain_small,
                                      random_temperature, smooth_temperature)
from oracle.oracle import Oracle
from adamantine_amd.thermal import (LAUNCH_LINE_ATOMIC, LAUNCH_LINE_COLORED,
                                    LAUNCH_PLANE)

pytestmark = pytest.mark.gpu

VMULT_TOL = 1e-12
STEP_TOL = 1e-10


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def gpu_vmult(op, T, t):
    torch = dev()
    op.set_time(t)
    dst = op.initialize_dof_vector()
    op.vmult(dst, torch.tensor(T, device="cuda"))
    torch.cuda.synchronize()
    return dst.cpu().numpy()


def check_vmult(p, T, times=(0.0, 2e-4)):
    from adamantine_amd.thermal import ThermalOperatorDevice
    op = ThermalOperatorDevice(p)
    o = Oracle(p)
    for t in times:
        o.set_time(t)
        got = gpu_vmult(op, T, t)
        want = o.vmult(T)
        assert rel(got, want) < VMULT_TOL, (t, rel(got, want))
    lq, pw = op.qp_state()
    olq, opw = o.qp_state()
    np.testing.assert_array_equal(lq, olq)
    np.testing.assert_array_equal(pw, opw)
    return op, o


def test_inverse_mass_bit_exact():
    from adamantine_amd.thermal import ThermalOperatorDevice
    for p in (plain_small(), plain_large(n_cells=(12, 6, 8))):
        p.set_material_height(p.extent[2] * 0.74)  # inactive top layers
        op = ThermalOperatorDevice(p)
        got = op.get_inverse_mass_matrix().cpu().numpy()
        np.testing.assert_array_equal(got, Oracle(p).inverse_mass())


def test_vmult_plain_small():
    p = plain_small()
    check_vmult(p, smooth_temperature(p), times=(0.0, 5e-4, 1e-3))
    check_vmult(p, random_temperature(p), times=(5e-4,))


def test_vmult_slp_powder_transitions():
    """Random temperatures across solidus/liquidus: the powder and liquid
    ratios at every quadrature point (cells and boundary faces) match the
    oracle bit for bit after one and after two applications."""
    from adamantine_amd.thermal import ThermalOperatorDevice
    p = plain_large(n_cells=(16, 8, 8))
    op = ThermalOperatorDevice(p)
    o = Oracle(p)
    for k, t in enumerate((0.0, 2e-4)):
        T = random_temperature(p, seed=11 + k, low=1200.0, high=2000.0)
        o.set_time(t)
        assert rel(gpu_vmult(op, T, t), o.vmult(T)) < VMULT_TOL
        lq, pw = op.qp_state()
        olq, opw = o.qp_state()
        np.testing.assert_array_equal(pw, opw)
        np.testing.assert_array_equal(lq, olq)
        assert (opw < 1.0).sum() > 0 and (opw == 1.0).sum() > 0
    import ctypes as C
    from adamantine_amd import thermal as TH
    fp = np.zeros(op.n_cells * 6 * op.n_face_q)
    TH.check(TH.load_library().adm_get_face_state(
        op.handle, fp.ctypes.data_as(C.POINTER(C.c_double))))
    # the faces with a boundary term: the top face (id 5) of the top layer
    top = np.arange(op.n_cells) >= op.n_cells - p.n_cells[0] * p.n_cells[1]
    got_fp = fp.reshape(op.n_cells, 6, -1)
    np.testing.assert_array_equal(got_fp[top, 5], o.face_powder()[top, 5])
    assert (got_fp[top, 5] < 1.0).any() and (got_fp[top, 5] == 1.0).any()
    assert not got_fp[~top].any() and not got_fp[top][:, :5].any()


def test_vmult_full_size_plain_large():
    """One application on the full 2,154,945-node workload."""
    from oracle import oracle as orc
    import os
    orc.set_threads(os.cpu_count() or 1)
    p = plain_large()
    check_vmult(p, random_temperature(p, low=300.0, high=2000.0),
                times=(1e-4,))


def test_vmult_add_and_constant():
    """vmult_add == vmult + dst; a constant field with sources off and
    adiabatic boundaries maps to ~0 (tests/test_thermal_operator.cc:41-146)."""
    torch = dev()
    from adamantine_amd.thermal import ThermalOperatorDevice
    p = plain_small()
    op = ThermalOperatorDevice(p)
    op.set_time(0.0)
    src = torch.full((p.n_dofs,), 500.0, dtype=torch.float64, device="cuda")
    d1 = op.initialize_dof_vector()
    op.vmult(d1, src)
    d2 = op.initialize_dof_vector(1.0)
    op.vmult_add(d2, src)
    torch.cuda.synchronize()
    a, b = d1.cpu().numpy(), d2.cpu().numpy()
    assert np.abs(a).max() < 1e-12 * 500.0
    assert np.allclose(b, a + 1.0, rtol=0, atol=1e-15)


def test_multi_material_polynomial(golden):
    """Two materials, state-major polynomial sums (n_materials = -1 layout)
    with the reference's polynomial fixture."""
    db = read_info(golden("material_property_polynomial.info"))
    m = material_from_database(db["materials"], n_states=3, p_order=4)
    for mid in (0, 1):  # the fixture has no specific heat
        for s in ("solid", "liquid", "powder"):
            m.set_poly(mid, "specific_heat", s, [2.0, 0.5])
    p = Problem(n_cells=(6, 4, 5), extent=(1.2, 0.8, 1.0), fe_degree=2,
                material=m)
    nc = p.n_cells_total
    cid = np.arange(nc)
    p.cell_material = (cid % 2).astype(np.uint8)
    st = np.zeros((nc, 3))
    st[:, 0] = 1.0
    powder = (cid % 2 == 1) & (cid % 3 == 0)  # material 1 has a powder state
    st[powder, 0] = 0.0
    st[powder, 2] = 1.0
    p.cell_state = st
    T = random_temperature(p, low=0.5, high=1.5)
    check_vmult(p, T, times=(0.0,))


def test_deposition_rotation():
    p = plain_large(n_cells=(8, 6, 4))
    for k in ("thermal_conductivity_y",):
        for s in ("solid", "liquid", "powder"):
            t = p.material.table[0, 2, 0].copy()
            p.material.set_table(0, k, s, [(a, 0.7 * b) for a, b in t[:8]])
    rng = np.random.RandomState(3)
    th = rng.uniform(0, np.pi, p.n_cells_total)
    p.deposition_cos, p.deposition_sin = np.cos(th), np.sin(th)
    check_vmult(p, random_temperature(p, low=300.0, high=1500.0),
                times=(0.0,))


def run_evolve(p, method, dt, n, T0):
    torch = dev()
    from adamantine_amd.thermal import ThermalPhysicsDevice
    ph = ThermalPhysicsDevice(p, method=method)
    ph.setup()
    sol = torch.tensor(T0, device="cuda")
    t_end = ph.evolve(0.0, dt, n, sol)
    torch.cuda.synchronize()
    return ph, sol.cpu().numpy(), t_end


def test_evolve_plain_small_100_fe_steps():
    p = plain_small()
    T0 = np.full(p.n_dofs, 300.0)
    ph, got, t_end = run_evolve(p, FORWARD_EULER, 1e-6, 100, T0)
    want, t_ref = Oracle(p).evolve(FORWARD_EULER, 0.0, 1e-6, 100, T0)
    assert t_end == t_ref
    assert rel(got, want) < STEP_TOL
    assert got.max() > 1000.0  # the beam melted something


@pytest.mark.parametrize("method", [RK_FOURTH_ORDER, RK_THIRD_ORDER])
def test_evolve_slp_rk(method):
    p = plain_large(n_cells=(24, 12, 8))
    T0 = np.full(p.n_dofs, 300.0)
    n = 60
    ph, got, _ = run_evolve(p, method, 5e-7, n, T0)
    o = Oracle(p)
    want, _ = o.evolve(method, 0.0, 5e-7, n, T0)
    assert rel(got, want) < STEP_TOL
    _, pw = ph.operator.qp_state()
    _, opw = o.qp_state()
    # quadrature points whose powder turned (partly) into liquid/solid
    np.testing.assert_array_equal(pw < 1.0, opw < 1.0)
    np.testing.assert_allclose(pw, opw, rtol=0, atol=1e-10)


def test_convective_gold_gpu(golden):
    """tests/test_thermal_physics.hh:512-624 on the device: 3-D 100x1x1 Q2,
    convective x+ face, forward Euler to t = 1 (gold values, 1e-6)."""
    from tests.test_oracle_physics import convective_problem
    p = convective_problem()
    T0 = np.full(p.n_dofs, 400.0)
    # the reference loop runs while t < 1.0
    t, n = 0.0, 0
    while t < 1.0:
        t += 2.5e-4
        n += 1
    ph, got, _ = run_evolve(p, FORWARD_EULER, 2.5e-4, n, T0)
    gold = np.sort(np.loadtxt(golden("convective_bc_gold.txt")))
    np.testing.assert_allclose(np.sort(got), gold, rtol=1e-6)


def test_energy_balance_gpu():
    """Exact discrete energy balance with a cube source (forward Euler)."""
    from tests.test_oracle_physics import material
    from tests import fe_reference as fe
    rho, cp = 0.5, 4.0
    cube = Beam(type=SOURCE_CUBE, start_time=0.0, end_time=1.0, value=6.0,
                min_point=(0.9,) * 3, max_point=(2.1,) * 3)
    p = Problem(n_cells=(4, 4, 4), extent=(4.0, 4.0, 4.0), fe_degree=2,
                material=material(rho, cp, 2.0, ky=2.0), beams=[cube])
    T0 = np.full(p.n_dofs, 10.0)
    ph, T, t_end = run_evolve(p, FORWARD_EULER, 0.01, 150, T0)
    mass = 1.0 / ph.operator.get_inverse_mass_matrix().cpu().numpy()
    t, n_on = 0.0, 0
    for _ in range(150):
        n_on += (t > 0.0) and (t < 1.0)
        t = t + 0.01
    _, _, W, X = fe.element_tensors(3, 2, [1.0, 1.0, 1.0])
    qint = 0.0
    for cc in fe.cells((4, 4, 4)):
        pts = X + np.array(cc, float)
        inside = np.all((pts >= 0.9) & (pts <= 2.1), axis=1)
        qint += W[inside].sum() * 6.0
    want = (mass * T0).sum() + 0.01 * n_on * qint / (rho * cp)
    assert abs((mass * T).sum() - want) <= 1e-12 * want


def test_state_roundtrip():
    from adamantine_amd.thermal import ThermalOperatorDevice
    p = plain_large(n_cells=(8, 4, 8))
    op = ThermalOperatorDevice(p)
    cs = op.set_state_to_material_properties()
    np.testing.assert_array_equal(cs, p.cell_state)
    st = p.cell_state.copy()
    st[:, 1] = 0.25
    st[:, 0] = 1.0 - 0.25 - st[:, 2]
    op.get_state_from_material_properties(st)
    np.testing.assert_allclose(op.set_state_to_material_properties(), st,
                               rtol=0, atol=1e-15)


def test_invalid_arguments():
    torch = dev()
    from adamantine_amd.thermal import AdmError, ThermalOperatorDevice
    op = ThermalOperatorDevice(plain_small())
    bad = torch.zeros(op.m(), dtype=torch.float32, device="cuda")
    with pytest.raises(AdmError):
        op.vmult(bad, bad)


def test_degree_three():
    """FE degree 3 (16 lanes per cell, 4 cells per wavefront): lumped mass
    bit-exact, vmult and phase state against the oracle on a SLP table
    problem with powder, rotation and radiative/convective faces, and an RK4
    step."""
    torch = dev()
    from adamantine_amd.thermal import ThermalOperatorDevice, ThermalPhysicsDevice
    p = plain_large(n_cells=(9, 5, 6))
    p.fe_degree = 3
    ang = np.random.RandomState(5).uniform(0, np.pi, p.n_cells_total)
    p.deposition_cos, p.deposition_sin = np.cos(ang), np.sin(ang)
    op = ThermalOperatorDevice(p)
    np.testing.assert_array_equal(op.get_inverse_mass_matrix().cpu().numpy(),
                                  Oracle(p).inverse_mass())
    o = Oracle(p)
    for k, t in enumerate((0.0, 1e-4)):
        T = random_temperature(p, seed=21 + k, low=1200.0, high=2000.0)
        o.set_time(t)
        assert rel(gpu_vmult(op, T, t), o.vmult(T)) < VMULT_TOL
        lq, pw = op.qp_state()
        olq, opw = o.qp_state()
        np.testing.assert_array_equal(pw, opw)
        np.testing.assert_array_equal(lq, olq)
    ph = ThermalPhysicsDevice(p, RK_FOURTH_ORDER)
    ph.setup()
    T0 = smooth_temperature(p)
    sol = torch.tensor(T0, device="cuda")
    ph.evolve(0.0, 1e-7, 3, sol)
    want, _ = Oracle(p).evolve(RK_FOURTH_ORDER, 0.0, 1e-7, 3, T0.copy())
    torch.cuda.synchronize()
    assert rel(sol.cpu().numpy(), want) < STEP_TOL


def _five_axis_problem(golden, fe_degree=2):
    """plain_small's material on [-1, 1]^3 (m) with a Goldak beam along the
    reference's scan_path_arcs.txt, whose rotation changes along the path;
    the source is evaluated at R(t) x, so it sits at R(t)^T c(t)."""
    import os
    from adamantine_amd.problem import read_scan_path
    p = plain_small(n_cells=(14, 14, 14))
    p.extent = (2.0, 2.0, 2.0)
    p.origin = (-1.0, -1.0, -1.0)
    p.fe_degree = fe_degree
    b = p.beams[0]
    b.segments = read_scan_path(golden("scan_path_arcs.txt"), "event_series")
    b.radius, b.depth, b.max_power = 0.15, 0.3, 2e7
    return p


def test_five_axis_beam_state(golden):
    """the library's host scan path (position, power, slerped rotation)
    equals the oracle's bit for bit along a path whose rotation changes"""
    from adamantine_amd.thermal import ThermalOperatorDevice
    from oracle import oracle as orc
    p = _five_axis_problem(golden)
    op = ThermalOperatorDevice(p)
    d, keep = p.to_c()
    seg = 0
    for t in np.linspace(0.0, 3.2, 67):
        got = op.beam_state_at(0, t)
        want = orc.beam_state(d.beams[0], t, seg)
        for name, _ in type(got)._fields_:
            a, b = getattr(got, name), getattr(want, name)
            if hasattr(a, "__len__"):
                assert list(a) == list(b), (t, name, list(a), list(b))
            else:
                assert a == b, (t, name, a, b)
        if t <= 3.0:
            assert got.rotated == 1


def test_five_axis_vmult_and_evolve(golden):
    """vmult with the rotated source (and the culling of cells around the
    rotated support) and RK4 steps against the oracle"""
    torch = dev()
    from adamantine_amd.thermal import ThermalOperatorDevice, ThermalPhysicsDevice
    p = _five_axis_problem(golden)
    op = ThermalOperatorDevice(p)
    o = Oracle(p)
    T = random_temperature(p, seed=4, low=300.0, high=900.0)
    for t in (0.5, 1.5, 2.2, 2.5, 2.9):
        o.set_time(t)
        got = gpu_vmult(op, T, t)
        want = o.vmult(T)
        assert rel(got, want) < VMULT_TOL, (t, rel(got, want))
        # the source contributes: the result moves with the beam
        o2 = Oracle(p)
        o2.set_time(5.0)   # after the path: no source
        assert rel(want, o2.vmult(T)) > 1e-6
    ph = ThermalPhysicsDevice(p, RK_FOURTH_ORDER)
    ph.setup()
    T0 = np.full(p.n_dofs, 300.0)
    sol = torch.tensor(T0, device="cuda")
    ph.evolve(1.4, 1e-2, 5, sol)
    want, _ = Oracle(p).evolve(RK_FOURTH_ORDER, 1.4, 1e-2, 5, T0.copy())
    torch.cuda.synchronize()
    got = sol.cpu().numpy()
    assert rel(got, want) < STEP_TOL
    assert got.max() > 301.0


@pytest.mark.parametrize("p,nc", [(4, (5, 4, 3)), (5, (4, 3, 3))])
def test_degree_four_and_five(p, nc):
    """FE degrees 4 and 5 (25 and 36 lanes per cell, 2 and 1 cells per
    wavefront): lumped mass bit-exact, vmult and phase state against the
    oracle on the SLP table problem with powder, rotation and
    radiative/convective faces, and RK4 steps."""
    torch = dev()
    from adamantine_amd.thermal import ThermalOperatorDevice, ThermalPhysicsDevice
    prob = plain_large(n_cells=nc)
    prob.fe_degree = p
    ang = np.random.RandomState(p).uniform(0, np.pi, prob.n_cells_total)
    prob.deposition_cos, prob.deposition_sin = np.cos(ang), np.sin(ang)
    op = ThermalOperatorDevice(prob)
    np.testing.assert_array_equal(op.get_inverse_mass_matrix().cpu().numpy(),
                                  Oracle(prob).inverse_mass())
    o = Oracle(prob)
    for k, t in enumerate((0.0, 1e-4)):
        T = random_temperature(prob, seed=31 + k, low=1200.0, high=2000.0)
        o.set_time(t)
        assert rel(gpu_vmult(op, T, t), o.vmult(T)) < VMULT_TOL
        lq, pw = op.qp_state()
        olq, opw = o.qp_state()
        np.testing.assert_array_equal(pw, opw)
        np.testing.assert_array_equal(lq, olq)
    ph = ThermalPhysicsDevice(prob, RK_FOURTH_ORDER)
    ph.setup()
    T0 = smooth_temperature(prob)
    sol = torch.tensor(T0, device="cuda")
    ph.evolve(0.0, 1e-7, 3, sol)
    want, _ = Oracle(prob).evolve(RK_FOURTH_ORDER, 0.0, 1e-7, 3, T0.copy())
    torch.cuda.synchronize()
    assert rel(sol.cpu().numpy(), want) < STEP_TOL


def test_launch_configs_and_autotune():
    """Every launch configuration of the cell kernel (the line kernel as 8
    colour launches or one atomic launch, the plane kernel; 2, 3 or 4
    wavefronts per block) meets the vmult and phase-state bars against the
    oracle; the autotuner picks one of them and leaves the phase state as it
    found it."""
    torch = dev()
    from adamantine_amd.thermal import ThermalOperatorDevice
    p = plain_large(n_cells=(24, 12, 8))
    T = random_temperature(p, seed=8, low=1200.0, high=2000.0)
    t = 1e-4
    o = Oracle(p)
    o.set_time(t)
    want = o.vmult(T)
    olq, opw = o.qp_state()
    for col in (LAUNCH_LINE_COLORED, LAUNCH_LINE_ATOMIC, LAUNCH_PLANE):
        for w in (2, 3, 4):
            op = ThermalOperatorDevice(p)
            op.set_launch_config(col, w)
            got = gpu_vmult(op, T, t)
            assert rel(got, want) < VMULT_TOL, (col, w, rel(got, want))
            lq, pw = op.qp_state()
            np.testing.assert_array_equal(pw, opw)
            np.testing.assert_array_equal(lq, olq)
    op = ThermalOperatorDevice(p)
    op.set_time(t)
    lq0, pw0 = op.qp_state()
    res = op.autotune(torch.tensor(T, device="cuda"), reps=3)
    print("  autotune:", res, flush=True)
    assert res["waves_per_block"] in (2, 3, 4)
    assert all(v > 0 for v in res["ms"].values())
    lq1, pw1 = op.qp_state()
    np.testing.assert_array_equal(lq0, lq1)
    np.testing.assert_array_equal(pw0, pw1)
    assert rel(gpu_vmult(op, T, t), want) < VMULT_TOL


def test_evolve_graph_replays():
    """adm_evolve with use_graph: one step captured as a HIP graph and
    replayed, the beams read from a per-stage table. Against the direct
    loop: forward Euler on the adiabatic workload bit for bit (colour
    launches, no face term); RK4 with radiative/convective faces on the
    plane path (fused stage updates, no atomics) bit for bit as well;
    three graph calls in a row, with an eager vmult between them, each
    checked against the direct twin."""
    torch = dev()
    from adamantine_amd.thermal import ThermalPhysicsDevice
    p = plain_small()
    a = ThermalPhysicsDevice(p, FORWARD_EULER)
    b = ThermalPhysicsDevice(p, FORWARD_EULER)
    a.setup()
    b.setup()
    # fixed accumulation order for the bit-for-bit comparison
    a.operator.set_launch_config(LAUNCH_LINE_COLORED, 4)
    b.operator.set_launch_config(LAUNCH_LINE_COLORED, 4)
    sa = torch.full((p.n_dofs,), 300.0, dtype=torch.float64, device="cuda")
    sb = sa.clone()
    ta = a.evolve(0.0, 1e-6, 60, sa)
    tb = b.evolve(0.0, 1e-6, 60, sb, use_graph=True)
    torch.cuda.synchronize()
    assert ta == tb
    assert sa.max().item() > 1000.0
    np.testing.assert_array_equal(sa.cpu().numpy(), sb.cpu().numpy())
    q = plain_large(n_cells=(24, 12, 8))
    a = ThermalPhysicsDevice(q, RK_FOURTH_ORDER)
    b = ThermalPhysicsDevice(q, RK_FOURTH_ORDER)
    a.setup()
    b.setup()
    T0 = smooth_temperature(q)
    sa = torch.tensor(T0, device="cuda")
    sb = sa.clone()
    t = 0.0
    for rep in range(3):
        ta = a.evolve(t, 5e-7, 20, sa)
        tb = b.evolve(t, 5e-7, 20, sb, use_graph=True)
        torch.cuda.synchronize()
        assert ta == tb
        got, want = sb.cpu().numpy(), sa.cpu().numpy()
        np.testing.assert_array_equal(got, want)
        # eager work between the replays
        dst = b.operator.initialize_dof_vector()
        b.operator.vmult(dst, sb)
        a.operator.vmult(dst, sa)
        torch.cuda.synchronize()
        t = ta
    la, pa = a.operator.qp_state()
    lb, pb = b.operator.qp_state()
    np.testing.assert_array_equal(pa < 1.0, pb < 1.0)
