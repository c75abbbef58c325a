"""Batched ensemble on the device against the oracle, member by member.

Each member is checked against an oracle built from that member's own
problem (its absorption efficiency, scaled conductivities, active layers and
powder), i.e. against what the reference computes with one ThermalPhysics
per member. Bars as for one operator: vmult 1e-12, steps 1e-10, lumped mass
and phase state bit-exact."""
import numpy as np
import pytest

from adamantine_amd.problem import RK_FOURTH_ORDER
from adamantine_amd.workloads import ensemble_workload
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

VMULT_TOL = 1e-12
STEP_TOL = 1e-10


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def small(n_members=5):
    return ensemble_workload(n_members=n_members, n_cells=(16, 8, 8))


def test_ensemble_inverse_mass_bit_exact():
    from adamantine_amd.ensemble import EnsembleDevice, member_problem
    base, members = small()
    ens = EnsembleDevice(base, members)
    got = ens.get_inverse_mass_matrix().cpu().numpy()
    for j, m in enumerate(members):
        np.testing.assert_array_equal(
            got[j], Oracle(member_problem(base, m)).inverse_mass())


def test_ensemble_vmult_per_member():
    import torch
    from adamantine_amd.ensemble import EnsembleDevice, member_problem
    base, members = small()
    print("ensemble vmult: 5 members, 16x8x8 Q2", flush=True)
    ens = EnsembleDevice(base, members)
    T = np.stack([np.random.RandomState(20 + j).uniform(300, 2000, ens.n_dofs)
                  for j in range(len(members))])
    oracles = [Oracle(member_problem(base, m)) for m in members]
    for t in (0.0, 2e-4):
        ens.set_time(t)
        dst = ens.initialize_dof_vector()
        ens.vmult(dst, torch.tensor(T, device="cuda"))
        torch.cuda.synchronize()
        got = dst.cpu().numpy()
        for j, o in enumerate(oracles):
            o.set_time(t)
            assert rel(got[j], o.vmult(T[j])) < VMULT_TOL, (t, j)
    for j, o in enumerate(oracles):
        lq, pw = ens.qp_state(j)
        olq, opw = o.qp_state()
        np.testing.assert_array_equal(pw, opw)
        np.testing.assert_array_equal(lq, olq)


def test_ensemble_evolve_per_member_and_member_loop():
    """40 RK4 steps with the beam on: each member against its oracle, and
    the batched ensemble against the serial loop of single-member solvers
    (the reference's run_ensemble schedule) on the device."""
    import torch
    from adamantine_amd.ensemble import EnsembleDevice, member_problem
    from adamantine_amd.thermal import ThermalPhysicsDevice
    base, members = small()
    n, dt = 40, 5e-7
    ens = EnsembleDevice(base, members, method=RK_FOURTH_ORDER)
    ens.compute_inverse_mass_matrix()
    sol = ens.initialize_dof_vector(300.0)
    t_end = ens.evolve(0.0, dt, n, sol)
    torch.cuda.synchronize()
    got = sol.cpu().numpy()
    assert got.max() > 310.0  # the source heated some members
    for j, m in enumerate(members):
        p = member_problem(base, m)
        o = Oracle(p)
        want, t_ref = o.evolve(RK_FOURTH_ORDER, 0.0, dt, n,
                               np.full(ens.n_dofs, 300.0))
        assert t_end == t_ref
        assert rel(got[j], want) < STEP_TOL, j
        _, pw = ens.qp_state(j)
        _, opw = o.qp_state()
        np.testing.assert_array_equal(pw < 1.0, opw < 1.0)
        ph = ThermalPhysicsDevice(p, method=RK_FOURTH_ORDER)
        ph.setup()
        s1 = ph.initialize_dof_vector(300.0)
        ph.evolve(0.0, dt, n, s1)
        torch.cuda.synchronize()
        assert rel(got[j], s1.cpu().numpy()) < 1e-13, j


def test_ensemble_rejects_bad_vectors():
    import torch
    from adamantine_amd.ensemble import EnsembleDevice
    base, members = small(2)
    ens = EnsembleDevice(base, members)
    bad = torch.zeros(ens.n_dofs, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        ens.vmult(bad, bad)


def test_ensemble_launch_configs_and_autotune():
    """The batched launch under each launch configuration meets the vmult
    and phase-state bars of every member against its oracle; the ensemble
    autotuner picks one and leaves every member's phase state as it was."""
    import torch
    from adamantine_amd.ensemble import EnsembleDevice, member_problem
    base, members = small(3)
    print("ensemble launch configurations: 3 members, 16x8x8 Q2", flush=True)
    T = np.stack([np.random.RandomState(40 + j).uniform(300, 2000,
                                                        base.n_dofs)
                  for j in range(len(members))])
    t = 2e-4
    want, wq = [], []
    for m in members:
        o = Oracle(member_problem(base, m))
        o.set_time(t)
        want.append(o.vmult(T[len(want)]))
        wq.append(o.qp_state())
    from adamantine_amd.thermal import (LAUNCH_LINE_ATOMIC,
                                        LAUNCH_LINE_COLORED, LAUNCH_PLANE)
    for col in (LAUNCH_LINE_COLORED, LAUNCH_LINE_ATOMIC, LAUNCH_PLANE):
        for w in (2, 3, 4):
            ens = EnsembleDevice(base, members)
            ens.set_launch_config(col, w)
            ens.set_time(t)
            dst = ens.initialize_dof_vector()
            ens.vmult(dst, torch.tensor(T, device="cuda"))
            torch.cuda.synchronize()
            got = dst.cpu().numpy()
            for j in range(len(members)):
                assert rel(got[j], want[j]) < VMULT_TOL, (col, w, j)
                lq, pw = ens.qp_state(j)
                np.testing.assert_array_equal(pw, wq[j][1])
                np.testing.assert_array_equal(lq, wq[j][0])
    ens = EnsembleDevice(base, members)
    ens.set_time(t)
    before = [ens.qp_state(j) for j in range(len(members))]
    res = ens.autotune(torch.tensor(T, device="cuda"), reps=2)
    print("  autotune:", res, flush=True)
    assert res["waves_per_block"] in (2, 3, 4)
    for j in range(len(members)):
        lq, pw = ens.qp_state(j)
        np.testing.assert_array_equal(lq, before[j][0])
        np.testing.assert_array_equal(pw, before[j][1])


@pytest.mark.parametrize("method", ["fe", "rk4"])
def test_ensemble_graph_replays(method):
    """adm_ensemble_evolve_graph: one batched step captured as a HIP graph
    and replayed, the beams and every member's power modifier read from
    per-stage tables. Against the direct batched loop (plane path: no
    atomics) bit for bit, with the beam on, over three graph calls with
    eager work between them; and on the line kernel's colour launches."""
    import torch
    from adamantine_amd.ensemble import EnsembleDevice
    from adamantine_amd.problem import FORWARD_EULER
    from adamantine_amd.thermal import LAUNCH_LINE_COLORED
    m = FORWARD_EULER if method == "fe" else RK_FOURTH_ORDER
    base, members = small()
    a = EnsembleDevice(base, members, method=m)
    b = EnsembleDevice(base, members, method=m)
    a.compute_inverse_mass_matrix()
    b.compute_inverse_mass_matrix()
    T0 = np.stack([np.random.RandomState(60 + j).uniform(300, 1900, a.n_dofs)
                   for j in range(len(members))])
    sa = torch.tensor(T0, device="cuda")
    sb = sa.clone()
    t = 1e-4
    for rep in range(3):
        ta = a.evolve(t, 5e-7, 12, sa)
        tb = b.evolve(t, 5e-7, 12, sb, use_graph=True)
        torch.cuda.synchronize()
        assert ta == tb
        np.testing.assert_array_equal(sb.cpu().numpy(), sa.cpu().numpy())
        # eager work between the replays
        dst = a.initialize_dof_vector()
        a.vmult(dst, sa)
        b.vmult(dst, sb)
        torch.cuda.synchronize()
        t = ta
    assert np.abs(sa.cpu().numpy() - T0).max() > 1.0
    for e in (a, b):
        e.set_launch_config(LAUNCH_LINE_COLORED, 4)
    ta = a.evolve(t, 5e-7, 6, sa)
    tb = b.evolve(t, 5e-7, 6, sb, use_graph=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(sb.cpu().numpy(), sa.cpu().numpy())
