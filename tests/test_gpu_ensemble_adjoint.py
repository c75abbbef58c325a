"""Every member's gradient from one batched adjoint run
(adm_ensemble_adjoint_gradient), with material activations inside the
window (source/ThermalPhysics.templates.hh:595-825 transposed): each
member's misfit and its gradient with respect to its absorption efficiency,
its conductivity factor, the beam radius and its initial temperature
against central differences of the oracle, which replays the same steps
and activations (cells_to_activate + add_material before the given steps).

The initial field straddles the solidus and liquidus, so the carried powder
ratio varies from point to point inside cells when the activations average
it. The latent heat switches 1/(rho c_p) by a jump there, so J jumps where a
point crosses the solidus or liquidus: a difference step of 1e-4 K on the
initial field put one such crossing inside the interval on one member (a
quotient of 40.3 against a derivative of 3.5). The steps here are 1e-6
(relative for the parameters, K for T0), which makes such a crossing
unlikely and keeps the quotients' rounding near 1e-8; the bar is 1e-5."""
import numpy as np
import pytest

from adamantine_amd.problem import FORWARD_EULER, RK_FOURTH_ORDER
from oracle.oracle import Oracle

from tests.test_gpu_adjoint import hot_field, rel, small_problem

pytestmark = pytest.mark.gpu

TOL = 1e-5


def setup(method):
    """10x6x5 Q2 box, top layer inactive at the start (powder in the top
    two layers and in the inactive one), three members with their own
    absorption efficiency and conductivity factor; member 0 deposits the
    top layer before step 3, member 2 half of it before step 5 and the rest
    before step 6, member 1 nothing."""
    from adamantine_amd.ensemble import Member
    p = small_problem("q2_slp", powder=False, method=method, rotate=False)
    hz = p.extent[2] / p.n_cells[2]
    p.set_material_height(4 * hz)
    p.set_powder_layer(p.extent[2], 3 * hz)
    members = [Member(absorption_efficiency=[0.3 + 0.03 * j],
                      max_power=[p.beams[0].max_power],
                      conductivity_scale=1.0 + 0.05 * j) for j in range(3)]
    X, Y, Z = p.extent
    full = np.array([[0.0, 0.0, 4.2 * hz, X, Y, 4.8 * hz]])
    half = np.array([[0.0, 0.0, 4.2 * hz, 0.5 * X, Y, 4.8 * hz]])
    events = [(0, 3, full, [np.cos(0.4)], [np.sin(0.4)], 1550.0),
              (2, 5, half, [1.0], [0.0], 1500.0),
              (2, 6, full, [np.cos(-0.7)], [np.sin(-0.7)], 1600.0)]
    return p, members, events


def agree(adj, fd):
    """The adjoint against a difference quotient: zero where the quotient
    is (member 1 never activates the layer the beam heats), else to TOL."""
    if fd == 0.0:
        return adj == 0.0
    return rel(adj, fd) < TOL


def oracle_run(p, member, j, events, method, t0, dt, n, T0, beam=None):
    from adamantine_amd.ensemble import member_problem
    o = Oracle(member_problem(p, member))
    if beam is not None:
        o.set_beam_parameters(0, *beam)
    T, t = T0.copy(), t0
    for k in range(n):
        for (m, step, boxes, cs, sn, new_t) in events:
            if m == j and step == k:
                b = o.cells_to_activate(boxes)
                assert (b >= 0).sum() > 0
                o.add_material(b, np.asarray(cs), np.asarray(sn), new_t, T)
        T, t = o.evolve(method, t, dt, 1, T)
    return T


@pytest.mark.parametrize("method", [FORWARD_EULER, RK_FOURTH_ORDER])
def test_ensemble_adjoint_across_activation(method):
    import copy
    import torch
    from adamantine_amd.adjoint import ensemble_adjoint_gradient
    from adamantine_amd.ensemble import EnsembleDevice
    print("ensemble adjoint: 3 members of 10x6x5 Q2, activations, method",
          method, flush=True)
    p, members, events = setup(method)
    t0, dt, n = 0.0, 1e-6, 8
    N = len(members)
    T0 = np.stack([hot_field(p, 1500.0, 1720.0, seed=3 + j)
                   for j in range(N)])
    X = p.node_coordinates()
    top = np.nonzero(np.isclose(X[:, 2], p.extent[2] * 4 / 5))[0]
    dof = top[::5]
    ref = [oracle_run(p, members[j], j, events, method, t0, dt, n, T0[j])
           for j in range(N)]
    obs = np.stack([ref[j][dof] - 2.0 + np.random.RandomState(20 + j)
                    .uniform(-1, 1, dof.size) for j in range(N)])
    ens = EnsembleDevice(p, members, method=method)
    ens.compute_inverse_mass_matrix()
    res = ensemble_adjoint_gradient(ens, t0, dt, n,
                                    torch.tensor(T0, device="cuda"), dof, obs,
                                    events=events, checkpoint_every=3)
    torch.cuda.synchronize()
    # the forward sweep is the ensemble's evolve with the activations
    ens2 = EnsembleDevice(p, members, method=method)
    ens2.compute_inverse_mass_matrix()
    sol = torch.tensor(T0, device="cuda")
    t = t0
    for k in range(n):
        for (m, step, boxes, cs, sn, new_t) in events:
            if step == k:
                ens2.activate(m, boxes, cs, sn, new_t, sol)
        t = ens2.evolve(t, dt, 1, sol)
    torch.cuda.synchronize()
    np.testing.assert_allclose(res.T_final.cpu().numpy(), sol.cpu().numpy(),
                               rtol=1e-13, atol=0)
    # ... and the ensemble is left in its final state (cells, phase state)
    for j in range(N):
        lq, pw = ens.qp_state(j)
        lq2, pw2 = ens2.qp_state(j)
        np.testing.assert_allclose(pw, pw2, rtol=0, atol=1e-12)
        np.testing.assert_allclose(ens.face_state(j), ens2.face_state(j),
                                   rtol=0, atol=1e-12)
    np.testing.assert_array_equal(ens.get_inverse_mass_matrix().cpu().numpy(),
                                  ens2.get_inverse_mass_matrix().cpu().numpy())
    b = p.beams[0]
    for j in range(N):
        r = ref[j][dof] - obs[j]
        Jo = 0.5 * float(np.dot(r, r))
        assert rel(res.J[j], Jo) < 1e-7
        eta = members[j].absorption_efficiency[0]
        base = [eta, b.max_power, b.radius, b.depth]

        def J(beam=None, member=None, T=None):
            Tj = oracle_run(p, member or members[j], j, events, method, t0,
                            dt, n, T0[j] if T is None else T, beam)
            rr = Tj[dof] - obs[j]
            return 0.5 * float(np.dot(rr, rr))
        d = 1e-6 * eta
        fd = (J(beam=[eta + d] + base[1:]) - J(beam=[eta - d] + base[1:])) / \
            (2 * d)
        # the beam rides on the top of the box: it heats only the members
        # that deposited the top layer
        assert (fd != 0.0) == (j != 1)
        print(f"  member {j} eta: fd {fd:.10e} adjoint {res.grad[j, 0]:.10e}",
              flush=True)
        assert agree(res.grad[j, 0], fd)
        d = 1e-6 * b.radius
        fd = (J(beam=base[:2] + [b.radius + d, b.depth]) -
              J(beam=base[:2] + [b.radius - d, b.depth])) / (2 * d)
        print(f"  member {j} radius: fd {fd:.10e} adjoint "
              f"{res.grad[j, 2]:.10e}", flush=True)
        assert agree(res.grad[j, 2], fd)
        e = 1e-6
        up, dn = copy.deepcopy(members[j]), copy.deepcopy(members[j])
        up.conductivity_scale *= 1 + e
        dn.conductivity_scale *= 1 - e
        fd = (J(member=up) - J(member=dn)) / (2 * e)
        print(f"  member {j} k-scale: fd {fd:.10e} adjoint "
              f"{res.grad[j, 1]:.10e}", flush=True)
        assert abs(fd) > 0 and rel(res.grad[j, 1], fd) < TOL
        v = np.random.RandomState(40 + j).uniform(-1, 1, p.n_dofs)
        h = 1e-6
        fd = (J(T=T0[j] + h * v) - J(T=T0[j] - h * v)) / (2 * h)
        adj = float(np.dot(res.lambda0[j].cpu().numpy(), v))
        print(f"  member {j} T0: fd {fd:.10e} adjoint {adj:.10e}", flush=True)
        assert abs(fd) > 0 and rel(adj, fd) < TOL
