"""EnKF analysis of a device ensemble against the oracle's restatement of
DataAssimilator::update_ensemble (pinned to the reference's Kalman-gain,
covariance and update tests in test_oracle_enkf.py).

Bars: the ensemble-space product (FP64 MFMA) against a float64 numpy
product at 1e-13 relative; the analysed states and parameters against the
oracle at 1e-10 relative to the size of the update (the oracle solves in
extended precision, the library in FP64)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DBL_MAX = np.finfo(np.float64).max


def ensemble(n_members, n_cells=(6, 4, 3)):
    from adamantine_amd.ensemble import EnsembleDevice, Member
    from adamantine_amd.workloads import plain_large
    base = plain_large(n_cells=n_cells)
    rng = np.random.RandomState(7)
    members = [Member(absorption_efficiency=[0.3 + 0.01 * j],
                      max_power=[400.0],
                      conductivity_scale=float(rng.uniform(0.9, 1.1)))
               for j in range(n_members)]
    return base, EnsembleDevice(base, members)


@pytest.mark.parametrize("N", [5, 40, 64])
def test_ensemble_transform_matches_float64_product(N):
    import torch
    from adamantine_amd.ensemble import ensemble_transform
    print("transform N =", N, flush=True)
    base, ens = ensemble(N)
    rng = np.random.RandomState(N)
    X = rng.uniform(300, 2000, (N, ens.n_dofs))
    Tm = np.eye(N) + rng.normal(scale=0.1, size=(N, N))
    dev = torch.tensor(X, device="cuda")
    ensemble_transform(ens, dev, Tm)
    torch.cuda.synchronize()
    want = (X.T @ Tm).T
    got = dev.cpu().numpy()
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


def oracle_analysis(base, X, params, obs_dof, y, R, noise, fn, cutoff):
    from tests.test_oracle_enkf import L, p
    N, n = X.shape
    npar = 0 if params is None else params.shape[1]
    aug = X if params is None else np.concatenate([X, params], axis=1)
    aug = np.ascontiguousarray(aug)
    coords = np.ascontiguousarray(base.node_coordinates())
    L().oracle_update_ensemble(N, aug.shape[1], n, p(aug), p(coords), 3, fn,
                               cutoff, len(obs_dof), p(obs_dof), p(y), p(R),
                               p(noise))
    return aug[:, :n], (None if params is None else aug[:, n:])


@pytest.mark.parametrize("fn,cutoff,diag", [
    ("none", DBL_MAX, True),        # ensemble-space (Woodbury) path
    ("none", DBL_MAX, False),       # m x m solve, non-diagonal R
    ("step_function", 0.5e-3, True),
    ("gaspari_cohn", 0.8e-3, True)])
def test_analysis_matches_oracle(fn, cutoff, diag):
    import torch
    from adamantine_amd.ensemble import LOCALIZATION, DataAssimilator
    print("analysis", fn, cutoff, diag, flush=True)
    N = 8
    base, ens = ensemble(N)
    rng = np.random.RandomState(3)
    X = 1000.0 + rng.normal(scale=30.0, size=(N, ens.n_dofs))
    params = np.ascontiguousarray(
        np.stack([0.3 + rng.normal(scale=0.05, size=N),
                  1.0 + rng.normal(scale=0.1, size=N)], axis=1))
    da = DataAssimilator(fn, cutoff)
    pts = np.stack([rng.uniform(0, 4e-3, 12), rng.uniform(0, 2e-3, 12),
                    np.full(12, 1e-3)], axis=1)
    obs = da.update_dof_mapping(ens, pts)
    m = obs.size
    y = 1000.0 + rng.normal(scale=5.0, size=m)
    R = np.diag(rng.uniform(1.0, 4.0, m))
    if not diag:
        R = R + 0.2 * (np.eye(m, k=1) + np.eye(m, k=-1))
    noise = da.fill_noise_vectors(N, R)
    want_X, want_p = oracle_analysis(base, X.copy(), params.copy(), obs, y, R,
                                     noise, LOCALIZATION[fn], cutoff)
    dev = torch.tensor(X, device="cuda")
    got_p = params.copy()
    da.update_ensemble(ens, dev, y, R, params=got_p, noise=noise)
    torch.cuda.synchronize()
    got_X = dev.cpu().numpy()
    scale = np.abs(want_X - X).max()
    assert scale > 1e-3  # the analysis moved the states
    assert np.abs(got_X - want_X).max() <= 1e-10 * scale + 1e-12 * 1000
    pscale = np.abs(want_p - params).max()
    np.testing.assert_allclose(got_p, want_p, rtol=0, atol=1e-10 * pscale)
    # observed values move towards the data on average
    before = np.abs(X[:, obs] - y).mean()
    after = np.abs(got_X[:, obs] - y).mean()
    assert after < before


def test_nearest_dofs_are_support_points():
    from adamantine_amd.ensemble import DataAssimilator
    base, ens = ensemble(2)
    X = base.node_coordinates()
    idx = np.array([0, 17, 123, ens.n_dofs - 1])
    da = DataAssimilator()
    got = da.update_dof_mapping(ens, X[idx] + 1e-9)
    np.testing.assert_array_equal(got, idx)
