"""The oracle's EnKF analysis against the reference's own test values
(tests/test_data_assimilator.cc) and closed forms.

Pins:
  - Kalman-gain forecast shifts, 12 values at 1e-4 (:46-167);
  - sample covariance: zero for identical members, the 16 dense values,
    step-function localisation to the diagonal, Gaspari-Cohn bounds, the
    augmented-parameter rows (:367-572);
  - update_ensemble moves observed values towards the data (:652-770);
  - noise statistics: sample covariance of 1000 draws within 20% (:574-650);
  - Gaspari-Cohn closed values and continuity at r = 1 and 2;
  - one observation, no localisation: K = P h / (h^T P h + r) in closed form.
"""
import ctypes as C

import numpy as np
import pytest

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)


def L():
    from oracle import oracle as orc
    lib = orc.lib()
    lib.oracle_gaspari_cohn.restype = C.c_double
    lib.oracle_gaspari_cohn.argtypes = [C.c_double]
    lib.oracle_sample_covariance.argtypes = [
        C.c_int, C.c_int64, C.c_int64, _dp, _dp, C.c_int, C.c_int, C.c_double,
        _dp]
    lib.oracle_apply_kalman_gain.argtypes = [
        C.c_int, C.c_int64, C.c_int64, _dp, _dp, C.c_int, C.c_int, C.c_double,
        C.c_int, _ip, _dp, _dp, _dp]
    lib.oracle_update_ensemble.argtypes = [
        C.c_int, C.c_int64, C.c_int64, _dp, _dp, C.c_int, C.c_int, C.c_double,
        C.c_int, _ip, _dp, _dp, _dp]
    lib.oracle_da_rng_size.restype = C.c_size_t
    lib.oracle_fill_noise.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp]
    return lib


def p(a):
    return None if a is None else a.ctypes.data_as(
        _ip if a.dtype == np.int64 else _dp)


DBL_MAX = np.finfo(np.float64).max
# support points of the single 2-D Q1 cell [0,1]^2 in deal.II's order
UNIT_SQUARE = np.array([[0., 0.], [1., 0.], [0., 1.], [1., 1.]])
FN = {"none": 0, "step_function": 1, "gaspari_cohn": 2}


def covariance(ens, n_sim, fn="none", cutoff=DBL_MAX, coords=UNIT_SQUARE):
    ens = np.ascontiguousarray(ens, dtype=np.float64)
    N, n = ens.shape
    P = np.zeros((n, n))
    L().oracle_sample_covariance(N, n, n_sim, p(ens), p(np.ascontiguousarray(
        coords, dtype=np.float64)), 2, FN[fn], cutoff, p(P))
    return P


ENS3 = np.array([[1.0, 3.0, 6.0, 9.0], [1.5, 3.2, 6.3, 9.7],
                 [1.1, 3.1, 6.1, 9.1]])
OBS_DOF = np.array([1, 3], dtype=np.int64)
Y = np.array([2.5, 9.5])
R2 = np.diag([0.002, 0.001])


def test_kalman_gain_reference_values():
    """tests/test_data_assimilator.cc:46-167."""
    innov = np.array([[Y[k] - ENS3[j, OBS_DOF[k]] for k in range(2)]
                      for j in range(3)])
    innov += np.array([[0.0008, -0.0005], [-0.001, 0.0002], [0.0002, -0.0009]])
    out = np.zeros((3, 4))
    L().oracle_apply_kalman_gain(3, 4, 4, p(ENS3), p(UNIT_SQUARE), 2, 0,
                                 DBL_MAX, 2, p(OBS_DOF), p(R2), p(innov),
                                 p(out))
    want = [[0.21352564, -0.14600986, -0.02616469, 0.45321598],
            [-0.27786325, -0.32946285, -0.31226298, -0.24346351],
            [0.12767094, -0.20319395, -0.09290565, 0.34824753]]
    # BOOST tt::tolerance(1e-4) is relative
    np.testing.assert_allclose(out, want, rtol=1e-4)


def test_sample_covariance_reference_values():
    """tests/test_data_assimilator.cc:367-572."""
    v = np.array([2.0, 4.0, 5.0, 7.0])
    v1 = np.array([2.1, 4.3, 5.2, 7.4])
    np.testing.assert_allclose(covariance([v, v], 4), 0.0, atol=1e-10)
    dense = np.outer([0.1, 0.3, 0.2, 0.4], [0.1, 0.3, 0.2, 0.4]) / 2.0
    np.testing.assert_allclose(covariance([v, v1], 4, "step_function", 100.0),
                               dense, atol=1e-10)
    np.testing.assert_allclose(
        covariance([v, v1], 4, "step_function", 1e-6), np.diag(np.diag(dense)),
        atol=1e-10)
    gc = covariance([v, v1], 4, "gaspari_cohn", 3.0)
    np.testing.assert_allclose(np.diag(gc), np.diag(dense), atol=1e-10)
    off = ~np.eye(4, dtype=bool)
    assert (gc[off] > 0).all() and (gc[off] < dense[off]).all()
    v2 = np.array([2.0, 4.0, 5.0, 7.0, 1.0, 1.5])
    v3 = np.array([2.1, 4.3, 5.2, 7.4, 1.1, 1.4])
    c4 = covariance([v2, v3], 4, "step_function", 1e-6)
    np.testing.assert_allclose(c4[:4, :4], np.diag(np.diag(dense)), atol=1e-10)
    np.testing.assert_allclose(c4[4, :], [0.005, 0.015, 0.01, 0.02, 0.005,
                                          -0.005], atol=1e-10)
    np.testing.assert_allclose(c4[5, :], [-0.005, -0.015, -0.01, -0.02,
                                          -0.005, 0.005], atol=1e-10)


def test_update_moves_towards_data():
    """tests/test_data_assimilator.cc:652-770 (no noise)."""
    ens = ENS3.copy()
    L().oracle_update_ensemble(3, 4, 4, p(ens), p(UNIT_SQUARE), 2, 0, DBL_MAX,
                               2, p(OBS_DOF), p(Y), p(R2), None)
    for j in range(3):
        for k in range(2):
            assert abs(Y[k] - ens[j, OBS_DOF[k]]) <= abs(Y[k] - ENS3[j, OBS_DOF[k]])


@pytest.mark.parametrize("diagonal", [True, False])
def test_noise_statistics(diagonal):
    """tests/test_data_assimilator.cc:574-650: 1000 draws reproduce R within
    20%."""
    R = (np.diag([0.1, 1.0, 0.2]) if diagonal else
         np.array([[0.1, 0.3, 0.0], [0.3, 1.0, 0.0], [0.0, 0.0, 0.2]]))
    lib = L()
    state = (C.c_char * lib.oracle_da_rng_size())()
    out = np.zeros((1000, 3))
    lib.oracle_fill_noise(state, 1000, 3, p(R), p(out))
    Rt = np.cov(out.T)
    for (a, b) in [(0, 0), (1, 1), (2, 2), (0, 1), (1, 0)]:
        if R[a, b] != 0:
            assert abs(Rt[a, b] - R[a, b]) <= 0.2 * abs(R[a, b])


def test_gaspari_cohn_closed_form():
    gc = L().oracle_gaspari_cohn
    assert gc(0.0) == 1.0
    r1 = 1 - 5 / 3 + 5 / 8 + 0.5 - 0.25
    assert abs(gc(1.0 - 1e-12) - r1) < 1e-10 and abs(gc(1.0) - r1) < 1e-12
    assert abs(gc(2.0 - 1e-9)) < 1e-8 and gc(2.0) == 0.0 and gc(3.0) == 0.0
    # monotone decreasing on [0, 2)
    r = np.linspace(0, 1.999, 400)
    v = np.array([gc(x) for x in r])
    assert (np.diff(v) <= 1e-15).all()


def test_single_observation_closed_form():
    """m = 1, no localisation: K = P h / (h^T P h + r)."""
    rng = np.random.RandomState(4)
    ens = rng.normal(size=(6, 5))
    r = np.array([[0.3]])
    obs = np.array([2], dtype=np.int64)
    innov = rng.normal(size=(6, 1))
    out = np.zeros((6, 5))
    L().oracle_apply_kalman_gain(6, 5, 5, p(ens), None, 2, 0, DBL_MAX, 1,
                                 p(obs), p(r), p(innov), p(out))
    P = np.cov(ens.T)
    K = P[:, 2] / (P[2, 2] + 0.3)
    np.testing.assert_allclose(out, innov * K[None, :], rtol=1e-12, atol=1e-14)
