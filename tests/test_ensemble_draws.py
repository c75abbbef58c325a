"""Ensemble parameter draws (source/ensemble_management.cc:18-52, 166-265).

Pins: the reference's exact draws (tests/test_ensemble_management.cc:40-66,
keys alpha.beta ~ N(100, 10) then gamma.delta ~ N(200, 20), 6 members, one
rank and two ranks) and the C++ standard's known answer for the engine
(the 10000th output of a default-constructed mt19937 is 4123659995,
[rand.predef]). Both the oracle's restatement and the library's host code
(which calls the standard library itself) must reproduce them bit for bit.
No GPU is used."""
import ctypes as C

import numpy as np

from adamantine_amd.ensemble import draw_member_values, normal_random_vector

# tests/test_ensemble_management.cc:40-66
ALPHA_BETA = [101.34529658472329, 98.536182188102771, 104.60650182383064,
              81.286156895893981, 101.63711684234313, 97.857466117910562]
GAMMA_DELTA = [205.97190519115932, 183.44111635285606, 200.20430868691838,
               221.10932887760762, 189.06317994350883, 223.49135354909728]
KEYS = [("alpha.beta", 100.0, 10.0), ("gamma.delta", 200.0, 20.0)]


def oracle_lib():
    from oracle import oracle as orc
    L = orc.lib()
    L.oracle_normal_random_vector.argtypes = [
        C.c_uint32, C.c_uint32, C.c_double, C.c_double,
        C.POINTER(C.c_double)]
    L.oracle_mt19937.argtypes = [C.c_uint32, C.POINTER(C.c_uint32)]
    return L


def oracle_normal(length, n_rejected, mean, stddev):
    out = np.zeros(length)
    oracle_lib().oracle_normal_random_vector(
        length, n_rejected, mean, stddev,
        out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def test_oracle_engine_known_answer():
    out = (C.c_uint32 * 10000)()
    oracle_lib().oracle_mt19937(10000, out)
    assert out[9999] == 4123659995


def test_oracle_reference_draws():
    np.testing.assert_array_equal(oracle_normal(6, 0, 100.0, 10.0), ALPHA_BETA)
    np.testing.assert_array_equal(oracle_normal(6, 6, 200.0, 20.0), GAMMA_DELTA)


def test_library_reference_draws_one_rank():
    v = draw_member_values(KEYS, 6)
    np.testing.assert_array_equal(v["alpha.beta"], ALPHA_BETA)
    np.testing.assert_array_equal(v["gamma.delta"], GAMMA_DELTA)


def test_library_reference_draws_two_ranks():
    for rank in (0, 1):
        v = draw_member_values(KEYS, 6, first_local_member=3 * rank,
                               local_ensemble_size=3)
        np.testing.assert_array_equal(v["alpha.beta"],
                                      ALPHA_BETA[3 * rank:3 * rank + 3])
        np.testing.assert_array_equal(v["gamma.delta"],
                                      GAMMA_DELTA[3 * rank:3 * rank + 3])


def test_rejection_of_negative_draws():
    """N(0.1, 1) draws: about 46% are redrawn; library == oracle over a long
    stream, and every value is non-negative."""
    a = normal_random_vector(4000, 17, 0.1, 1.0)
    b = oracle_normal(4000, 17, 0.1, 1.0)
    np.testing.assert_array_equal(a, b)
    assert (a >= 0).all()
    # the half-normal shape: mean of |N(0.1,1)| restricted to >= 0 is ~0.85
    assert 0.75 < a.mean() < 0.95


def test_zero_stddev_and_errors():
    from adamantine_amd.thermal import AdmError
    np.testing.assert_array_equal(normal_random_vector(5, 3, 2.5, 0.0),
                                  np.full(5, 2.5))
    import pytest
    with pytest.raises(AdmError):
        normal_random_vector(3, 0, 1.0, -1.0)
