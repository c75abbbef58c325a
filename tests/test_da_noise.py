"""Observation-noise stream of the library's DataAssimilator (host code, no
GPU) against the oracle's restatement (pinned statistically in
test_oracle_enkf.py and by the engine/normal known answers of
test_ensemble_draws.py)."""
import ctypes as C

import numpy as np
import pytest

from adamantine_amd.ensemble import DataAssimilator


def oracle_noise(R, counts):
    from tests.test_oracle_enkf import L, p
    lib = L()
    state = (C.c_char * lib.oracle_da_rng_size())()
    outs = []
    for c in counts:
        out = np.zeros((c, R.shape[0]))
        lib.oracle_fill_noise(state, c, R.shape[0], p(R), p(out))
        outs.append(out)
    return outs


@pytest.mark.parametrize("diagonal", [True, False])
def test_noise_stream_matches_oracle(diagonal):
    R = (np.diag([0.1, 1.0, 0.2, 4.0]) if diagonal else
         np.array([[0.1, 0.3, 0.0, 0.0], [0.3, 1.0, 0.0, 0.1],
                   [0.0, 0.0, 0.2, 0.0], [0.0, 0.1, 0.0, 4.0]]))
    da = DataAssimilator()
    got = [da.fill_noise_vectors(c, R) for c in (3, 5, 1)]  # stream persists
    want = oracle_noise(R, (3, 5, 1))
    for g, w in zip(got, want):
        if diagonal:
            np.testing.assert_array_equal(g, w)
        else:
            np.testing.assert_allclose(g, w, rtol=1e-14, atol=1e-15)


def test_unknown_localization_is_an_error():
    with pytest.raises(ValueError):
        DataAssimilator("bogus")
