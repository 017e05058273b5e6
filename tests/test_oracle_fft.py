"""Oracle FFT against the DFT definition and numpy.fft golden vectors."""

import os

import numpy as np
import pytest

from tests.conftest import GOLDEN


def brute_dft(x, inverse=False):
    n = len(x)
    j = np.arange(n)
    sign = 1.0 if inverse else -1.0
    w = np.exp(sign * 2j * np.pi * np.outer(j, j) / n)
    return w @ x


@pytest.mark.parametrize("n", list(range(1, 41)) + [49, 63, 64, 67, 121, 126])
def test_fft_matches_brute_force_dft(oracle_mod, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = brute_dft(x)
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(oracle_mod.fft(x) - ref).max() < 1e-13 * scale
    assert np.abs(oracle_mod.fft(x, True) - brute_dft(x, True)).max() < 1e-13 * scale


def test_fft_matches_numpy_fixtures(oracle_mod):
    data = np.load(os.path.join(GOLDEN, "fft_vectors.npz"))
    for key in data.files:
        if not key.startswith("x_"):
            continue
        n = int(key[2:])
        x = data[key]
        scale = max(1.0, np.abs(data[f"fwd_{n}"]).max())
        assert np.abs(oracle_mod.fft(x) - data[f"fwd_{n}"]).max() < 1e-13 * scale, n
        assert np.abs(oracle_mod.fft(x, True) - data[f"inv_{n}"]).max() < 1e-13 * scale, n
    x3 = data["x3"]
    assert np.abs(oracle_mod.fft3(x3) - data["fwd3"]).max() < 1e-12


@pytest.mark.parametrize("n", [6, 12, 16, 97, 128, 194, 1024])
def test_inverse_of_forward_is_n_identity_and_parseval(oracle_mod, n):
    rng = np.random.default_rng(7 * n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = oracle_mod.fft(x)
    np.testing.assert_allclose(oracle_mod.fft(X, True) / n, x, atol=1e-13)
    assert np.sum(np.abs(X) ** 2) == pytest.approx(n * np.sum(np.abs(x) ** 2), rel=1e-12)
