"""The oracle's transforms against numpy.fft (FP64 pocketfft) and brute force."""
import numpy as np
import pytest


@pytest.mark.parametrize("n", [1, 2, 4, 8, 64, 1024, 4096, 3, 5, 6, 12, 1000, 1155, 7 * 1024, 32774])
def test_fft_vs_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    scale = max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(oracle.fft(x, -1) - ref)) < 1e-12 * scale
    assert np.max(np.abs(oracle.fft(x, +1) - np.fft.ifft(x) * n)) < 1e-12 * scale


@pytest.mark.parametrize("n", [2, 8, 16, 48, 64])
def test_fft_brute_force(oracle, n):
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    dft = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
    assert np.max(np.abs(oracle.fft(x, -1) - dft)) < 1e-12 * np.max(np.abs(dft))


@pytest.mark.parametrize("n", [128, 4096, 1 << 16, 3000])
def test_rfft_vs_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    ref = np.fft.rfft(x)
    got = oracle.rfft(x)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))
