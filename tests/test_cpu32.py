"""The FP32 CPU form of the search (oracle/fdas_cpu32.c), used as the CPU
baseline: its radix-4 Stockham FFT against numpy's FP64 FFT and a brute-force
DFT, and its per-step summed planes against the FP64 oracle."""
import numpy as np
import pytest


@pytest.mark.parametrize("n", [4, 8, 64, 512, 2048, 8192, 1 << 15])
def test_fft32_against_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    ref = np.fft.fft(x.astype(np.complex128))
    got = oracle.fft32(x, -1)
    assert np.max(np.abs(got - ref)) <= 2e-6 * np.log2(n) * np.sqrt(n)
    inv = oracle.fft32(x, +1)
    assert np.max(np.abs(inv - np.fft.ifft(x.astype(np.complex128)) * n)) <= 2e-6 * np.log2(n) * np.sqrt(n)


def test_fft32_brute_force(oracle):
    n = 128
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    k = np.arange(n)
    dft = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x.astype(np.complex128)
    assert np.max(np.abs(oracle.fft32(x, -1) - dft)) <= 1e-4


def test_cpu32_planes_match_oracle(oracle):
    from gpu_accelsearch_amd import synth
    n = 1 << 18
    x = synth.plain_series(n, seed=19, r0=9000.3, z0=6.0, power=70.0)
    c = oracle.Cpu32(x, n, 6.4e-5, 30, 4, 1024)
    ref = oracle.Search(x, n, 6.4e-5, 30, 4, 1024)
    assert (c.uselen, c.nsteps, c.numz) == (ref.uselen, ref.nsteps, ref.numz)
    for st in (0, synth.steps_covering(9000.3, ref.rlo, ref.uselen), ref.nsteps - 1):
        got = c.step_plane(st).astype(np.float64)
        want = ref.step_plane(st)
        assert np.max(np.abs(got - want) / np.maximum(want, 1.0)) <= 1e-4
    wall, hits, tn = c.run_steps(np.arange(c.nsteps), powcut=(20.0, 25.0, 35.0))
    assert wall > 0 and tn > 0 and hits > 0
    c.close()
    ref.close()
