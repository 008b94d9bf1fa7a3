"""Synthetic DM-trial time series for the search's workloads.

* ``plain_series``      -- N(0,1) noise plus one accelerated sinusoid
                          (2^20 samples, r0 = 50000.3, z0 = +24, ~15 sigma).
* ``production_series`` -- N(0,1) noise + red noise (PSD 1 + 10 (f/0.1 Hz)^-2 below
                          5 Hz) + accelerated pulsars with 5 %-duty Gaussian pulses.

Signals use the phase phi(u) = (r0 - z0/2) u + z0 u^2 / 2 cycles, u = t/T, so r0 is
the mean Fourier frequency over the observation and z0 the drift in bins.  Series
are generated in FP64 and cast to float32.
"""
from __future__ import annotations

import numpy as np

DT = 6.4e-5


def chirp_phase(n: int, r0: float, z0: float) -> np.ndarray:
    u = np.arange(n, dtype=np.float64) / n
    return (r0 - 0.5 * z0) * u + 0.5 * z0 * u * u


def plain_series(n: int = 1 << 20, seed: int = 1, r0: float = 50000.3, z0: float = 24.0,
                 power: float = 120.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n)
    amp = np.sqrt(4.0 * power / n)   # normalised Fourier power of a sinusoid = A^2 n / 4
    x += amp * np.cos(2.0 * np.pi * chirp_phase(n, r0, z0) + 0.3)
    return x.astype(np.float32)


def red_noise(n: int, rng, dt: float = DT, amp: float = 10.0, f0: float = 0.1,
              fmax: float = 5.0) -> np.ndarray:
    """Gaussian noise with PSD amp*(f/f0)^-2 relative to unit white noise, f < fmax."""
    nf = n // 2 + 1
    f = np.arange(nf) / (n * dt)
    ratio = np.zeros(nf)
    sel = (f > 0) & (f < fmax)
    ratio[sel] = amp * (f[sel] / f0) ** -2
    spec = np.sqrt(ratio * n / 2.0) * (rng.standard_normal(nf) + 1j * rng.standard_normal(nf))
    return np.fft.irfft(spec, n)


def pulse_train(n: int, r0: float, z0: float, duty: float, power1: float) -> np.ndarray:
    """Gaussian pulses (FWHM = duty of the period) whose first harmonic has
    normalised power ``power1``; mean removed."""
    phi = chirp_phase(n, r0, z0)
    ph = phi - np.floor(phi)
    sigma = duty / (2.0 * np.sqrt(2.0 * np.log(2.0)))
    d = np.minimum(ph, 1.0 - ph)
    g = np.exp(-0.5 * (d / sigma) ** 2)
    # first Fourier coefficient of the periodic Gaussian: 2 sqrt(2 pi) sigma exp(-2 pi^2 sigma^2)
    a1 = 2.0 * np.sqrt(2.0 * np.pi) * sigma * np.exp(-2.0 * np.pi ** 2 * sigma ** 2)
    amp = np.sqrt(4.0 * power1 / n) / a1
    g = amp * (g - g.mean())
    return g


PRODUCTION_PULSARS = (
    # (r0 of the first harmonic, z0 of the first harmonic, power of the first harmonic)
    (125000.25, 20.0, 12.0),
    (260000.75, -18.0, 10.0),
    (470000.5, 6.0, 14.0),
)


def production_series(n: int = 1 << 23, seed: int = 2, pulsars=PRODUCTION_PULSARS,
                      dt: float = DT, duty: float = 0.05) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n)
    x += red_noise(n, rng, dt)
    for r0, z0, p1 in pulsars:
        x += pulse_train(n, r0, z0, duty, p1)
    return x.astype(np.float32)


def steps_covering(r: float, rlo: int, uselen: int) -> int:
    """Index of the step whose fundamental columns contain Fourier frequency r."""
    return int((r - rlo) // (uselen // 2))


def log_uniform_lengths(count: int, lo: int, hi: int, seed: int = 3) -> list:
    """Even series lengths drawn log-uniformly in [lo, hi]."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(np.log2(lo), np.log2(hi), count)
    return [int(round(2.0 ** v / 2.0)) * 2 for v in u]


def noise_pool(n: int, seed: int = 5, dt: float = DT) -> np.ndarray:
    """One long N(0,1) + red-noise series; requests take windows of it."""
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(n) + red_noise(n, rng, dt)).astype(np.float32)


def serving_stream(count: int, rate: float, max_len: int = 1 << 25, min_len: int = 1 << 20,
                   seed: int = 6, pulsar_frac: float = 0.05, precisions=("fp32", "i16")):
    """Request descriptions of a Poisson stream: arrival time (s), length, zmax,
    numharm, storage precision (FP32 or the reduced mode within the 1e-3 power
    tolerance) and whether a pulsar is injected."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    lens = log_uniform_lengths(count, min_len, max_len, seed + 1)
    out = []
    for i in range(count):
        out.append({"t": float(t[i]), "n": lens[i], "zmax": int(rng.choice([50, 200, 600])),
                    "numharm": int(rng.choice([4, 8, 16])),
                    "precision": str(rng.choice(list(precisions))),
                    "pulsar": bool(rng.uniform() < pulsar_frac),
                    "offset": int(rng.integers(0, max_len - lens[i] + 1))})
    return out


def add_chirp(x: np.ndarray, r0: float, z0: float, power: float) -> np.ndarray:
    """x plus one accelerated sinusoid of normalised Fourier power ``power``."""
    n = len(x)
    amp = np.sqrt(4.0 * power / n)
    return (x + amp * np.cos(2.0 * np.pi * chirp_phase(n, r0, z0))).astype(np.float32)


def batch_workload(count: int, lo: int = 1 << 21, hi: int = 1 << 23, seed: int = 4,
                   pulsar_frac: float = 0.1):
    """Series of a DM-trial batch: lengths log-uniform in [lo, hi], windows of one
    N(0,1) + red-noise series, a fraction carrying an accelerated sinusoid."""
    rng = np.random.default_rng(seed)
    lengths = log_uniform_lengths(count, lo, hi, seed)
    pool = noise_pool(hi, seed=seed + 1)
    xs = []
    for n in lengths:
        x = pool[:n]
        if rng.uniform() < pulsar_frac:
            x = add_chirp(x, rng.uniform(2e4, 0.4 * n / 2), rng.uniform(-150, 150), 60.0)
        xs.append(x)
    return lengths, xs
