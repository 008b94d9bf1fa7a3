"""Order statistic, fast inverse square root and whitening of the oracle.

Median and inverse square root are pinned bit-exactly by golden vectors from the
reference's own source/bpd_math.c (tests/golden/make_ref_vectors.py).  The
whitening (source/bpd_noise.c, which needs CUDA headers and is not built here) is
pinned by closed forms: constant spectra, power-of-two scale invariance, a
monotone spectrum whose block medians are known in advance (block schedule
derived by hand), and a linear spectrum that whitening must flatten.
"""
import numpy as np
import pytest

LN2 = np.log(2.0)


def test_median_matches_reference(oracle, golden):
    vals, offs, out = golden["median_vals"], golden["median_offs"], golden["median_out"]
    L = oracle.lib()
    for i in range(len(out)):
        a = np.ascontiguousarray(vals[offs[i]:offs[i + 1]])
        got = L.or_lower_median(a, len(a))
        assert np.float32(got).view(np.uint32) == out[i].view(np.uint32), i


def test_rsqrt_matches_reference(oracle, golden):
    L = oracle.lib()
    x, y = golden["rsqrt_in"], golden["rsqrt_out"]
    got = np.array([L.or_quake_rsqrt(float(v)) for v in x], np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), y.view(np.uint32))


def test_next_pow2_matches_reference(oracle, golden):
    L = oracle.lib()
    assert [L.or_next_pow2(int(v)) for v in golden["pow2_in"]] == list(golden["pow2_out"])


def test_rsqrt_error_bound(oracle):
    L = oracle.lib()
    x = np.geomspace(1e-30, 1e30, 20001).astype(np.float32)
    got = np.array([L.or_quake_rsqrt(float(v)) for v in x], np.float64)
    rel = np.abs(got * np.sqrt(x.astype(np.float64)) - 1.0)
    assert rel.max() < 1.8e-3


def _quake(x):
    x = np.float32(x)
    i = np.array([x], np.float32).view(np.int32)[0]
    i = np.int32(0x5F3759DF) - (i >> 1)
    y = np.array([i], np.int32).view(np.float32)[0]
    h = np.float32(np.float32(np.float32(0.5) * x) * y) * y
    return np.float32(y * np.float32(np.float32(1.5) - np.float32(h)))


def test_deredden_constant_spectrum(oracle):
    nb = 20000
    amp = np.float32(3.0)
    spec = np.full(nb, amp + 0j, np.complex64)
    out = oracle.deredden(spec, nb)
    norm = _quake(np.float32(np.float64(np.float32(9.0)) / LN2))
    assert out[0] == 1.0 + 0j
    np.testing.assert_array_equal(out[1:].real, np.float32(amp * norm))
    assert np.all(out[1:].imag == 0)


def test_deredden_scale_invariance(oracle):
    """Scaling amplitudes by 2^k scales medians by 4^k and the fast inverse
    square root by 2^-k exactly, so the whitened spectrum is unchanged."""
    rng = np.random.default_rng(3)
    nb = 50000
    f = np.arange(nb) + 1.0
    spec = (np.sqrt(1 + 30 / f) * (rng.standard_normal(nb) + 1j * rng.standard_normal(nb))).astype(np.complex64)
    a = oracle.deredden(spec, nb)
    for k in (-3, 2, 5):
        b = oracle.deredden((spec * np.float32(2.0 ** k)).astype(np.complex64), nb)
        np.testing.assert_array_equal(a[1:].view(np.uint64), b[1:].view(np.uint64))


def test_deredden_block_schedule_kat(oracle):
    """Amplitude b+10 (power (b+10)^2, exact and increasing): the lower median of
    block [s, s+len) is the power at s + (len-1)//2.  Blocks (start, len), derived
    by hand from len = int(6 ln start): (1,6) (7,11) (18,17) (35,21) (56,24)."""
    nb = 3000
    spec = (np.arange(nb) + 10.0).astype(np.complex64)
    out = oracle.deredden(spec, nb)
    blocks = [(1, 6), (7, 11), (18, 17), (35, 21), (56, 24)]
    med = [np.float32((s + (n - 1) // 2 + 10) ** 2) for s, n in blocks]
    # segment 0: bins [1, 4) with the first block's mean, no slope
    m0 = np.float32(np.float64(med[0]) / LN2)
    for b in range(1, 4):
        assert out[b].real == np.float32(np.float32(b + 10) * _quake(m0))
    # segments j = 1..4: bins [prev_hi, start_j + len_j//2), slope towards block j
    lo = 4
    mean_old = m0
    for j in range(1, 5):
        s, n = blocks[j]
        mean_new = np.float64(med[j]) / LN2
        slope = np.float32((mean_new - np.float64(mean_old)) / (0.5 * (blocks[j - 1][1] + n)))
        hi = s + n // 2
        for b in range(lo, hi):
            t = np.float32(slope * np.float32(b - lo))
            want = np.float32(np.float32(b + 10) * _quake(np.float32(mean_old + t)))
            assert out[b].real == want, (j, b)
        lo = hi
        mean_old = np.float32(mean_new)


def test_deredden_flattens_linear_spectrum(oracle):
    """A power spectrum linear in bin index is flattened to ln2 * (local power /
    interpolated median) ~ ln2 up to the fast rsqrt's error and a half-bin offset."""
    nb = 100000
    b = np.arange(nb, dtype=np.float64)
    power = 1000.0 + 0.5 * b
    spec = np.sqrt(power).astype(np.complex64)
    out = oracle.deredden(spec, nb)
    p = np.abs(out[200:nb - 300].astype(np.complex128)) ** 2
    assert np.max(np.abs(p / LN2 - 1.0)) < 5e-3


def test_segment_normalisation(oracle):
    """source/bpd_search.cu:157-174: only the first numdata amplitudes are scaled by
    rsqrt(median/ln2); bins outside the spectrum read as zero."""
    rng = np.random.default_rng(5)
    nb = 10000
    spec = (rng.standard_normal(nb) + 1j * rng.standard_normal(nb)).astype(np.complex64)
    L, lobin, numdata = 2048, nb - 900, 1000
    amps, med = oracle.segment_amplitudes(spec, lobin, numdata, L)
    raw = np.concatenate([spec[lobin:], np.zeros(numdata - 900, np.complex64)])
    pw = (raw.real * raw.real + raw.imag * raw.imag).astype(np.float32)
    assert med == np.sort(pw)[(numdata - 1) // 2]
    c = _quake(np.float32(np.float64(med) / LN2))
    np.testing.assert_array_equal(amps[:900], (spec[lobin:] * c).astype(np.complex64))
    assert np.all(amps[900:numdata] == 0) and np.all(amps[numdata:] == 0)
    amps2, _ = oracle.segment_amplitudes(spec, 100, numdata, L)
    np.testing.assert_array_equal(amps2[numdata:], spec[100 + numdata:100 + L // 2])
