"""Correlation, powers and harmonic summing of the oracle."""
import numpy as np
import pytest

from gpu_accelsearch_amd import synth


@pytest.mark.parametrize("z,L", [(0, 1024), (24, 2048), (-50, 2048), (200, 8192)])
def test_fft_correlation_equals_direct(oracle, z, L):
    """In the kept columns the circular FFT correlation equals the direct sum
    sum_t s[n+t] conj(tap[2hw+t]) over the spread segment (no wrap-around)."""
    rng = np.random.default_rng(L + z)
    amps = (rng.standard_normal(L // 2) + 1j * rng.standard_normal(L // 2)).astype(np.complex64)
    hw_max = oracle.lib().or_z_response_halfwidth(abs(z) + 10)
    off = 2 * hw_max
    numrs = L - 4 * hw_max - 2
    row = oracle.template_row(z, L)
    K = np.fft.fft(row)
    Kd = np.ascontiguousarray(np.stack([K.real, K.imag], -1).reshape(-1))
    a = np.ascontiguousarray(amps.view(np.float32))
    p_fft = np.zeros(numrs)
    p_dir = np.zeros(numrs)
    oracle.lib().or_row_powers_fft(a, Kd, L, off, numrs, p_fft)
    oracle.lib().or_row_powers_direct(a, z, L, off, numrs, p_dir)
    assert np.max(np.abs(p_fft - p_dir)) < 1e-9 * max(1.0, p_dir.max())


def test_noise_powers_are_unit_mean(oracle):
    """Median normalisation makes noise powers Exp(1): at z = 0 the even columns are
    the normalised input bins, mean 1 and median ln 2."""
    n = 1 << 18
    x = np.random.default_rng(8).standard_normal(n).astype(np.float32)
    S = oracle.Search(x, n, 6.4e-5, 20, 1, 1024)
    vals = np.concatenate([S.step_plane(st)[10, ::2] for st in range(2, 60, 3)])
    assert abs(vals.mean() - 1.0) < 0.03
    assert abs(np.median(vals) - np.log(2.0)) < 0.03
    S.close()


def test_injected_signal_recovered(oracle):
    n = 1 << 20
    x = synth.plain_series(n)
    S = oracle.Search(x, n, 6.4e-5, 50, 1, 2048)
    k = synth.steps_covering(50000.3, S.rlo, S.uselen)
    P = S.step_plane(k)
    j, i = np.unravel_index(np.argmax(P), P.shape)
    assert -50 + 2 * j == 24
    assert abs(S.rlo + k * (S.uselen // 2) + 0.5 * i - 50000.3) <= 0.25
    assert P.max() > 60
    S.close()


def test_stage_partial_sums(oracle):
    """Stage planes are the running sum in reference order; the last equals the plane."""
    n = 1 << 18
    x = synth.plain_series(n, seed=4, r0=8000.2, z0=6.0, power=80.0)
    S = oracle.Search(x, n, 6.4e-5, 30, 8, 2048)
    plane, st = S.step_plane(5, stages=True)
    assert st.shape[0] == 4
    np.testing.assert_array_equal(st[-1], plane)
    assert np.all(st[1] >= st[0]) and np.all(st[2] >= st[1]) and np.all(st[3] >= st[2])
    S.close()


def test_harmonic_sum_collects_planted_harmonics(oracle):
    """A pulse train with fundamental (r1, z1) gives its largest 4-harmonic sum at
    the cell (4 r1, 4 z1), and that sum exceeds the fundamental-only power there."""
    n = 1 << 20
    r1, z1 = 3000.5, 2.0
    x = np.random.default_rng(6).standard_normal(n)
    x += synth.pulse_train(n, r1, z1, 0.05, 25.0)
    x = x.astype(np.float32)
    S = oracle.Search(x, n, 6.4e-5, 20, 4, 2048)
    k = synth.steps_covering(4 * r1, S.rlo, S.uselen)
    plane, st = S.step_plane(k, stages=True)
    j, i = np.unravel_index(np.argmax(plane), plane.shape)
    assert -20 + 2 * j == 4 * z1
    assert abs(S.rlo + k * (S.uselen // 2) + 0.5 * i - 4 * r1) <= 0.5
    assert plane[j, i] > 2.0 * st[0][j, i]
    S.close()


def test_search_candidates_match_stage_planes(oracle):
    """or_search_candidates (OpenMP over steps) against the stage partial sums of
    or_step_plane, step by step: the cells above each cut, and the cells within
    the margin of a cut flagged ambiguous."""
    from gpu_accelsearch_amd import synth
    n = 1 << 18
    x = synth.plain_series(n, seed=17, r0=9000.3, z0=6.0, power=70.0)
    ref = oracle.Search(x, n, 6.4e-5, 30, 4, 1024)
    cut, margin = (14.0, 18.0, 26.0, 0.0, 0.0), 0.05
    cands = ref.candidates(cut, margin)
    assert len(cands) > 0 and cands["ambiguous"].sum() > 0
    assert np.all(np.diff(cands["step"]) >= 0)
    for st in range(ref.nsteps):
        _, S = ref.step_plane(st, stages=True)
        above, amb = set(), set()
        for s in range(3):
            P = S[s]
            near = np.abs(P - cut[s]) <= margin * np.maximum(P, 1.0)
            z, i = np.nonzero(near)
            amb |= {(int(a), int(b), s) for a, b in zip(z, i)}
            z, i = np.nonzero((P > cut[s]) & ~near)
            above |= {(int(a), int(b), s) for a, b in zip(z, i)}
        mine = cands[cands["step"] == st]
        assert {(int(c["z_index"]), int(c["r_index"]), int(c["stage"])) for c in mine
                if not c["ambiguous"]} == above
        assert {(int(c["z_index"]), int(c["r_index"]), int(c["stage"])) for c in mine
                if c["ambiguous"]} == amb
    ref.close()
