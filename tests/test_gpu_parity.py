"""GPU parity: every HIP stage against the CPU oracle (oracle/), through the C ABI.

Tolerances (FP32 on the device vs the FP64 oracle):
  powers      max |dP| / max(P_ref, 1) <= 1e-4
  FFTs        max |dX| / rms(X_ref)    <= 2e-6
  whitening   bit-exact (same float input, same float/double operation order)
  medians     bit-exact (exact order statistic)
"""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

POWER_TOL = 1e-4


def power_err(p, pref):
    return float(np.max(np.abs(p.astype(np.float64) - pref) / np.maximum(pref, 1.0)))


@pytest.fixture(scope="module")
def fd():
    import gpu_accelsearch_amd as fd
    fd.lib()
    return fd


@pytest.mark.parametrize("n", [64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 32768,
                               1 << 16, 1 << 20])
def test_c2c_fft(fd, n):
    print_case("c2c", n)
    rng = np.random.default_rng(n)
    batch = 3 if n <= 16384 else 1
    x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))).astype(np.complex64)
    d = fd.DeviceArray.from_host(x)
    for inverse in (False, True):
        out = (fd.inverseComplexToComplexFFT if inverse else fd.complexToComplexFFT)(d, n, batch)
        y = out.to_host().reshape(batch, n)
        ref = np.fft.ifft(x.astype(np.complex128), axis=1) * n if inverse else \
            np.fft.fft(x.astype(np.complex128), axis=1)
        err = np.max(np.abs(y - ref)) / np.sqrt(np.mean(np.abs(ref) ** 2))
        assert err < 2e-6 * np.log2(n), (n, inverse, err)
        out.free()


@pytest.mark.parametrize("n", [256, 4096, 1 << 16, 1 << 20, 1 << 23, 3 << 18, (1 << 20) + 2,
                               2_000_006])
def test_r2c_fft(fd, oracle, n):
    print_case("r2c", n)
    rng = np.random.default_rng(7 + n)
    x = rng.standard_normal(n).astype(np.float32)
    y = fd.realToComplexFFT(x)
    ref = np.fft.rfft(x.astype(np.float64))
    assert y.shape == ref.shape
    err = np.max(np.abs(y - ref)) / np.sqrt(np.mean(np.abs(ref) ** 2))
    assert err < 2e-6 * np.log2(n), err
    if n <= 4096:   # the oracle's own transform agrees as well
        assert np.allclose(oracle.rfft(x), ref, rtol=0, atol=1e-9 * np.abs(ref).max())


@pytest.mark.parametrize("nbins", [16, 100, 5000, 1 << 19])
def test_deredden_bit_exact(fd, oracle, nbins):
    print_case("deredden", nbins)
    rng = np.random.default_rng(nbins)
    f = np.arange(nbins) + 1.0
    amp = np.sqrt(1.0 + 50.0 * (f / 30.0) ** -2)
    spec = (amp * (rng.standard_normal(nbins) + 1j * rng.standard_normal(nbins))).astype(np.complex64)
    got = fd.deredden(spec, nbins)
    want = oracle.deredden(spec, nbins)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_template_bank(fd, oracle):
    print_case("templates", "stages=4 zmax=200 uselen=7734")
    fd.accel.fdas_set_uselen(7734)
    try:
        bank = fd.createSubharmonicInfo(4, 200)
        for h in range(bank.total):
            e = bank.entries[h]
            got = bank.bank(h)
            L, Z = e.kernelFFTLength, e.totalKernelsTemplates
            for k in sorted({0, Z // 3, Z // 2, Z - 1}):
                row = oracle.template_row(-e.zmax + 2 * k, L)
                ref = np.fft.fft(row)
                err = np.max(np.abs(got[k] - ref)) / np.sqrt(np.mean(np.abs(ref) ** 2))
                assert err < 2e-6 * np.log2(L), (h, k, err)
        bank.free()
    finally:
        fd.accel.fdas_set_uselen(32000)


def _plan_vs_oracle(fd, oracle, series, n, zmax, numharm, fftlen, steps):
    P = fd.SearchParams(n=n, dt=6.4e-5, zmax=zmax, numharm=numharm, fftlen=fftlen)
    with fd.AccelSearch(P) as s:
        ref = oracle.Search(series, n, 6.4e-5, zmax, numharm, fftlen)
        assert (s.uselen, s.rlo, s.nsteps, s.numz) == (ref.uselen, ref.rlo, ref.nsteps, ref.numz)
        s.prepare(series)
        spec = s.spectrum()
        rspec = ref.spectrum()
        serr = np.max(np.abs(spec - rspec)) / np.sqrt(np.mean(np.abs(rspec) ** 2))
        assert serr < 2e-5, serr
        worst = 0.0
        for st in steps:
            got = s.run_steps(st, 1)[0]
            want = ref.step_plane(st)
            e = power_err(got, want)
            worst = max(worst, e)
            assert e <= POWER_TOL, (st, e)
        ref.close()
        return worst


def test_plan_plain_fp32(fd, oracle):
    """2^20 samples, zmax 50, fftlen 2048, one harmonic: every step vs the oracle."""
    from gpu_accelsearch_amd import synth
    n = 1 << 20
    x = synth.plain_series(n)
    print_case("plan C1 all steps")
    P = fd.SearchParams(n=n, zmax=50, numharm=1, fftlen=2048)
    with fd.AccelSearch(P) as s:
        ref = oracle.Search(x, n, 6.4e-5, 50, 1, 2048)
        s.prepare(x)
        planes = s.run_steps(0, s.nsteps)
        worst = 0.0
        for st in range(s.nsteps):
            worst = max(worst, power_err(planes[st], ref.step_plane(st)))
        assert worst <= POWER_TOL, worst
        k = synth.steps_covering(50000.3, s.rlo, s.uselen)
        j, i = np.unravel_index(np.argmax(planes[k]), planes[k].shape)
        assert -50 + 2 * j == 24 and abs(s.rlo + k * (s.uselen // 2) + 0.5 * i - 50000.3) <= 0.5
        ref.close()


def test_plan_production_steps(fd, oracle):
    """2^23 samples with red noise and pulsars, zmax 200, 8 harmonics: selected steps."""
    from gpu_accelsearch_amd import synth
    n = 1 << 23
    x = synth.production_series(n)
    uselen, rlo = 7734, 536
    steps = {0, 1083}
    for r0, z0, _ in synth.PRODUCTION_PULSARS:
        steps.add(synth.steps_covering(8 * r0, rlo, uselen))
    print_case("plan C2 steps", sorted(steps))
    worst = _plan_vs_oracle(fd, oracle, x, n, 200, 8, 8192, sorted(steps))
    print("C2 max power error", worst)


def test_plan_harmonics_4_16(fd, oracle):
    """numharm 2, 4 and 16 on a short series (every stage boundary of the sum)."""
    from gpu_accelsearch_amd import synth
    n = 1 << 18
    x = synth.plain_series(n, seed=5, r0=9000.7, z0=-10.0, power=60.0)
    # (16, 2048): the 1/16 harmonic runs at FFT length 256 (128-point segments)
    for numharm, fftlen in ((2, 1024), (4, 2048), (16, 4096), (16, 2048)):
        print_case("plan numharm", numharm)
        _plan_vs_oracle(fd, oracle, x, n, 60, numharm, fftlen, [0, 3, 7])


def test_plan_uneven_length(fd, oracle):
    """Series lengths that are not powers of two (chirp-z R2C path)."""
    from gpu_accelsearch_amd import synth
    for n, steps in ((1_000_002, [0, 100, 500]), (3 << 19, [1, 700])):
        x = synth.plain_series(n, seed=12, r0=30000.4, z0=-6.0, power=70.0)
        print_case("plan uneven n", n)
        _plan_vs_oracle(fd, oracle, x, n, 50, 2, 2048, steps)


def test_stage_pipeline_matches_plan(fd, oracle):
    """Reference-named stage entry points (AccelSearch.cu:217-241) on one step: the
    summed plane against the FP64 oracle (1e-4) and against the fused plan."""
    from gpu_accelsearch_amd import synth
    n = 1 << 20
    x = synth.plain_series(n, seed=3, r0=70000.2, z0=-30.0)
    zmax, numharm, fftlen = 50, 4, 2048
    print_case("stage pipeline")
    P = fd.SearchParams(n=n, zmax=zmax, numharm=numharm, fftlen=fftlen)
    ref = oracle.Search(x, n, 6.4e-5, zmax, numharm, fftlen)
    with fd.AccelSearch(P) as s:
        s.prepare(x)
        spec = s.spectrum()
        fd.accel.fdas_set_uselen(s.uselen)
        try:
            bank = fd.createSubharmonicInfo(3, zmax)
            pipe = fd.StagePipeline(bank)
            for st in (0, 11, synth.steps_covering(70000.2, s.rlo, s.uselen), s.nsteps - 1):
                startr = s.rlo + st * (s.uselen // 2)
                got = pipe.step(spec, startr, -zmax, s.uselen, s.nbins)
                want = s.run_steps(st, 1)[0]
                assert power_err(got, want.astype(np.float64)) <= 2e-5
                err = power_err(got, ref.step_plane(st))
                print(f"stage pipeline step {st}: max |dP|/max(P,1) vs oracle {err:.2e}")
                assert err <= 1e-4, err
            pipe.close()
            bank.free()
        finally:
            fd.accel.fdas_set_uselen(32000)
            ref.close()


def test_candidates_match_oracle(fd, oracle):
    """Threshold + compaction agree with the oracle's stage partial sums."""
    from gpu_accelsearch_amd import synth
    n = 1 << 20
    x = synth.plain_series(n, seed=9, r0=40000.6, z0=8.0, power=150.0)
    powcut = (30.0, 40.0, 50.0, 0.0, 0.0)
    print_case("candidates")
    P = fd.SearchParams(n=n, zmax=40, numharm=4, fftlen=2048, powcut=powcut)
    with fd.AccelSearch(P) as s:
        ref = oracle.Search(x, n, 6.4e-5, 40, 4, 2048)
        s.search(x)
        cands = s.candidates()
        assert not s.overflowed
        got = {(int(c["step"]), int(c["z_index"]), int(c["r_index"]), int(c["stage"])) for c in cands}
        want, margin_ok = set(), True
        steps = sorted({int(c["step"]) for c in cands} | {synth.steps_covering(r, s.rlo, s.uselen)
                                                        for r in (40000.6, 80001.2, 120001.8, 160002.4)})
        for st in steps:
            _, stages = ref.step_plane(st, stages=True)
            for sg in range(3):
                z, i = np.nonzero(stages[sg] > powcut[sg])
                want |= {(st, int(a), int(b), sg) for a, b in zip(z, i)}
                near = np.abs(stages[sg] - powcut[sg]) <= POWER_TOL * np.maximum(stages[sg], 1.0)
                margin_ok &= not near.any()
        got_sel = {g for g in got if g[0] in steps}
        assert len(want) > 0
        if margin_ok:
            assert got_sel == want
        assert got == got_sel, "candidates outside the checked steps"
        ref.close()


def test_search_is_deterministic(fd):
    from gpu_accelsearch_amd import synth
    n = 1 << 20
    x = synth.plain_series(n, seed=4)
    print_case("determinism")
    P = fd.SearchParams(n=n, zmax=50, numharm=2, fftlen=2048, chunk_steps=64)
    with fd.AccelSearch(P) as s:
        s.prepare(x)
        a = s.run_steps(100, 40)
        s.prepare(x)
        b = s.run_steps(100, 40)
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
        c = s.run_steps(120, 5)
        np.testing.assert_array_equal(c.view(np.uint32), a[20:25].view(np.uint32))


def test_plan_zmax_zero(fd, oracle):
    """zmax 0: one template (the z = 0 interpolation kernel), harmonics 1 and 4."""
    from gpu_accelsearch_amd import synth
    n = 1 << 19
    x = synth.plain_series(n, seed=14, r0=21000.3, z0=0.0, power=90.0)
    for numharm, fftlen in ((1, 1024), (4, 2048)):
        print_case("plan zmax 0, numharm", numharm)
        _plan_vs_oracle(fd, oracle, x, n, 0, numharm, fftlen, [0, 100, 250])


def test_plan_largest_shape(fd, oracle):
    """The largest serving shape: 2^25 samples, zmax 600 (601 templates),
    16 harmonics, fftlen 16384 (correlation lengths up to 16384)."""
    from gpu_accelsearch_amd import synth
    n = 1 << 25
    rng = np.random.default_rng(15)
    x = rng.standard_normal(n).astype(np.float32)
    x = synth.add_chirp(x, 400000.3, 250.0, 200.0)
    st = synth.steps_covering(400000.3, 2147, 14942)
    print_case("plan 2^25 zmax 600 numharm 16 steps", [0, st])
    _plan_vs_oracle(fd, oracle, x, n, 600, 16, 16384, [0, st])


@pytest.mark.parametrize("zmax,numharm,fftlen,want", [
    (200, 8, 8192, [8192, 4096, 2560, 6144, 1280, 4096, 5120, 7168]),
    (200, 8, 4096, [4096, 2048, 1280, 3072, 768, 2048, 2560, 4096]),
    (0, 8, 8192, [8192, 5120, 2560, 7168, 1280, 4096, 6144, 8192]),
    (600, 16, 16384, [16384, 10240, 4096, 16384, 2048, 6144, 12288, 16384, 1280, 3072, 5120, 7168,
                      10240, 12288, 16384, 16384]),
])
def test_sub_harmonic_lengths(fd, oracle, zmax, numharm, fftlen, want):
    """Sub-harmonics run at 5/8, 3/4 or 7/8 of the reference's power-of-two FFT
    length where every step's bins and template span fit (the kept columns do
    not depend on the length): both length modes against the oracle and against
    each other."""
    from gpu_accelsearch_amd import synth
    n = 1 << 21
    x = synth.plain_series(n, seed=21, r0=60000.3, z0=-0.6 * zmax, power=90.0)
    steps = [0, 1, 5, 100]
    planes = {}
    for mode in ("shortest", "reference"):
        print_case("lengths", mode, zmax, numharm, fftlen)
        P = fd.SearchParams(n=n, zmax=zmax, numharm=numharm, fftlen=fftlen, fft_lengths=mode)
        with fd.AccelSearch(P) as s:
            got = [h[2] for h in s.harmonics]
            if mode == "shortest":
                assert got == want, got
            else:
                assert all(L & (L - 1) == 0 for L in got), got
            s.prepare(x)
            planes[mode] = [s.run_steps(st, 1)[0] for st in steps]
    ref = oracle.Search(x, n, 6.4e-5, zmax, numharm, fftlen)
    for i, st in enumerate(steps):
        want_plane = ref.step_plane(st)
        for mode in planes:
            assert power_err(planes[mode][i], want_plane) <= POWER_TOL, (mode, st)
        assert power_err(planes["shortest"][i], planes["reference"][i].astype(np.float64)) <= 2e-5
    ref.close()


def test_length_modes_same_candidates(fd):
    """A whole production trial (2^23, zmax 200, 8 harmonics, pulsars and red
    noise) gives the same candidate list under both sub-harmonic length modes,
    apart from cells within FP32 rounding of a threshold, with powers that agree."""
    from gpu_accelsearch_amd import synth
    n = 1 << 23
    x = synth.production_series(n)
    powcut = (40.0, 45.0, 55.0, 70.0, 0.0)
    lists = {}
    for mode in ("shortest", "reference"):
        print_case("candidates, lengths", mode)
        P = fd.SearchParams(n=n, zmax=200, numharm=8, fftlen=8192, powcut=powcut, fft_lengths=mode)
        with fd.AccelSearch(P) as s:
            s.search(x)
            c = s.candidates()
            assert not s.overflowed
            lists[mode] = {(int(a["step"]), int(a["z_index"]), int(a["r_index"]), int(a["stage"])): float(a["power"])
                           for a in c}
    a, b = lists["shortest"], lists["reference"]
    assert len(a) > 0
    for key in set(a) ^ set(b):
        p = a.get(key, b.get(key))
        assert abs(p - powcut[key[3]]) <= 2e-5 * max(p, 1.0), (key, p)
    for key in set(a) & set(b):
        assert abs(a[key] - b[key]) <= 2e-5 * max(b[key], 1.0), (key, a[key], b[key])
