"""Reduced-precision storage against the FP64 oracle.  Segments and templates are
stored in 16 bits (FP16, BF16, or int16 block floating point I16) with power-of-
two exponents, and the power planes in 16 bits (P16); FFT arithmetic and power
accumulation stay FP32.

The power tolerance for reduced precision is 1e-3 of max(P_ref, 1).  I16 meets
it and is tested against it, as is I16_MMA (the same 16-bit quantisation, time-
domain, correlated exactly in integers on the matrix cores; tests/test_gpu_mma.py
covers its other shapes).  FP16 and BF16 do not: their 11- and 8-bit
significands round every stored template and segment value, and that rounding
adds up in each correlation output.  Their tests assert the stated limit of
each mode: the largest error measured over these workloads, with a small
margin.  Each test also checks candidate agreement wherever the oracle's stage
sums are farther than the mode's bound from the threshold."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

# I16 and I16_MMA: the 1e-3 tolerance (measured at most 5.7e-4 and 6.6e-4 over
# these tests and tests/test_gpu_mma.py).
IN_TOL = {"i16": 1e-3, "i16mma": 1e-3}
# Stated limits: largest |dP| / max(P_ref, 1) measured over these tests, plus a
# small margin.  FP16 1.6e-3, BF16 1.27e-2.
STATED_LIMIT = {"fp16": 1.8e-3, "bf16": 1.4e-2}
BOUND = {**IN_TOL, **STATED_LIMIT}


def power_err(p, pref):
    return float(np.max(np.abs(p.astype(np.float64) - pref) / np.maximum(pref, 1.0)))


@pytest.fixture(scope="module")
def fd():
    import gpu_accelsearch_amd as fd
    fd.lib()
    return fd


def _plain(fd, oracle, mode):
    from gpu_accelsearch_amd import synth
    n = 1 << 20
    x = synth.plain_series(n, seed=21, r0=44000.7, z0=12.0, power=90.0)
    print_case("precision plain", mode)
    P = fd.SearchParams(n=n, zmax=50, numharm=4, fftlen=2048, precision=mode)
    with fd.AccelSearch(P) as s:
        ref = oracle.Search(x, n, 6.4e-5, 50, 4, 2048)
        s.prepare(x)
        worst = 0.0
        for st in (0, 40, synth.steps_covering(4 * 44000.7, s.rlo, s.uselen), s.nsteps - 1):
            worst = max(worst, power_err(s.run_steps(st, 1)[0], ref.step_plane(st)))
        ref.close()
    print(f"{mode} max relative power error {worst:.3e}")
    return worst


def test_precision_plain_in_tolerance(fd, oracle):
    assert _plain(fd, oracle, "i16") <= IN_TOL["i16"]


@pytest.mark.parametrize("mode", ["fp16", "bf16"])
def test_precision_plain_stated_limit(fd, oracle, mode):
    assert _plain(fd, oracle, mode) <= STATED_LIMIT[mode]


def _production(fd, oracle, mode):
    """2^23 samples, zmax 200, 8 harmonics: powers on the pulsar steps and the
    candidate sets there."""
    from gpu_accelsearch_amd import synth
    n = 1 << 23
    x = synth.production_series(n)
    powcut = (40.0, 45.0, 55.0, 70.0, 0.0)
    print_case("precision production", mode)
    P = fd.SearchParams(n=n, zmax=200, numharm=8, fftlen=8192, precision=mode, powcut=powcut)
    with fd.AccelSearch(P) as s:
        ref = oracle.Search(x, n, 6.4e-5, 200, 8, 8192)
        steps = sorted({synth.steps_covering(8 * r0, s.rlo, s.uselen)
                        for r0, _, _ in synth.PRODUCTION_PULSARS})
        s.prepare(x)
        worst = 0.0
        for st in steps:
            worst = max(worst, power_err(s.run_steps(st, 1)[0], ref.step_plane(st)))
        print(f"{mode} production max relative power error {worst:.3e}")
        assert worst <= BOUND[mode], worst
        s.search(x)
        got = {(int(c["step"]), int(c["z_index"]), int(c["r_index"]), int(c["stage"]))
               for c in s.candidates() if int(c["step"]) in steps}
        want, ambiguous = set(), set()
        for st in steps:
            _, stages = ref.step_plane(st, stages=True)
            for sg in range(4):
                z, i = np.nonzero(stages[sg] > powcut[sg])
                want |= {(st, int(a), int(b), sg) for a, b in zip(z, i)}
                z, i = np.nonzero(np.abs(stages[sg] - powcut[sg]) <=
                                  BOUND[mode] * np.maximum(stages[sg], 1.0))
                ambiguous |= {(st, int(a), int(b), sg) for a, b in zip(z, i)}
        assert len(want) > 0
        assert got - ambiguous == want - ambiguous
        print(f"{mode}: {len(want)} oracle candidates, {len(got)} found, "
              f"{len(ambiguous)} within the error of the threshold")
        ref.close()


def test_precision_production_in_tolerance(fd, oracle):
    _production(fd, oracle, "i16")


@pytest.mark.parametrize("mode", ["fp16", "bf16"])
def test_precision_production_stated_limit(fd, oracle, mode):
    _production(fd, oracle, mode)


# Precision workload: 2^24 samples of N(0,1) + red noise + 5 accelerated pulsars
# whose first harmonics span ~6 to ~30 sigma; zmax 300, 8 harmonics, fftlen 8192.
C4_PULSARS = ((120000.3, 30.0, 21.0), (250000.6, -25.0, 45.0), (410000.2, 12.0, 90.0),
              (600000.8, -35.0, 200.0), (900000.4, 20.0, 455.0))
C4_POWCUT = (30.0, 36.0, 45.0, 60.0, 0.0)
C4_BOUND = {"fp32": 1e-4, **BOUND}


@pytest.fixture(scope="module")
def c4(oracle):
    from gpu_accelsearch_amd import synth
    n = 1 << 24
    x = synth.production_series(n, seed=8, pulsars=C4_PULSARS)
    ref = oracle.Search(x, n, 6.4e-5, 300, 8, 8192)
    steps = sorted({synth.steps_covering(8 * r0, ref.rlo, ref.uselen) for r0, _, _ in C4_PULSARS})
    planes = {st: ref.step_plane(st, stages=True) for st in steps}
    ref.close()
    return x, n, steps, planes


def _workload(fd, c4, mode):
    """Max relative power error and candidate agreement per storage mode."""
    x, n, steps, planes = c4
    print_case("precision workload 2^24 zmax 300", mode)
    P = fd.SearchParams(n=n, zmax=300, numharm=8, fftlen=8192, precision=mode, powcut=C4_POWCUT)
    with fd.AccelSearch(P) as s:
        assert s.uselen == 7470
        s.prepare(x)
        worst = worst_big = 0.0
        for st in steps:
            got = s.run_steps(st, 1)[0].astype(np.float64)
            want = planes[st][0]
            worst = max(worst, power_err(got, want))
            big = want >= 1.0
            worst_big = max(worst_big, float(np.max(np.abs(got[big] - want[big]) / want[big])))
        s.search(x)
        got = {(int(c["step"]), int(c["z_index"]), int(c["r_index"]), int(c["stage"]))
               for c in s.candidates() if int(c["step"]) in steps}
    want, ambiguous = set(), set()
    for st in steps:
        stages = planes[st][1]
        for sg in range(4):
            z, i = np.nonzero(stages[sg] > C4_POWCUT[sg])
            want |= {(st, int(a), int(b), sg) for a, b in zip(z, i)}
            z, i = np.nonzero(np.abs(stages[sg] - C4_POWCUT[sg]) <=
                              C4_BOUND[mode] * np.maximum(stages[sg], 1.0))
            ambiguous |= {(st, int(a), int(b), sg) for a, b in zip(z, i)}
    agree = got - ambiguous == want - ambiguous
    print(f"PRECISION {mode}: max |dP|/max(P,1) {worst:.3e}, max |dP|/P (P>=1) {worst_big:.3e}, "
          f"oracle candidates {len(want)}, found {len(got)}, within error of threshold "
          f"{len(ambiguous)}, agree {agree}")
    assert worst <= C4_BOUND[mode], worst
    assert len(want) > 0 and agree


@pytest.mark.parametrize("mode", ["fp32", "i16", "i16mma"])
def test_precision_workload_in_tolerance(fd, c4, mode):
    _workload(fd, c4, mode)


@pytest.mark.parametrize("mode", ["fp16", "bf16"])
def test_precision_workload_stated_limit(fd, c4, mode):
    _workload(fd, c4, mode)
