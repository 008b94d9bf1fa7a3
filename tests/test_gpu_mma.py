"""The direct correlation on the matrix cores (precision "i16mma",
FDAS_PREC_I16_MMA) against the FP64 oracle.

Each kept column is computed without an FFT, as an exact integer sum of the
quantised segment values and template taps (int8 halves, 32-bit accumulation),
and scaled back to an FP32 power; the planes are P16 and the harmonic sum is that
of the other reduced-precision modes.  The power tolerance is that of reduced
precision, 1e-3 of max(P_ref, 1) (the oracle's direct form: or_row_powers_direct,
restating source/bpd_search.cu:243-307).  Candidates must agree with the oracle
wherever its stage sums are farther than that from a cut.  Batches, chunking and
graph replay must be bit-identical to single plans."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

TOL = 1e-3


def power_err(p, pref):
    return float(np.max(np.abs(p.astype(np.float64) - pref) / np.maximum(pref, 1.0)))


@pytest.fixture(scope="module")
def fd():
    import gpu_accelsearch_amd as fd
    fd.lib()
    return fd


@pytest.mark.parametrize("zmax,numharm,fftlen", [(50, 4, 2048), (0, 1, 1024), (200, 16, 8192),
                                                 (100, 2, 4096)])
def test_mma_steps_match_oracle(fd, oracle, zmax, numharm, fftlen):
    """Summed planes of the first, a middle, the pulsar's and the last step."""
    from gpu_accelsearch_amd import synth
    n = 1 << 20
    r0 = 44000.7
    x = synth.plain_series(n, seed=23, r0=r0, z0=min(12.0, float(zmax) / 4), power=90.0)
    print_case("mma steps", zmax, numharm, fftlen)
    P = fd.SearchParams(n=n, zmax=zmax, numharm=numharm, fftlen=fftlen, precision="i16mma")
    with fd.AccelSearch(P) as s:
        ref = oracle.Search(x, n, 6.4e-5, zmax, numharm, fftlen)
        assert (s.uselen, s.rlo, s.nsteps) == (ref.uselen, ref.rlo, ref.nsteps)
        s.prepare(x)
        steps = sorted({0, s.nsteps // 2, synth.steps_covering(numharm * r0, s.rlo, s.uselen),
                        s.nsteps - 1})
        worst = 0.0
        for st in steps:
            if not 0 <= st < s.nsteps:
                continue
            worst = max(worst, power_err(s.run_steps(st, 1)[0], ref.step_plane(st)))
        ref.close()
    print(f"MMA zmax {zmax} numharm {numharm} fftlen {fftlen}: max |dP|/max(P,1) {worst:.3e}")
    assert worst <= TOL, worst


def test_mma_production_steps_and_candidates(fd, oracle):
    """2^23 samples, zmax 200, 8 harmonics: the pulsar steps' powers and candidates."""
    from gpu_accelsearch_amd import synth
    n = 1 << 23
    x = synth.production_series(n)
    powcut = (40.0, 45.0, 55.0, 70.0, 0.0)
    print_case("mma production")
    P = fd.SearchParams(n=n, zmax=200, numharm=8, fftlen=8192, precision="i16mma", powcut=powcut)
    with fd.AccelSearch(P) as s:
        ref = oracle.Search(x, n, 6.4e-5, 200, 8, 8192)
        steps = sorted({synth.steps_covering(8 * r0, s.rlo, s.uselen)
                        for r0, _, _ in synth.PRODUCTION_PULSARS})
        s.prepare(x)
        worst = 0.0
        for st in steps:
            worst = max(worst, power_err(s.run_steps(st, 1)[0], ref.step_plane(st)))
        print(f"MMA production max relative power error {worst:.3e}")
        assert worst <= TOL, worst
        s.search(x)
        got = {(int(c["step"]), int(c["z_index"]), int(c["r_index"]), int(c["stage"]))
               for c in s.candidates() if int(c["step"]) in steps}
        want, ambiguous = set(), set()
        for st in steps:
            _, stages = ref.step_plane(st, stages=True)
            for sg in range(4):
                z, i = np.nonzero(stages[sg] > powcut[sg])
                want |= {(st, int(a), int(b), sg) for a, b in zip(z, i)}
                z, i = np.nonzero(np.abs(stages[sg] - powcut[sg]) <= TOL * np.maximum(stages[sg], 1.0))
                ambiguous |= {(st, int(a), int(b), sg) for a, b in zip(z, i)}
        ref.close()
    print(f"MMA production: {len(want)} oracle candidates, {len(got)} found, "
          f"{len(ambiguous)} within the error of a cut")
    assert len(want) > 0
    assert got - ambiguous == want - ambiguous


def test_mma_batch_graph_matches_single_plans(fd):
    """Uneven trials in one batch (chunks mixing trials), eager and graph replay,
    bit-identical to single-series plans."""
    from gpu_accelsearch_amd import synth
    lengths = [1 << 19, 600_002, (1 << 19) + 2]
    xs = [synth.plain_series(n, seed=60 + t, r0=20000.3 + 3000.1 * t, z0=-10.0 + 8.0 * t, power=90.0)
          for t, n in enumerate(lengths)]
    powcut = (26.0, 32.0, 40.0, 0.0, 0.0)
    print_case("mma batch graph", lengths)

    def cset(c):
        return sorted((int(a["trial"]), int(a["step"]), int(a["z_index"]), int(a["r_index"]),
                       int(a["stage"]), float(a["power"])) for a in c)

    single = []
    for t, (n, x) in enumerate(zip(lengths, xs)):
        P = fd.SearchParams(n=n, zmax=50, numharm=4, fftlen=2048, precision="i16mma", powcut=powcut)
        with fd.AccelSearch(P) as s:
            s.search(x)
            c = s.candidates()
            c["trial"] = t
            single += cset(c)
    single.sort()
    assert len(single) > 0
    P = fd.SearchParams(n=0, zmax=50, numharm=4, fftlen=2048, precision="i16mma", powcut=powcut,
                        chunk_steps=29)
    with fd.BatchSearch(P, lengths, graph=True) as b:
        for t, x in enumerate(xs):
            b.upload(t, x)
        runs = []
        for _ in range(3):   # eager (first), then graph replays
            b.run()
            runs.append(cset(b.candidates()))
    for r in runs:
        assert r == single
