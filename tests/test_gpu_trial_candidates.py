"""Whole-trial candidate agreement with the FP64 oracle: every cell of every step
whose stage sum passes its cut (the search hook of source/bpd_search.cu:350-356,
candidates emitted after harmonics 1, 2, 4, 8) must be found, and nothing else,
except cells whose oracle stage sum lies within the mode's power tolerance of a
cut (1e-4 of max(P, 1) in FP32, 1e-3 in the reduced-precision mode), which may
fall either way.  The oracle runs every step (OpenMP over steps)."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

POWCUT = (22.0, 28.0, 34.0, 31.0, 0.0)


@pytest.fixture(scope="module")
def fd():
    import gpu_accelsearch_amd as fd
    fd.lib()
    return fd


def _sets(oracle_cands, trial=None):
    want = {(int(c["step"]), int(c["z_index"]), int(c["r_index"]), int(c["stage"]))
            for c in oracle_cands if not c["ambiguous"]}
    amb = {(int(c["step"]), int(c["z_index"]), int(c["r_index"]), int(c["stage"]))
           for c in oracle_cands if c["ambiguous"]}
    return want, amb


def _got(cands, trial=0):
    return {(int(c["step"]), int(c["z_index"]), int(c["r_index"]), int(c["stage"]))
            for c in cands if int(c["trial"]) == trial}


def test_production_trial_fp32_candidates(fd, oracle):
    """2^23 samples, zmax 200, 8 harmonics, fftlen 8192, every step."""
    from gpu_accelsearch_amd import synth
    n = 1 << 23
    x = synth.production_series(n)
    print_case("whole trial fp32")
    P = fd.SearchParams(n=n, zmax=200, numharm=8, fftlen=8192, powcut=POWCUT)
    with fd.AccelSearch(P) as s:
        s.search(x)
        got = _got(s.candidates())
        nsteps = s.nsteps
    ref = oracle.Search(x, n, 6.4e-5, 200, 8, 8192)
    assert ref.nsteps == nsteps
    want, amb = _sets(ref.candidates(POWCUT, 1e-4))
    ref.close()
    print(f"WHOLE TRIAL fp32: {nsteps} steps, oracle {len(want)} candidates "
          f"(+{len(amb)} within 1e-4 of a cut), found {len(got)}, "
          f"agree {got - amb == want}")
    assert len(want) > 0
    assert got - amb == want


def test_batched_reduced_precision_candidates(fd, oracle):
    """Two 2^23-sample trials in one I16 batch at fftlen 4096, replayed from a
    captured hipGraph; each trial's candidates against the oracle at that fftlen."""
    from gpu_accelsearch_amd import synth
    n = 1 << 23
    xs = [synth.production_series(n),
          synth.production_series(n, seed=5, pulsars=((90000.6, -40.0, 13.0), (333000.2, 70.0, 11.0)))]
    print_case("whole trial batched i16 fftlen 4096")
    P = fd.SearchParams(n=n, zmax=200, numharm=8, fftlen=4096, precision="i16", powcut=POWCUT)
    with fd.BatchSearch(P, [n, n], graph=True) as b:
        for t, x in enumerate(xs):
            b.upload(t, x)
        b.run()          # eager run + capture
        b.run()          # replay
        cands = b.candidates()
    for t, x in enumerate(xs):
        ref = oracle.Search(x, n, 6.4e-5, 200, 8, 4096)
        want, amb = _sets(ref.candidates(POWCUT, 1e-3))
        ref.close()
        got = _got(cands, t)
        print(f"WHOLE TRIAL batched i16 trial {t}: oracle {len(want)} candidates "
              f"(+{len(amb)} within 1e-3 of a cut), found {len(got)}, agree {got - amb == want}")
        assert len(want) > 0
        assert got - amb == want
