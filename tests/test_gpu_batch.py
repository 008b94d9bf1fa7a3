"""Batches of DM trials: one plan over many series of uneven lengths.

A batch computes each (trial, step) item independently of what else shares its
launch, so its planes and candidates must be bit-identical to single-series
plans of the same trials, and those agree with the oracle to 1e-4 (relative to
max(P_ref, 1)).  Graph replay must reproduce the eager run bit for bit, also
after the inputs change between replays.
"""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

POWER_TOL = 1e-4
ZMAX, NUMHARM, FFTLEN = 50, 4, 2048
LENGTHS = [1 << 18, 300_000, (1 << 18) + 2, 250_002, 1 << 19]
POWCUT = (28.0, 36.0, 45.0, 0.0, 0.0)


@pytest.fixture(scope="module")
def fd():
    import gpu_accelsearch_amd as fd
    fd.lib()
    return fd


def _series(t, n, other=0):
    from gpu_accelsearch_amd import synth
    return synth.plain_series(n, seed=40 + t + 100 * other, r0=9000.7 + 1700.3 * t + 611.1 * other,
                              z0=-20.0 + 9.0 * t - 3.0 * other, power=90.0)


def _params(fd, n=0, chunk=0):
    return fd.SearchParams(n=n, zmax=ZMAX, numharm=NUMHARM, fftlen=FFTLEN, chunk_steps=chunk,
                           powcut=POWCUT)


def _cand_set(c):
    return {(int(a["trial"]), int(a["step"]), int(a["z_index"]), int(a["r_index"]),
             int(a["stage"]), float(a["power"])) for a in c}


def test_batch_matches_single_plans_and_oracle(fd, oracle):
    xs = [_series(t, n) for t, n in enumerate(LENGTHS)]
    print_case("batch uneven lengths", LENGTHS)
    # chunk of 37 items: launches mix steps of different trials
    with fd.BatchSearch(_params(fd, chunk=37), LENGTHS) as b:
        assert b.info.ntrials == len(LENGTHS) and b.info.chunk_steps == 37
        for t, x in enumerate(xs):
            b.upload(t, x)
        b.prepare()
        firsts = [tr.first_step for tr in b.trials]
        assert firsts[0] == 0
        want_cands = set()
        for t, (n, x) in enumerate(zip(LENGTHS, xs)):
            tr = b.trials[t]
            with fd.AccelSearch(_params(fd, n=n)) as s:
                assert (tr.n, tr.nbins, tr.rlo, tr.nsteps) == (n, n // 2, s.rlo, s.nsteps)
                s.prepare(x)
                np.testing.assert_array_equal(b.spectrum(t).view(np.uint32),
                                              s.spectrum().view(np.uint32))
                single = s.run_steps(0, s.nsteps)
                batch = b.run_steps(tr.first_step, tr.nsteps)
                np.testing.assert_array_equal(batch.view(np.uint32), single.view(np.uint32))
                s.search(x)
                c = s.candidates()
                c["trial"] = t
                want_cands |= _cand_set(c)
            ref = oracle.Search(x, n, 6.4e-5, ZMAX, NUMHARM, FFTLEN)
            for st in (0, tr.nsteps // 2, tr.nsteps - 1):
                want = ref.step_plane(st)
                e = float(np.max(np.abs(batch[st] - want) / np.maximum(want, 1.0)))
                assert e <= POWER_TOL, (t, st, e)
            ref.close()
        print_case("batch run + candidates")
        b.run()
        got = _cand_set(b.candidates())
        assert not b.overflowed
        assert len(got) > 0 and {g[0] for g in got} == set(range(len(LENGTHS)))
        assert got == want_cands


def test_batch_graph_replay(fd):
    xs = [_series(t, n) for t, n in enumerate(LENGTHS)]
    ys = [_series(t, n, other=1) for t, n in enumerate(LENGTHS)]
    print_case("batch eager reference")
    with fd.BatchSearch(_params(fd, chunk=64), LENGTHS) as e:
        for t, x in enumerate(xs):
            e.upload(t, x)
        e.run()
        cx = _cand_set(e.candidates())
        for t, y in enumerate(ys):
            e.upload(t, y)
        e.run()
        cy = _cand_set(e.candidates())
        planes_y = e.run_steps(0, 20)
    assert cx != cy and len(cx) > 0 and len(cy) > 0
    print_case("batch graph replay")
    with fd.BatchSearch(_params(fd, chunk=64), LENGTHS, graph=True) as g:
        for rep, (data, want) in enumerate(((xs, cx), (xs, cx), (ys, cy), (xs, cx), (ys, cy))):
            for t, x in enumerate(data):
                g.upload(t, x)
            g.run()
            got = _cand_set(g.candidates())
            assert got == want, rep
            g.synchronize()
        # the graph leaves the spectra of the last inputs behind
        np.testing.assert_array_equal(g.run_steps(0, 20).view(np.uint32), planes_y.view(np.uint32))


def test_template_banks_shared(fd):
    print_case("shared template banks")
    a = fd.AccelSearch(_params(fd, n=1 << 18))
    b = fd.BatchSearch(_params(fd), [1 << 18, 1 << 19])
    try:
        assert a.info.template_bytes > 0
        assert a.info.template_bytes == b.info.template_bytes
        # plan-owned buffers do not include the shared banks
        assert b.info.device_bytes > 0
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("precision,tol", [("fp32", POWER_TOL), ("i16", 1e-3)])
def test_batch_workload_numharm16_autotuned(fd, oracle, precision, tol):
    """Uneven lengths log-uniform in [2^21, 2^23], red noise, some trials with a
    pulsar; zmax 200, 16 harmonics, fftlen chosen by the autotuner; FP32 and the
    reduced-precision mode within its 1e-3 power tolerance (I16 segments and
    templates, P16 power planes)."""
    from gpu_accelsearch_amd import synth
    lengths = synth.log_uniform_lengths(10, 1 << 21, 1 << 23, seed=4)
    pool = synth.noise_pool(1 << 23, seed=9)
    xs = [pool[:n].copy() for n in lengths]
    with_pulsar = (2, 7)
    for t in with_pulsar:
        xs[t] += synth.pulse_train(lengths[t], 21000.4 + 3000 * t, 11.0, 0.05, 40.0).astype(np.float32)
    powcut = (35.0, 40.0, 50.0, 65.0, 90.0)
    print_case("batch workload autotune", precision)
    base = fd.SearchParams(n=lengths[0], zmax=200, numharm=16, fftlen=0, powcut=powcut,
                           precision=precision)
    rep = fd.autotune(base, fftlens=(4096, 8192, 16384))
    print("autotune", rep["fftlen"], rep["ms_per_trial"])
    assert rep["fftlen"] in (4096, 8192, 16384)
    print_case("batch workload run", lengths)
    with fd.BatchSearch(base, lengths, graph=True) as b:
        assert b.info.fftlen[0] == rep["fftlen"]
        for t, x in enumerate(xs):
            b.upload(t, x)
        b.run()
        cands = b.candidates()
        b.prepare()
        for t in with_pulsar:
            tr = b.trials[t]
            ref = oracle.Search(xs[t], lengths[t], 6.4e-5, 200, 16, rep["fftlen"])
            st = synth.steps_covering(16 * (21000.4 + 3000 * t), tr.rlo, b.uselen)
            want, stages = ref.step_plane(st, stages=True)
            got = b.run_steps(tr.first_step + st, 1)[0]
            e = float(np.max(np.abs(got - want) / np.maximum(want, 1.0)))
            print(f"{precision} trial {t} step {st} max power error {e:.2e}")
            assert e <= tol
            # candidates of that step match the oracle's stage sums away from the margin
            mine = {(int(c["z_index"]), int(c["r_index"]), int(c["stage"]))
                    for c in cands if int(c["trial"]) == t and int(c["step"]) == st}
            exp, amb = set(), set()
            for sg in range(5):
                z, i = np.nonzero(stages[sg] > powcut[sg])
                exp |= {(int(a), int(b_), sg) for a, b_ in zip(z, i)}
                z, i = np.nonzero(np.abs(stages[sg] - powcut[sg]) <= tol * np.maximum(stages[sg], 1.0))
                amb |= {(int(a), int(b_), sg) for a, b_ in zip(z, i)}
            assert len(exp) > 0 and mine - amb == exp - amb
            ref.close()
        assert {int(c["trial"]) for c in cands} >= set(with_pulsar)


def test_batch_from_presto_files(fd, tmp_path):
    """DM trials read from .dat/.inf files search exactly like in-memory arrays."""
    from gpu_accelsearch_amd import synth
    lengths = [300_001, 262_144, 400_000]
    bases = []
    for t, n in enumerate(lengths):
        x = synth.plain_series(n, seed=60 + t, r0=8000.3 + 900 * t, z0=6.0, power=90.0)
        base = tmp_path / f"trial_DM{t}.00"
        x.tofile(f"{base}.dat")
        with open(f"{base}.inf", "w") as f:
            f.write(f" Data file name without suffix          =  {base.name}\n"
                    f" Number of bins in the time series      =  {n}\n"
                    f" Width of each time series bin (sec)    =  6.4e-05\n"
                    f" Dispersion measure (cm-3 pc)           =  {t}.0\n")
        bases.append(str(base))
    P = fd.SearchParams(n=0, zmax=50, numharm=4, fftlen=2048, powcut=POWCUT)
    print_case("batch from files")
    with fd.batch_from_files(P, bases) as b:
        assert [tr.n for tr in b.trials] == [300_000, 262_144, 400_000]
        assert [i["dm"] for i in b.infos] == [0.0, 1.0, 2.0]
        b.run()
        got = _cand_set(b.candidates())
    xs = [fd.readData(f"{base}.dat")[:n - (n & 1)] for base, n in zip(bases, lengths)]
    with fd.BatchSearch(P, [len(x) for x in xs]) as b2:
        for t, x in enumerate(xs):
            b2.upload(t, x)
        b2.run()
        assert got == _cand_set(b2.candidates()) and len(got) > 0


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_graph_replay_production_shape(fd, precision):
    """Graph replay at the production shape (zmax 200, 8 harmonics, fftlen 8192:
    sub-harmonics at 7168, 6144, 5120, 4096, 2560 and 1280, the two-region LDS
    exchange) reproduces the eager candidates bit for bit on every replay, with
    the inputs changed between replays."""
    from gpu_accelsearch_amd import synth
    lengths = [1 << 21, (1 << 21) - 6]
    xs = [synth.plain_series(n, seed=60 + t, r0=70000.3 + 900.0 * t, z0=-70.0 + 50.0 * t, power=100.0)
          for t, n in enumerate(lengths)]
    ys = [synth.plain_series(n, seed=80 + t, r0=90000.7 - 700.0 * t, z0=60.0 - 30.0 * t, power=100.0)
          for t, n in enumerate(lengths)]
    P = fd.SearchParams(n=0, zmax=200, numharm=8, fftlen=8192, precision=precision,
                        powcut=(40.0, 45.0, 55.0, 70.0, 0.0))
    print_case("production-shape eager", precision)
    with fd.BatchSearch(P, lengths) as e:
        assert list(e.info.fftlen[:8]) == [8192, 4096, 2560, 6144, 1280, 4096, 5120, 7168]
        want = []
        for data in (xs, ys):
            for t, x in enumerate(data):
                e.upload(t, x)
            e.run()
            want.append(_cand_set(e.candidates()))
            assert not e.overflowed
    assert want[0] != want[1] and len(want[0]) > 0 and len(want[1]) > 0
    print_case("production-shape graph replay", precision)
    with fd.BatchSearch(P, lengths, graph=True) as g:
        for rep, k in enumerate((0, 0, 1, 0, 1)):
            for t, x in enumerate((xs, ys)[k]):
                g.upload(t, x)
            g.run()
            assert _cand_set(g.candidates()) == want[k], rep
            g.synchronize()
