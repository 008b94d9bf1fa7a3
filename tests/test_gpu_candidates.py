"""Candidate optimisation on the GPU: the response it maximises against the
oracle's quadrature and against the search's own plane, and recovery of
injected signals (build-defined stage: the reference's optimiser is a stub,
source/bpd_candidates.cu:22-26)."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fd():
    import gpu_accelsearch_amd as fd
    fd.lib()
    return fd


def test_rz_power_matches_oracle_quadrature(fd, oracle):
    from gpu_accelsearch_amd import synth
    n = 1 << 20
    x = synth.plain_series(n, seed=31, r0=30000.37, z0=5.3, power=150.0)
    print_case("rz power vs oracle")
    with fd.AccelSearch(fd.SearchParams(n=n, zmax=50, numharm=1, fftlen=2048)) as s:
        s.prepare(x)
        spec = s.spectrum()
        rng = np.random.default_rng(3)
        r = np.concatenate([rng.uniform(1000, 500000, 30), 30000.37 + rng.uniform(-2, 2, 10), [1234.0, 2000.5]])
        z = np.concatenate([rng.uniform(-50, 50, 30), rng.uniform(-0.09, 0.09, 5),
                            5.3 + rng.uniform(-1, 1, 5), [0.0, -12.0]])
        got = s.rz_power(r, z)
        want = np.array([oracle.rz_power(spec, a, b) for a, b in zip(r, z)])
        err = np.abs(got - want) / np.maximum(want, 1e-3 * want.mean())
        assert err.max() < 1e-6, (err.max(), r[err.argmax()], z[err.argmax()])


def test_rz_power_matches_search_plane(fd):
    """At search cells, plane = (segment normalisation)^2 x response power."""
    from gpu_accelsearch_amd import synth
    n = 1 << 20
    x = synth.plain_series(n, seed=32, r0=60000.2, z0=-14.0, power=150.0)
    print_case("rz power vs plane")
    with fd.AccelSearch(fd.SearchParams(n=n, zmax=50, numharm=1, fftlen=2048)) as s:
        s.prepare(x)
        for st in (5, synth.steps_covering(60000.2, s.rlo, s.uselen)):
            plane = s.run_steps(st, 1)[0]
            j, i = np.nonzero(plane > 2.0)
            sel = np.random.default_rng(st).choice(len(j), 60, replace=False)
            j, i = j[sel], i[sel]
            r = s.rlo + st * (s.uselen // 2) + 0.5 * i
            z = -50.0 + 2.0 * j
            ratio = plane[j, i] / s.rz_power(r, z)
            assert np.ptp(ratio) / ratio.mean() < 1e-4, (st, ratio.min(), ratio.max())


def test_optimizer_recovers_injected_signal(fd):
    from gpu_accelsearch_amd import synth
    n = 1 << 20
    r0, z0 = 30000.37, 5.3
    x = synth.plain_series(n, seed=33, r0=r0, z0=z0, power=150.0)
    print_case("optimiser single harmonic")
    with fd.AccelSearch(fd.SearchParams(n=n, zmax=50, numharm=1, fftlen=2048,
                                        powcut=(30.0, 0, 0, 0, 0))) as s:
        s.search(x)
        acc = s.sift(sigma_min=3.0)
        assert len(acc) >= 1
        top = acc[0]
        assert abs(top["r"] - r0) < 0.5 and abs(top["z"] - z0) < 2.0
        opt = s.optimize(acc[:1])[0]
        print("grid", top["r"], top["z"], top["power"], "opt", opt["opt_r"], opt["opt_z"], opt["opt_power"])
        assert abs(opt["opt_r"] - r0) < 0.05 and abs(opt["opt_z"] - z0) < 0.6
        assert opt["opt_power"] >= 0.97 * top["power"]
        assert opt["harm_power"][0] == pytest.approx(opt["opt_power"], rel=1e-5)
        T = n * 6.4e-5
        assert opt["freq"] == pytest.approx(opt["opt_r"] / T)
        assert opt["opt_sigma"] > 3.0


def test_pulse_train_sift_and_optimise(fd):
    """8 harmonics of a pulse train: the sifted list keeps the fundamental and
    drops the harmonics; the optimiser refines it using all 8."""
    from gpu_accelsearch_amd import synth
    n = 1 << 21
    r1, z1 = 12000.3, 3.0
    rng = np.random.default_rng(34)
    x = (rng.standard_normal(n) + synth.pulse_train(n, r1, z1, 0.05, 40.0)).astype(np.float32)
    print_case("pulse train sift")
    P = fd.SearchParams(n=n, zmax=50, numharm=8, fftlen=2048)
    with fd.AccelSearch(P) as s0:
        powcut = s0.stage_powcuts(6.0)
    for st, pc in enumerate(powcut[:4]):
        ni = fd.numindep(s0.rlo, s0.nbins, s0.numz, st)
        assert abs(fd.candidate_sigma(pc, 1 << st, ni) - 6.0) < 1e-6
    P.powcut = powcut
    with fd.AccelSearch(P) as s:
        s.search(x)
        acc = s.sift(sigma_min=6.0)
        assert len(acc) >= 1
        top = acc[0]
        assert top["numharm"] >= 4 and abs(top["r"] - r1) < 0.3, acc[:3][["r", "z", "numharm", "sigma"]]
        for k in range(2, 9):
            assert not np.any((np.abs(acc["r"] - k * r1) < 1.0)), k
        opt = s.optimize(acc[:1])[0]
        print("pulse", top["r"], top["z"], top["numharm"], "->", opt["opt_r"], opt["opt_z"],
              opt["opt_power"], opt["harm_power"][:8])
        assert abs(opt["opt_r"] - r1) < 0.02 and abs(opt["opt_z"] - z1) < 0.3
        assert opt["opt_power"] >= 0.97 * top["power"]


def test_candidate_overflow_is_a_deterministic_error(fd):
    """More cells above the thresholds than the plan holds: no partial list (which
    cells got stored would depend on the order the GPU found them in), the same
    error and the same count, the true number found, on every run."""
    import ctypes as C
    from gpu_accelsearch_amd import synth
    n = 1 << 20
    x = synth.plain_series(n, seed=41, r0=20000.3, z0=4.0, power=60.0)
    powcut = (6.0, 9.0, 14.0, 0.0, 0.0)
    print_case("candidate overflow")
    with fd.AccelSearch(fd.SearchParams(n=n, zmax=50, numharm=4, fftlen=2048, powcut=powcut)) as s:
        s.search(x)
        total = len(s.candidates())
    assert total > 1000
    L = fd.lib()
    P = fd.SearchParams(n=n, zmax=50, numharm=4, fftlen=2048, powcut=powcut, max_candidates=100)
    with fd.AccelSearch(P) as s:
        for _ in range(2):
            s.search(x)
            cnt, ov = C.c_int(), C.c_int()
            buf = (fd._native.FdasCandidate * 100)()
            rc = L.fdas_get_candidates(s.handle, buf, 100, C.byref(cnt), C.byref(ov))
            assert rc == 5 and ov.value == 1 and cnt.value == total
            assert "exceed the capacity" in fd.last_error()
            with pytest.raises(fd.FdasError, match="candidate capacity exceeded"):
                s.candidates()
