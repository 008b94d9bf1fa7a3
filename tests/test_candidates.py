"""Candidate significance and sifting (host logic, no GPU).

The reference leaves candidate search, sorting, harmonic elimination and the
sigma/powcut/numindep settings as stubs (source/bpd_candidates.cu:4-33,
AccelSearch.cu:95-97), so this behaviour is build-defined; its statistics are
pinned by scipy/mpmath (Gamma(n, 1) tails, Gaussian equivalents), its geometry
by construction."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def fd():
    import gpu_accelsearch_amd as fd
    fd.lib()
    return fd


def test_sigma_matches_scipy(fd):
    from scipy.special import gammaincc, ndtri
    for n in (1, 2, 4, 8, 16):
        for P in (15.0, 25.0, 40.0, 60.0, 100.0):
            for N in (1.0, 1e3, 1e9):
                p1 = gammaincc(n, P)
                pt = -np.expm1(N * np.log1p(-p1))
                if not (1e-290 < pt < 0.999):
                    continue
                want = -ndtri(pt)
                got = fd.candidate_sigma(P, n, N)
                assert abs(got - want) <= 1e-9 * max(1.0, abs(want)), (n, P, N, got, want)


def test_sigma_extreme_tails_match_mpmath(fd):
    import mpmath as mp
    mp.mp.dps = 50
    for n, P, N in ((1, 800.0, 1.0), (1, 2000.0, 1e6), (8, 1500.0, 1e10), (16, 3000.0, 1e12)):
        p = mp.gammainc(n, P, mp.inf, regularized=True) * N
        s = mp.findroot(lambda x: mp.log(mp.erfc(x / mp.sqrt(2)) / 2) - mp.log(p),
                        mp.sqrt(-2 * mp.log(p)))
        got = fd.candidate_sigma(P, n, N)
        assert abs(got - float(s)) < 1e-9 * float(s), (n, P, N, got, float(s))


def test_power_for_sigma_inverts(fd):
    for n in (1, 2, 4, 8, 16):
        for sig in (2.0, 5.0, 8.0, 20.0):
            for N in (1.0, 1e6, 1e10):
                P = fd.power_for_sigma(sig, n, N)
                assert abs(fd.candidate_sigma(P, n, N) - sig) < 1e-7, (n, sig, N)
    # more harmonics or more trials need more power for the same sigma
    assert fd.power_for_sigma(6, 8, 1e9) > fd.power_for_sigma(6, 4, 1e9) > fd.power_for_sigma(6, 4, 1e6)


def test_numindep_formula(fd):
    assert fd.numindep(100, 100100, 201, 0) == pytest.approx(1e5 * 202 * 2 / 6.95)
    assert fd.numindep(100, 100100, 201, 3) == pytest.approx(1e5 * 202 * 2 / 6.95 / 8)
    assert fd.numindep(0, 5000, 1, 2) == pytest.approx(1250.0)


def _cand(fd, trial, rcol, zcol, stage, power, rlo, uselen, zmax):
    """Search cell of a first-harmonic (r, z) detection at a stage."""
    step = int((rcol - rlo) // (uselen // 2))
    ri = int(round((rcol - rlo - step * (uselen // 2)) / 0.5))
    zi = int(round((zcol + zmax) / 2))
    return (step, zi, ri, stage, power, trial)


def test_sift_collapses_duplicates_and_harmonics(fd):
    rlo, uselen, zmax, numz = 67, 1914, 50, 51
    trials = [(1 << 20, 1 << 19, rlo, 6.4e-5), (1 << 20, 1 << 19, rlo, 6.4e-5)]
    r0, z0 = 20000.5, 4.0
    rows = [
        _cand(fd, 0, r0, z0, 0, 80.0, rlo, uselen, zmax),             # fundamental, stage 0
        _cand(fd, 0, r0 + 0.5, z0, 0, 60.0, rlo, uselen, zmax),       # neighbouring cell
        _cand(fd, 0, 2 * r0, 2 * z0, 1, 120.0, rlo, uselen, zmax),    # same signal, 2 harmonics
        _cand(fd, 0, 2 * r0, 2 * z0, 0, 45.0, rlo, uselen, zmax),     # 2nd harmonic alone
        _cand(fd, 0, 1.5 * r0, 1.5 * z0, 0, 40.0, rlo, uselen, zmax), # 3/2 harmonic
        _cand(fd, 0, 31234.0, -20.0, 0, 50.0, rlo, uselen, zmax),     # unrelated
        _cand(fd, 1, r0, z0, 0, 70.0, rlo, uselen, zmax),             # other trial: kept
        _cand(fd, 0, 41000.0, 10.0, 0, 22.0, rlo, uselen, zmax),      # weak
    ]
    c = np.array(rows, dtype=fd.CANDIDATE_DTYPE)
    out = fd.sift_raw(trials, zmax, uselen, numz, c, sigma_min=3.0)
    assert list(out["sigma"]) == sorted(out["sigma"], reverse=True)
    got = {(int(a["trial"]), round(float(a["r"]), 3)) for a in out}
    assert got == {(0, r0), (1, r0), (0, 31234.0)}, out[["trial", "r", "z", "numharm", "sigma"]]
    top = out[0]
    assert top["numharm"] == 2 and top["trial"] == 0 and top["z"] == pytest.approx(z0)
    # properties from (r, z) and T = n dt
    T = (1 << 20) * 6.4e-5
    assert top["freq"] == pytest.approx(r0 / T) and top["fdot"] == pytest.approx(z0 / T ** 2)
    assert top["period"] == pytest.approx(T / r0)
    # without harmonic elimination the harmonics stay
    out2 = fd.sift_raw(trials, zmax, uselen, numz, c, sigma_min=3.0, max_ratio=-1)
    assert len(out2) == 5
    # with tiny tolerances only exact duplicates merge: the 2-harmonic detection
    # has the same first-harmonic (r, z) as the fundamental's
    n_all = len(fd.sift_raw(trials, zmax, uselen, numz, c, max_ratio=-1, r_tol=1e-3, z_tol=1e-3))
    assert n_all == len(rows) - 1
