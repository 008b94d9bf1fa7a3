"""Autotuner: every correlation launch shape agrees with the oracle (1e-4 of
max(P_ref, 1)) and, in reduced-precision storage, with the default shape; the
tuner's choices apply to later plans, and a plan keeps the shapes in effect when
it was made."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

POWER_TOL = 1e-4
ALL_LENGTHS = (256, 512, 1024, 2048, 4096, 8192, 16384, 768, 1536, 3072, 6144, 12288, 1280, 2560,
               5120, 10240, 7168)


@pytest.fixture(scope="module")
def fd():
    import gpu_accelsearch_amd as fd
    fd.lib()
    yield fd
    for L in ALL_LENGTHS:
        fd.lib().fdas_set_corr_variant(L, 0)


@pytest.mark.parametrize("n,zmax,numharm,fftlen", [
    (1 << 20, 50, 2, 2048),     # correlation lengths 2048 and 1024
    (1 << 20, 200, 2, 8192),    # 8192 and 4096
    (1 << 21, 200, 1, 16384),   # 16384
    (1 << 20, 200, 8, 8192),    # 8192, 7168, 6144, 5120, 4096, 2560, 1280
    (1 << 20, 200, 8, 4096),    # 4096, 3072, 2560, 2048, 1280, 768
    (1 << 21, 600, 4, 16384),   # 16384, 10240, 4096
    (1 << 20, 50, 16, 2048),    # 2048, 1536, 1280, 1024, 768, 512, 256
])
def test_every_corr_variant_matches_oracle(fd, oracle, n, zmax, numharm, fftlen):
    from gpu_accelsearch_amd import synth
    x = synth.plain_series(n, seed=21, r0=30000.3, z0=zmax / 2, power=80.0)
    ref = oracle.Search(x, n, 6.4e-5, zmax, numharm, fftlen)
    L = fd.lib()
    P = fd.SearchParams(n=n, zmax=zmax, numharm=numharm, fftlen=fftlen)
    try:
        with fd.AccelSearch(P) as s:
            lens = sorted({hm[2] for hm in s.harmonics})
            steps = [0, s.nsteps // 2]
        nv = max(L.fdas_corr_variant_count(l) for l in lens)
        want = [ref.step_plane(st) for st in steps]
        for v in range(nv):
            for l in lens:
                assert L.fdas_set_corr_variant(l, min(v, L.fdas_corr_variant_count(l) - 1)) == 0
            print_case("corr variant", v, "lengths", lens)
            with fd.AccelSearch(P) as s:   # a plan takes the shapes in effect at creation
                s.prepare(x)
                for st, w in zip(steps, want):
                    got = s.run_steps(st, 1)[0]
                    e = float(np.max(np.abs(got - w) / np.maximum(w, 1.0)))
                    assert e <= POWER_TOL, (v, st, e)
    finally:
        for l in ALL_LENGTHS:
            L.fdas_set_corr_variant(l, 0)
        ref.close()


def test_autotune_selects_and_applies(fd):
    p = fd.SearchParams(n=1 << 20, zmax=50, numharm=4, fftlen=0)
    print_case("autotune")
    rep = fd.autotune(p, fftlens=(1024, 2048, 4096))
    print(rep)
    assert rep["fftlen"] in (1024, 2048, 4096)
    times = rep["ms_per_trial"]
    assert all(times[k] > 0 for k in (1024, 2048, 4096))
    assert times[rep["fftlen"]] == min(times.values())
    assert fd.resolve_fftlen(p) == rep["fftlen"]
    for l, c in rep["corr"].items():
        us = c["us_per_variant"]
        assert fd.lib().fdas_get_corr_variant(l) == c["variant"]
        if len(us) > 1:   # the shape in use stays unless another is 3% faster
            assert 0.97 * us[c["variant"]] <= min(us)
        # split form only at powers of two, and reported as LDS form 1 there
        assert c["lds_form"] in (0, 1, 2, 4, 8)
        assert c["split"] == (c["lds_form"] == 1 and l & (l - 1) == 0)
    with fd.AccelSearch(p) as s:
        assert s.info.fftlen[0] == rep["fftlen"]


def test_every_harmsum_shape_is_bit_identical(fd):
    """The harmonic sum adds harmonics per cell in the same order whatever the
    tile shape, so every shape gives the same bits (planes and candidates)."""
    from gpu_accelsearch_amd import synth
    n = 1 << 20
    x = synth.plain_series(n, seed=23, r0=31000.3, z0=-12.0, power=90.0)
    L = fd.lib()
    P = fd.SearchParams(n=n, zmax=100, numharm=16, fftlen=4096, powcut=(30.0, 35.0, 45.0, 60.0, 80.0))
    ref = None
    default = L.fdas_get_harmsum_variant()
    try:
        for v in range(L.fdas_harmsum_variant_count()):
            assert L.fdas_set_harmsum_variant(v) == 0
            print_case("harmsum shape", v)
            with fd.AccelSearch(P) as s:
                s.prepare(x)
                planes = s.run_steps(0, s.nsteps)
                s.search(x)
                cands = s.candidates()
            key = (planes.view(np.uint32), sorted(map(tuple, cands.tolist())))
            if ref is None:
                ref = key
                assert len(key[1]) > 0
            else:
                np.testing.assert_array_equal(key[0], ref[0])
                assert key[1] == ref[1]
    finally:
        L.fdas_set_harmsum_variant(default)


def test_plan_keeps_its_launch_shapes(fd):
    """Changing the process-wide shapes leaves a plan's kernels (and so its bits)
    as they were when it was made; a plan made afterwards takes the new shapes."""
    from gpu_accelsearch_amd import synth
    n = 1 << 20
    x = synth.plain_series(n, seed=29, r0=21000.3, z0=8.0, power=80.0)
    L = fd.lib()
    P = fd.SearchParams(n=n, zmax=200, numharm=8, fftlen=8192)
    print_case("plan keeps its shapes")
    try:
        L.fdas_set_corr_variant(8192, 0)
        L.fdas_set_harmsum_variant(0)
        other_shape = next(v for v, sh in enumerate(fd.corr_variant_shapes(8192))
                           if sh["threads"] == 1024)   # 1024 threads x 8 values
        with fd.AccelSearch(P) as a:
            a.prepare(x)
            before = a.run_steps(0, 2).copy()
            L.fdas_set_corr_variant(8192, other_shape)
            L.fdas_set_harmsum_variant(6)
            after = a.run_steps(0, 2)
            np.testing.assert_array_equal(before.view(np.uint32), after.view(np.uint32))
            with fd.AccelSearch(P) as b:
                b.prepare(x)
                other = b.run_steps(0, 2)
            # a different transform shape rounds differently in FP32
            assert not np.array_equal(before.view(np.uint32), other.view(np.uint32))
            assert float(np.max(np.abs(other - before) / np.maximum(before, 1.0))) <= 1e-5
    finally:
        L.fdas_set_corr_variant(8192, 0)
        L.fdas_set_harmsum_variant(0)


# Oracle bounds of the reduced-precision storage modes: the 1e-3 tolerance for
# I16, the stated limits of FP16 and BF16 (as the precision tests)
PREC_TOL = {"fp16": 1.8e-3, "bf16": 1.4e-2, "i16": 1e-3}


@pytest.mark.parametrize("precision", ["bf16", "fp16", "i16"])
@pytest.mark.parametrize("fftlen", [8192, 4096])
def test_every_corr_variant_reduced_precision(fd, oracle, precision, fftlen):
    """In reduced-precision storage every correlation shape reads the same stored
    template rows and segments and computes in FP32, so its powers equal the
    default shape's to FP32 rounding; stored as P16 (relative step 2^-10) such a
    difference can move a stored power by one step, so the summed planes agree
    within 2^-10 of max(P, 1) plus the FP32 rounding (2e-5).  The split form
    stores its own packed bank (sums and differences of template values), so it
    is held to the mode's oracle tolerance instead, as the default shape is."""
    from gpu_accelsearch_amd import synth
    n = 1 << 20
    x = synth.plain_series(n, seed=22, r0=30000.3, z0=100, power=80.0)
    ref = oracle.Search(x, n, 6.4e-5, 200, 8, fftlen)
    L = fd.lib()
    try:
        P = fd.SearchParams(n=n, zmax=200, numharm=8, fftlen=fftlen, precision=precision)
        with fd.AccelSearch(P) as s:
            lens = sorted({hm[2] for hm in s.harmonics})
            steps = [0, s.nsteps // 2]
        shapes = {l: fd.corr_variant_shapes(l) for l in lens}
        nv = max(len(v) for v in shapes.values())
        want = [ref.step_plane(st) for st in steps]
        for l in lens:
            L.fdas_set_corr_variant(l, 0)
        with fd.AccelSearch(P) as s:
            s.prepare(x)
            base = [s.run_steps(st, 1)[0].copy() for st in steps]
        for b, w in zip(base, want):
            e = float(np.max(np.abs(b - w) / np.maximum(w, 1.0)))
            print(f"TUNE {precision} {fftlen} default shape oracle error {e:.3e}")
            assert e <= PREC_TOL[precision]
        for v in range(1, nv):
            split = False
            for l in lens:
                vl = min(v, len(shapes[l]) - 1)
                split |= shapes[l][vl]["split"]
                assert L.fdas_set_corr_variant(l, vl) == 0
            print_case("corr variant", v, precision, "lengths", lens, "split" if split else "")
            with fd.AccelSearch(P) as s:
                s.prepare(x)
                for st, b, w in zip(steps, base, want):
                    got = s.run_steps(st, 1)[0]
                    if split:
                        e = float(np.max(np.abs(got - w) / np.maximum(w, 1.0)))
                        print(f"TUNE {precision} {fftlen} split variant {v} oracle error {e:.3e}")
                        assert e <= PREC_TOL[precision], (v, st, e)
                    else:
                        e = float(np.max(np.abs(got - b) / np.maximum(b, 1.0)))
                        assert e <= 2.0 ** -10 + 2e-5, (v, st, e)
    finally:
        for l in ALL_LENGTHS:
            L.fdas_set_corr_variant(l, 0)
        ref.close()
