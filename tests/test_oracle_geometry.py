"""Search geometry of the oracle and of libfdas.so's host side, pinned by
hand-derived tables and the reference's own known answers."""
import math

import numpy as np
import pytest

# (U, zmax, fundamental FFT length) -- the USELEN comments of headers/bpd_common.h:14-28
REFERENCE_USELEN_KATS = [
    (32000, 300, 32768), (15660, 300, 16384), (7470, 300, 8192), (7960, 100, 8192),
    (3850, 100, 4096), (1820, 100, 2048), (3370, 300, 4096), (1320, 300, 2048),
]

# Production workload: zmax 200, fftlen 8192 -> U 7734; per harmonic
# [1/1,1/2,1/4,3/4,1/8,3/8,5/8,7/8]: zmax_h, numzs, halfwidth of the widest template, L_h
C2_TABLE = {
    "frac": [(1, 1), (1, 2), (1, 4), (3, 4), (1, 8), (3, 8), (5, 8), (7, 8)],
    "zmax": [200, 100, 50, 150, 26, 76, 126, 176],
    "numzs": [201, 101, 51, 151, 27, 77, 127, 177],
    "hw": [114, 56, 33, 82, 24, 44, 69, 98],
    "L": [8192, 4096, 4096, 8192, 2048, 4096, 8192, 8192],
}


def test_uselen_kats(oracle):
    L = oracle.lib()
    for U, zmax, fl in REFERENCE_USELEN_KATS:
        assert L.or_fft_length(U, 1, 1, zmax) == fl, (U, zmax)


def test_halfwidth_kats(oracle):
    L = oracle.lib()
    assert [L.or_z_response_halfwidth(z) for z in (0, 2, 50, 200, 300, 600, -200)] == \
        [16, 16, 33, 114, 180, 360, 114]


def test_nearest_int_rounds_half_away(oracle):
    L = oracle.lib()
    assert [L.or_nearest_int(v) for v in (0.5, 1.5, 2.5, -0.5, -1.5, 0.49, -0.49)] == \
        [1, 2, 3, -1, -2, 0, 0]
    assert L.or_calc_required_z(0.5, 2) == 2 and L.or_calc_required_z(0.5, -2) == -2
    assert L.or_calc_required_z(0.125, 200) == 26


def test_production_table(oracle):
    L = oracle.lib()
    assert L.or_max_uselen(8192, 200) == 7734
    assert oracle.harmonic_order(8) == C2_TABLE["frac"]
    for k, (c, m) in enumerate(C2_TABLE["frac"]):
        zh = L.or_calc_required_z(c / m, 200)
        assert zh == C2_TABLE["zmax"][k]
        assert (zh // 2) * 2 + 1 == C2_TABLE["numzs"][k]
        assert L.or_z_response_halfwidth(-zh) == C2_TABLE["hw"][k]
        assert L.or_fft_length(7734, m, c, 200) == C2_TABLE["L"][k]


@pytest.mark.parametrize("n,zmax,fftlen,uselen,rlo,nsteps", [
    (1 << 20, 50, 2048, 1914, 67, 547),
    (1 << 23, 200, 8192, 7734, 536, 1084),
    (1 << 24, 300, 8192, 7470, 1073, 2245),
])
def test_workload_geometry(oracle, n, zmax, fftlen, uselen, rlo, nsteps):
    L = oracle.lib()
    assert L.or_max_uselen(fftlen, zmax) == uselen
    assert math.floor(float(np.float32(n) * np.float32(6.4e-5))) == rlo
    assert L.or_num_steps(n // 2, rlo, uselen) == nsteps


def test_segment_lengths_round_to_half_fft(oracle):
    """nice_numdata == fftlen/2 for every harmonic at the first, middle and last step."""
    L = oracle.lib()
    for n, zmax, fftlen, nh in ((1 << 20, 50, 2048, 4), (1 << 23, 200, 8192, 16)):
        U = L.or_max_uselen(fftlen, zmax)
        rlo = math.floor(float(np.float32(n) * np.float32(6.4e-5)))
        ns = L.or_num_steps(n // 2, rlo, U)
        for st in (0, ns // 2, ns - 1):
            for g in oracle.step_geometry(nh, zmax, U, rlo + st * (U // 2)):
                assert L.or_next_pow2(g["numdata"]) == g["fftlen"] // 2


def test_product_geometry_matches_oracle(oracle):
    """libfdas.so's exported host geometry equals the oracle on a grid."""
    import gpu_accelsearch_amd as fd
    L = oracle.lib()
    for f in (1.0, 0.5, 0.25, 0.75, 0.125, 0.375, 0.625, 0.875, 1 / 16, 15 / 16):
        for z in range(-1200, 1201, 7):
            assert fd.calc_required_z(f, z) == L.or_calc_required_z(f, z)
        for r in np.linspace(0, 4.2e6, 257):
            assert fd.calc_required_r(f, r) == L.or_calc_required_r(f, r)
    for z in range(-1200, 1201, 3):
        assert fd.z_response_halfwidth(z) == L.or_z_response_halfwidth(z)
    assert fd.index_from_r(100.5, 90.0) == L.or_index_from_r(100.5, 90.0) == 21
    assert fd.index_from_z(10.0, -20.0) == L.or_index_from_z(10.0, -20.0) == 15
    assert fd.getTotalHarmonics(4) == 8


def test_harmonic_sum_index_maps(oracle):
    """Reference index maps (source/bpd_gpu_kernels.cu:139-153) stay in the planes,
    r indices are non-decreasing across columns, and the integer forms used by the
    fused kernel equal the FP64 forms."""
    L = oracle.lib()
    for n, zmax, fftlen, nh in ((1 << 23, 200, 8192, 16), (1 << 20, 50, 2048, 8)):
        U = L.or_max_uselen(fftlen, zmax)
        rlo = math.floor(float(np.float32(n) * np.float32(6.4e-5)))
        ns = L.or_num_steps(n // 2, rlo, U)
        for st in (0, 1, ns // 3, ns - 1):
            startr = rlo + st * (U // 2)
            g = oracle.step_geometry(nh, zmax, U, startr)
            fullrlo = int(startr)
            i = np.arange(U)
            for h in range(1, nh):
                f = g[h]["cur"] / g[h]["max"]
                rr = fullrlo + i * 0.5
                subr = np.floor(2 * rr * f + 0.5) * 0.5
                rind = ((subr - g[h]["rlo"]) * 2).astype(np.int64)
                lm = int(math.log2(g[h]["max"]))
                q = 2 * fullrlo + i
                rind_int = ((q * g[h]["cur"] + (1 << (lm - 1))) >> lm) - 2 * g[h]["rlo"]
                np.testing.assert_array_equal(rind, rind_int)
                assert rind.min() >= 0 and rind.max() < g[h]["numrs"]
                assert np.all(np.diff(rind) >= 0)
                zz = g[0]["zlo"] + 2 * np.arange(g[0]["numzs"])
                subz = np.trunc(zz * f).astype(np.int64)
                zind = ((subz - g[h]["zlo"]) * 0.5).astype(np.int64)
                subz_int = np.array([int(a * g[h]["cur"] / (1 << lm)) for a in zz])
                zind_int = (subz_int - g[h]["zlo"]) >> 1
                np.testing.assert_array_equal(zind, zind_int)
                assert zind.min() >= 0 and zind.max() < g[h]["numzs"]
