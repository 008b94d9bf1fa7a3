"""Fresnel integrals and templates of the oracle, pinned against scipy, mpmath,
the reference's own fresnl() (golden vectors) and the response integral."""
import math

import mpmath as mp
import numpy as np
import pytest
from scipy.special import fresnel as sp_fresnel


def test_fresnel_vs_mpmath(oracle):
    mp.mp.dps = 30
    for x in list(np.linspace(-6, 6, 241)) + [1.4999, 1.5, 1.5001, 7.3, 19.9, 40.0, -33.3]:
        s, c = oracle.fresnel(float(x))
        assert abs(s - float(mp.fresnels(x))) < 1e-14
        assert abs(c - float(mp.fresnelc(x))) < 1e-14


def test_fresnel_vs_scipy(oracle):
    xs = np.linspace(-60, 60, 6001)
    S, Cc = sp_fresnel(xs)
    got = np.array([oracle.fresnel(float(x)) for x in xs])
    # scipy's rational approximations are good to ~1e-12 at large |x|
    assert np.max(np.abs(got[:, 0] - S)) < 2e-12
    assert np.max(np.abs(got[:, 1] - Cc)) < 2e-12


def test_fresnel_vs_reference_golden(oracle, golden):
    """Agreement with the reference's own fresnl() (source/bpd_fresnl.c:160-200),
    including its saturation at |x| > 36974."""
    x, s_ref, c_ref = golden["fresnl_x"], golden["fresnl_s"], golden["fresnl_c"]
    got = np.array([oracle.fresnel(float(v)) for v in x])
    assert np.max(np.abs(got[:, 0] - s_ref)) < 5e-12
    assert np.max(np.abs(got[:, 1] - c_ref)) < 5e-12
    big = np.abs(x) > 36974
    assert np.all(np.abs(got[big]) == 0.5)


def _response_integral(r, z):
    f = lambda u: mp.exp(2j * mp.pi * (r * u + z * u * u / 2))
    return complex(mp.quad(f, [0, 0.25, 0.5, 0.75, 1]))


@pytest.mark.parametrize("z", [-200, -24, -2, 2, 10, 150, 300])
def test_z_response_is_the_chirp_integral(oracle, z):
    """Tap i of the drift-z response equals int_0^1 exp(2 pi i (r u + z u^2/2)) du at
    r = numkern/4 - z/2 - i/2 (roffset 0, two points per bin)."""
    mp.mp.dps = 20
    hw = oracle.lib().or_z_response_halfwidth(z)
    nk = 4 * hw
    resp = oracle.z_response(z, nk)
    for i in sorted({0, 1, nk // 4, nk // 2 - 1, nk // 2, nk // 2 + 1, 3 * nk // 4, nk - 1}):
        r = nk / 4 - z / 2 - i / 2
        assert abs(resp[i] - _response_integral(r, z)) < 1e-11, (z, i)


def test_r_response_closed_form(oracle):
    """z = 0: int_0^1 exp(2 pi i r u) du = e^{i pi r} sin(pi r)/(pi r); centre tap 1."""
    nk = 64
    resp = oracle.z_response(0.0, nk)
    r = nk / 4 - 0.5 * np.arange(nk)
    with np.errstate(invalid="ignore", divide="ignore"):
        want = np.exp(1j * np.pi * r) * np.sin(np.pi * r) / (np.pi * r)
    want[nk // 2] = 1.0
    assert np.max(np.abs(resp - want)) < 1e-12
    assert resp[nk // 2] == 1.0


def test_small_z_continuity(oracle):
    """The response is continuous in z: z = 1e-3 is close to the z = 0 kernel."""
    nk = 64
    a = oracle.z_response(0.0, nk)
    b = oracle.z_response(1e-3, nk)
    assert np.max(np.abs(a - b)) < 5e-3


@pytest.mark.parametrize("z,L", [(0, 1024), (-50, 2048), (200, 8192), (-176, 8192)])
def test_template_row_layout(oracle, z, L):
    """source/bpd_subharmonics.cu:259-265: taps[2hw:] at the start, taps[:2hw] at the
    end, zeros between, values rounded to float."""
    hw = oracle.lib().or_z_response_halfwidth(z)
    taps = oracle.z_response(z, 4 * hw)
    row = oracle.template_row(z, L)
    half = 2 * hw
    f32 = lambda a: a.real.astype(np.float32) + 1j * a.imag.astype(np.float32)
    np.testing.assert_array_equal(row[:half], f32(taps[half:]))
    np.testing.assert_array_equal(row[L - half:], f32(taps[:half]))
    assert np.all(row[half:L - half] == 0)


def test_rz_response_quadrature(oracle):
    """The optimiser's response R(q, z) = int_0^1 exp(2 pi i (q u + z u^2/2)) du:
    the oracle's Gauss-Legendre quadrature against mpmath, and against the
    Fresnel-form template taps (tap i = R(numkern/4 - z/2 - i/2, z))."""
    import mpmath as mp
    for q, z in ((0.3, 0.0), (2.7, -3.0), (-10.2, 40.0), (5.5, 150.5), (-300.2, 1100.0)):
        v = oracle.rz_response(q, z)
        w = complex(mp.quad(lambda u: mp.expj(2 * mp.pi * (q * u + z * u * u / 2)),
                            mp.linspace(0, 1, 200)))
        assert abs(v - w) < 1e-13, (q, z, v, w)
    for z in (-24.0, 2.0, 100.0, -600.0):
        hw = oracle.lib().or_z_response_halfwidth(z)
        nk = 4 * hw
        taps = oracle.z_response(z, nk)
        for i in range(0, nk, max(1, nk // 17)):
            want = oracle.rz_response(nk / 4 - z / 2 - i / 2, z)
            assert abs(taps[i] - want) < 1e-12, (z, i)
