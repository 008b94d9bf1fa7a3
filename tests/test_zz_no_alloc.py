"""fdas_deredden_noalloc: whitening of device bins with its median scratch in a
static device array (no allocation in the call), against the existing deredden
entry point (host bins, allocating) -- bit-identical, on consecutive calls where
a later call is smaller than the one before (the scratch keeps the previous
call's contents), and against the CPU oracle at the smallest size."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

MAX_BINS = 1 << 26


@pytest.fixture(scope="module")
def fd():
    import gpu_accelsearch_amd as fd
    fd.lib()
    return fd


def red_spectrum(nbins, seed):
    rng = np.random.default_rng(seed)
    f = np.arange(nbins) + 1.0
    amp = np.sqrt(1.0 + 50.0 * (f / 30.0) ** -2)
    return (amp * (rng.standard_normal(nbins) + 1j * rng.standard_normal(nbins))).astype(np.complex64)


def noalloc(fd, spec):
    d = fd.DeviceArray.from_host(spec)
    rc = fd.lib().fdas_deredden_noalloc(d.ptr, len(spec), None)
    assert rc == 0, fd.last_error()
    fd.synchronize()
    out = d.to_host()
    d.free()
    return out


@pytest.mark.parametrize("sizes", [((1 << 22) + 1234, 300_001, 4097), (MAX_BINS, 1 << 20)])
def test_noalloc_matches_deredden(fd, oracle, sizes):
    for k, nbins in enumerate(sizes):
        print_case("deredden_noalloc", nbins)
        spec = red_spectrum(nbins, 100 + k)
        got = noalloc(fd, spec)
        want = fd.deredden(spec, nbins)
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
        if nbins <= 300_001:
            ref = oracle.deredden(spec, nbins)
            np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_noalloc_rejects_sizes(fd):
    d = fd.DeviceArray((16,), np.complex64)
    for nbins in (15, MAX_BINS + 1):
        assert fd.lib().fdas_deredden_noalloc(d.ptr, nbins, None) == 1
        assert "bins outside" in fd.last_error()
    assert fd.lib().fdas_deredden_noalloc(None, 100, None) == 1
    d.free()
