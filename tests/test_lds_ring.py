"""The LDS exchange ring of the workgroup FFT (BlockFFT::run_ring in
gpu_accelsearch_amd/csrc/fft_block.hpp) has no race between its exchanges or
between successive calls: each workgroup thread runs as an OS thread with a real
barrier, thread 0 is held back after every barrier so the others run ahead, and
the transforms must still equal numpy's DFT (tests/fft_ring_race.cpp)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SLEEP_US = 1000

# (length, threads, transforms, ring slots, transforms per slot): the shapes the
# correlation kernels use (one transform at 2048/4096/8192; R transforms of
# L/R at 3072, 5120, 6144, 7168, 10240, 12288)
SHAPES = [(2048, 256, 1, 2, 1), (4096, 512, 1, 2, 1), (8192, 512, 1, 2, 1),
          (1024, 128, 3, 2, 1), (1024, 256, 3, 2, 1), (2048, 256, 3, 2, 1), (4096, 512, 3, 2, 1),
          (1024, 128, 5, 2, 1), (2048, 256, 5, 2, 1),
          (1024, 256, 7, 2, 1), (1024, 128, 7, 2, 1), (1024, 256, 7, 2, 2)]


@pytest.fixture(scope="module")
def ring(tmp_path_factory):
    d = tmp_path_factory.mktemp("ring")
    so = str(d / "ring.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-pthread", "-o", so,
                    os.path.join(ROOT, "tests", "fft_ring_race.cpp")], check=True)
    lib = C.CDLL(so)
    lib.ring_race.argtypes = [C.c_int] * 9 + [C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
    return lib


def run(lib, n, t, na, nr, k, sign, plant, calls, seed):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((calls, na, n)) + 1j * rng.standard_normal((calls, na, n))).astype(np.complex64)
    out = np.zeros_like(x)
    lead = C.c_int(-1)
    rc = lib.ring_race(n, t, sign, na, nr, k, plant, calls, SLEEP_US, x.ctypes.data, out.ctypes.data,
                       C.byref(lead))
    assert rc == 0
    xd = x.astype(np.complex128)
    ref = np.fft.fft(xd, axis=-1) if sign < 0 else np.fft.ifft(xd, axis=-1) * n
    err = np.abs(out - ref).max(axis=-1) / np.sqrt(np.mean(np.abs(ref) ** 2, axis=-1))
    return err, lead.value


@pytest.mark.parametrize("n,t,na,nr,k", SHAPES)
def test_ring_exchange_has_no_race(ring, n, t, na, nr, k):
    for sign in (1, -1):
        err, _ = run(ring, n, t, na, nr, k, sign, 0, 3, n + t + na + k)
        assert err.max() < 1e-6, (sign, err)


@pytest.mark.parametrize("n,t,na", [(8192, 512, 1), (2048, 256, 3)])
def test_ring_check_sees_a_missing_barrier(ring, n, t, na):
    """A planted mistake: on shapes whose ring needs the barrier between calls
    (the last exchange and the next call's first use one slot), running without
    it must give wrong transforms, so the check above can see a missing barrier."""
    err, lead = run(ring, n, t, na, 2, 1, 1, 1, 2, 7)
    assert lead == 1
    assert err.max() > 1e-2, err


# (length, threads, transforms, form): one region (1024, 8192, 16384), one region
# shared by R transforms (768, 1280, 2560), a region per transform (1536, and the
# seven lock-step transforms of the 7168 segment preparation)
FORMS = [(1024, 64, 1, 0), (8192, 512, 1, 0), (16384, 512, 1, 0), (256, 64, 3, 1), (256, 64, 5, 1),
         (512, 128, 5, 1), (512, 64, 3, 0), (512, 128, 7, 0)]


@pytest.mark.parametrize("n,t,na,form", FORMS)
def test_other_exchange_forms_have_no_race(ring, n, t, na, form):
    ring.form_race.argtypes = [C.c_int] * 7 + [C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(n + t + na)
    calls = 3
    for sign in (1, -1):
        x = (rng.standard_normal((calls, na, n)) + 1j * rng.standard_normal((calls, na, n))).astype(np.complex64)
        out = np.zeros_like(x)
        assert ring.form_race(n, t, sign, na, form, calls, SLEEP_US, x.ctypes.data, out.ctypes.data) == 0
        xd = x.astype(np.complex128)
        ref = np.fft.fft(xd, axis=-1) if sign < 0 else np.fft.ifft(xd, axis=-1) * n
        err = np.abs(out - ref).max() / np.sqrt(np.mean(np.abs(ref) ** 2))
        assert err < 1e-6, (sign, err)


@pytest.mark.parametrize("n,t", [(2048, 128), (4096, 256), (4096, 512), (8192, 512)])
def test_wavefront_split_has_no_race(ring, n, t):
    """WaveFFT (wavefront-local passes, one exchange across wavefronts) over
    successive calls that alternate between two LDS sets, as the correlation
    kernel's steps do."""
    ring.wave_race.argtypes = [C.c_int] * 6 + [C.c_void_p, C.c_void_p]
    calls = 4
    rng = np.random.default_rng(n + t + 11)
    for sign in (1, -1):
        x = (rng.standard_normal((calls, n)) + 1j * rng.standard_normal((calls, n))).astype(np.complex64)
        out = np.zeros_like(x)
        assert ring.wave_race(n, t, sign, 0, calls, SLEEP_US, x.ctypes.data, out.ctypes.data) == 0
        xd = x.astype(np.complex128)
        ref = np.fft.fft(xd, axis=-1) if sign < 0 else np.fft.ifft(xd, axis=-1) * n
        err = np.abs(out - ref).max() / np.sqrt(np.mean(np.abs(ref) ** 2))
        assert err < 1e-6, (sign, err)


def test_wavefront_split_check_sees_one_set_reused(ring):
    """A planted mistake: one LDS set for every call lets a wavefront start the
    next call's passes in its band while another still reads it."""
    ring.wave_race.argtypes = [C.c_int] * 6 + [C.c_void_p, C.c_void_p]
    n, t, calls = 4096, 256, 3
    rng = np.random.default_rng(5)
    x = (rng.standard_normal((calls, n)) + 1j * rng.standard_normal((calls, n))).astype(np.complex64)
    out = np.zeros_like(x)
    assert ring.wave_race(n, t, 1, 1, calls, SLEEP_US, x.ctypes.data, out.ctypes.data) == 0
    ref = np.fft.ifft(x.astype(np.complex128), axis=-1) * n
    err = np.abs(out - ref).max() / np.sqrt(np.mean(np.abs(ref) ** 2))
    assert err > 1e-2, err
