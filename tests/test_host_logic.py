"""Host-side logic of the product that needs no GPU: the C ABI loads and exports
every entry point declared in include/fdas.h, parameter validation rejects bad
plans, host data movement helpers, and the workgroup FFT's index arithmetic
(tests/fft_host_sim.cpp simulates a workgroup phase by phase)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "fdas.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?[\s\*](\w+)\s*\(", src, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "sizeof")))


def test_abi_exports_every_declared_symbol():
    import gpu_accelsearch_amd as fd
    L = fd.lib()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    from gpu_accelsearch_amd._native import PROTOTYPES
    bound = {p[0] for p in PROTOTYPES}
    assert set(names) <= bound, set(names) - bound


LLVM_BIN = "/opt/rocm/llvm/bin"


def test_power_stores_carry_no_immediate_offset(tmp_path):
    """The correlation kernels store powers through row-bounded buffer stores whose
    byte offset is col * size, a wrapped (huge) value for a negative column.  If the
    compiler split such an offset into a register part and an immediate, a negative
    register part plus the immediate would not wrap back into range in the
    hardware's check, and in-range powers would be dropped.  Every correlation
    kernel of the built library (gfx950 code object, disassembled) must store with
    a whole register offset."""
    lib = os.path.join(ROOT, "gpu_accelsearch_amd", "libfdas.so")
    tools = [os.path.join(LLVM_BIN, t) for t in ("llvm-objcopy", "clang-offload-bundler", "llvm-objdump")]
    if not os.path.exists(lib) or not all(os.path.exists(t) for t in tools):
        pytest.skip("needs the built library and the ROCm LLVM tools")
    fat = tmp_path / "fatbin"
    subprocess.run([tools[0], f"--dump-section=.hip_fatbin={fat}", lib, str(tmp_path / "copy.so")],
                   check=True)
    data = fat.read_bytes()
    starts = [m.start() for m in re.finditer(re.escape(b"__CLANG_OFFLOAD_BUNDLE__"), data)]
    stores = immediate = 0
    kernels = set()
    for k, p in enumerate(starts):
        part = tmp_path / f"bundle{k}"
        part.write_bytes(data[p:starts[k + 1] if k + 1 < len(starts) else len(data)])
        obj = tmp_path / f"gfx950_{k}.o"
        subprocess.run([tools[1], "--unbundle", "--type=o", "--targets=hipv4-amdgcn-amd-amdhsa--gfx950",
                        f"--input={part}", f"--output={obj}"], check=True)
        asm = subprocess.run([tools[2], "-d", "--mcpu=gfx950", str(obj)], check=True,
                             capture_output=True, text=True).stdout
        for fn in re.split(r"\n(?=[0-9a-f]{16} <)", asm):
            m = re.match(r"[0-9a-f]{16} <([^>]+)>:", fn)
            if not m or not re.search(r"corr_(wave|odd|sub|split)_kernel", m.group(1)):
                continue
            kernels.add(m.group(1))
            for line in fn.splitlines():
                if "buffer_store" in line:
                    stores += 1
                    immediate += bool(re.search(r"offset:\s*\d", line))
    assert len(kernels) > 100 and stores > 1000, (len(kernels), stores)
    assert immediate == 0, f"{immediate} of {stores} power stores carry an immediate offset"


@pytest.mark.parametrize("kw,code", [
    (dict(n=1001), 4), (dict(n=1000), 1), (dict(zmax=201), 1), (dict(zmax=1300), 1), (dict(numharm=3), 1),
    (dict(fftlen=512), 4), (dict(fftlen=6000), 4), (dict(uselen=7735), 1),
    (dict(uselen=8000), 1), (dict(zmax=400, precision="i16mma"), 4),
])
def test_plan_rejects_bad_parameters(kw, code):
    import gpu_accelsearch_amd as fd
    base = dict(n=1 << 20, zmax=200, numharm=8, fftlen=8192)
    base.update(kw)
    p = fd.SearchParams(**base).to_c()
    h = C.c_void_p()
    rc = fd.lib().fdas_plan_create(C.byref(p), C.byref(h))
    assert rc == code, (kw, rc, fd.last_error())
    assert h.value is None
    assert fd.last_error()


@pytest.mark.parametrize("lengths,code", [
    ([], 1), ([1 << 20, 1001], 4), ([1 << 20, 1000], 1), ([1 << 20, 100], 4),
    ([(1 << 27) + 2], 4),
])
def test_batch_rejects_bad_lengths(lengths, code):
    """Every trial is checked on the host before any device state is made."""
    import gpu_accelsearch_amd as fd
    p = fd.SearchParams(n=0, zmax=200, numharm=8, fftlen=8192).to_c()
    arr = (C.c_long * max(len(lengths), 1))(*lengths)
    h = C.c_void_p()
    rc = fd.lib().fdas_batch_create(C.byref(p), len(lengths), arr, C.byref(h))
    assert rc == code, (lengths, rc, fd.last_error())
    assert h.value is None
    assert fd.last_error()


def test_fftlen_defaults_and_corr_variants():
    """fftlen 0 resolves to the smallest valid length with template overhang
    4 hw(zmax) + 2 at most 1/8 of it; correlation shape tables are host state."""
    import gpu_accelsearch_amd as fd
    L = fd.lib()
    want = {(50, 4): 2048, (50, 16): 2048, (100, 8): 2048, (200, 8): 4096, (300, 8): 8192,
            (600, 16): 16384}
    for (zmax, numharm), fl in want.items():
        assert fd.resolve_fftlen(fd.SearchParams(n=1 << 20, zmax=zmax, numharm=numharm, fftlen=0)) == fl
    assert fd.resolve_fftlen(fd.SearchParams(n=1 << 20, zmax=200, numharm=8, fftlen=8192)) == 8192
    counts = {l: L.fdas_corr_variant_count(l) for l in (256, 512, 1024, 2048, 4096, 8192, 16384)}
    assert all(c >= 1 for c in counts.values()) and counts[8192] >= 2
    assert L.fdas_corr_variant_count(3000) == 0
    assert L.fdas_get_corr_variant(8192) == 0
    assert L.fdas_set_corr_variant(8192, counts[8192]) == 1
    assert L.fdas_set_corr_variant(8192, 1) == 0 and L.fdas_get_corr_variant(8192) == 1
    assert L.fdas_set_corr_variant(8192, 0) == 0


def test_uselen_setting():
    import gpu_accelsearch_amd as fd
    L = fd.lib()
    assert L.fdas_get_uselen() == 32000
    assert L.fdas_set_uselen(7735) == -1
    assert L.fdas_set_uselen(7734) == 0
    assert fd.calculateFFTLength(1, 1, 200) == 8192
    assert L.fdas_set_uselen(32000) == 0
    assert fd.calculateFFTLength(1, 1, 300) == 32768


def test_host_data_helpers():
    """get_fourier_amplitudes / spreadData (source/bpd_search.cu:16-67)."""
    import gpu_accelsearch_amd as fd
    L = fd.lib()
    spec = (np.arange(20) + 1j * np.arange(20)).astype(np.complex64)
    p = L.get_fourier_amplitudes(-3, 8, spec.ctypes.data, 20, 0)
    got = np.frombuffer((C.c_float * 16).from_address(p), np.float32).copy().view(np.complex64)
    C.CDLL(None).free(C.c_void_p(p))
    np.testing.assert_array_equal(got, np.concatenate([np.zeros(3, np.complex64), spec[:5]]))
    p = L.get_fourier_amplitudes(16, 8, spec.ctypes.data, 20, 0)
    got = np.frombuffer((C.c_float * 16).from_address(p), np.float32).copy().view(np.complex64)
    C.CDLL(None).free(C.c_void_p(p))
    np.testing.assert_array_equal(got, np.concatenate([spec[16:], np.zeros(4, np.complex64)]))
    out = np.full(20, 7 + 7j, np.complex64)
    L.spreadData(spec.ctypes.data, 5, 2, 3, out.ctypes.data)
    want = np.full(20, 7 + 7j, np.complex64)
    want[3:13] = 0
    want[3:13:2] = spec[:5]
    np.testing.assert_array_equal(out, want)


def test_generate_harmonics_ffdot_matches_oracle(oracle):
    """Per-step ffdotInfo from the exported host function vs the oracle (needs a
    SubharmonicInfo only for numzs/fftlen/halfwidth; built here by hand)."""
    import gpu_accelsearch_amd as fd
    from gpu_accelsearch_amd._native import FfdotInfo, SubharmonicInfo
    L = fd.lib()
    L.fdas_set_uselen(7734)
    try:
        H, stages, zmax = 8, 4, 200
        order = oracle.harmonic_order(H)
        shi = (SubharmonicInfo * H)()
        keep = []
        for h, (c, m) in enumerate(order):
            zh = fd.calc_required_z(c / m, zmax)
            shi[h].maxHarmonic, shi[h].currentHarmonic, shi[h].zmax = m, c, zh
            shi[h].totalKernelsTemplates = zh + 1
            shi[h].kernelFFTLength = fd.calculateFFTLength(m, c, zmax)
            hw = (C.c_ushort * (zh + 1))(*[fd.z_response_halfwidth(-zh + 2 * k) for k in range(zh + 1)])
            keep.append(hw)
            shi[h].kernelHalfWidth = C.cast(hw, C.POINTER(C.c_ushort))
        for startr in (536.0, 536.0 + 3867 * 500, 536.0 + 3867 * 1083):
            info = (FfdotInfo * H)()
            bo = (C.c_ushort * H)()
            L.generate_harmonics_ffdot(H, stages, shi, startr, startr + 3867 - 0.5, -zmax, info, bo)
            want = oracle.step_geometry(H, zmax, 7734, startr)
            for h in range(H):
                w = want[h]
                assert (info[h].currentHarmonic, info[h].maxHarmonic) == (w["cur"], w["max"])
                assert (info[h].numrs, info[h].numzs, info[h].rlo, info[h].zlo, info[h].fftlen) == \
                    (w["numrs"], w["numzs"], w["rlo"], w["zlo"], w["fftlen"])
                assert info[h].rhi == w["rhi"] and bo[h] == w["binoffset"]
    finally:
        L.fdas_set_uselen(32000)


@pytest.fixture(scope="module")
def fft_sim(tmp_path_factory):
    d = tmp_path_factory.mktemp("fftsim")
    so = str(d / "fft_sim.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", so,
                    os.path.join(ROOT, "tests", "fft_host_sim.cpp")], check=True)
    return C.CDLL(so)


@pytest.mark.parametrize("n,t", [(64, 16), (128, 32), (128, 64), (256, 64), (512, 64),
                                 (1024, 64), (1024, 128), (2048, 128), (2048, 256), (4096, 256),
                                 (4096, 512), (8192, 256), (8192, 512), (8192, 1024),
                                 (16384, 512), (16384, 1024)])
def test_workgroup_fft_index_arithmetic(fft_sim, n, t):
    """Every (length, threads) shape the library instantiates."""
    rng = np.random.default_rng(n)
    for sign in (-1, 1):
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        out = np.zeros(n, np.complex64)
        rc = fft_sim.fft_sim(n, t, sign, C.c_void_p(x.ctypes.data), C.c_void_p(out.ctypes.data))
        assert rc == 0
        ref = np.fft.fft(x.astype(np.complex128)) if sign < 0 else np.fft.ifft(x.astype(np.complex128)) * n
        assert np.max(np.abs(out - ref)) / np.sqrt(np.mean(np.abs(ref) ** 2)) < 1e-6


@pytest.mark.parametrize("n,t", [(2048, 128), (4096, 256), (4096, 512), (8192, 512), (16384, 1024)])
def test_wavefront_split_transform(fft_sim, n, t):
    """Length-n transform split across t/64 wavefronts (WaveFFT): inputs in the
    decimated order x[w + W (lane + 64 i)], outputs in natural order."""
    rng = np.random.default_rng(n + 3 * t)
    for sign in (-1, 1):
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        out = np.zeros(n, np.complex64)
        rc = fft_sim.fft_wave_sim(n, t, sign, C.c_void_p(x.ctypes.data), C.c_void_p(out.ctypes.data))
        assert rc == 0
        ref = np.fft.fft(x.astype(np.complex128)) if sign < 0 else np.fft.ifft(x.astype(np.complex128)) * n
        assert np.max(np.abs(out - ref)) / np.sqrt(np.mean(np.abs(ref) ** 2)) < 1e-6


@pytest.mark.parametrize("r,m,t", [(3, 128, 64), (3, 256, 64), (3, 512, 64), (3, 1024, 64),
                                   (3, 1024, 128), (3, 2048, 128), (3, 2048, 256), (5, 128, 64),
                                   (5, 256, 64), (5, 1024, 128), (5, 1024, 256), (7, 256, 64),
                                   (7, 512, 64), (7, 1024, 128), (7, 1024, 256)])
def test_length_rm_transform(fft_sim, r, m, t):
    """Length R*M (R = 3, 5, 7) = R lock-step length-M workgroup transforms plus the
    in-register radix-R step: forward from natural order into the decimated
    order, inverse back."""
    n = r * m
    rng = np.random.default_rng(n + t)
    for sign in (-1, 1):
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        out = np.zeros(n, np.complex64)
        rc = fft_sim.fft_odd_sim(r, m, t, sign, C.c_void_p(x.ctypes.data), C.c_void_p(out.ctypes.data))
        assert rc == 0
        ref = np.fft.fft(x.astype(np.complex128)) if sign < 0 else np.fft.ifft(x.astype(np.complex128)) * n
        assert np.max(np.abs(out - ref)) / np.sqrt(np.mean(np.abs(ref) ** 2)) < 1e-6


@pytest.mark.parametrize("zmax,numharm,fftlen,n", [
    (200, 8, 8192, 1 << 23), (200, 8, 4096, 1 << 23), (0, 8, 8192, 1 << 22), (50, 16, 2048, 1 << 20),
    (100, 16, 4096, 1 << 21), (600, 16, 16384, 1 << 22), (300, 4, 8192, 1 << 22), (200, 2, 1024, 1 << 20),
])
def test_sub_harmonic_lengths_fit_every_step(oracle, zmax, numharm, fftlen, n):
    """Each sub-harmonic's FFT length under the default (shortest) mode holds every
    step's normalised bins (numdata <= L/2) and the template span of the last kept
    column (numrs + 4 binoffset <= L, no wrap-around), by the oracle's own step
    geometry at every step of a trial.  The reference mode gives the reference's
    powers of two, and a shorter length is only ever 5/8, 3/4 or 7/8 of that."""
    import gpu_accelsearch_amd as fd
    P = fd.SearchParams(n=n, dt=6.4e-5, zmax=zmax, numharm=numharm, fftlen=fftlen)
    U, short = fd.harmonic_fftlens(P)
    U2, ref = fd.harmonic_fftlens(fd.SearchParams(n=n, dt=6.4e-5, zmax=zmax, numharm=numharm,
                                                  fftlen=fftlen, fft_lengths="reference"))
    assert U == U2
    order = oracle.harmonic_order(numharm)
    for h, (c, m) in enumerate(order):
        assert ref[h] & (ref[h] - 1) == 0
        assert short[h] in (ref[h], 5 * ref[h] // 8, 3 * ref[h] // 4, 7 * ref[h] // 8), (h, short[h], ref[h])
    assert short[0] == ref[0] == fftlen
    rlo = int(np.floor(np.float32(n) * np.float32(6.4e-5)))
    nsteps = 0
    startr = float(rlo)
    while startr + U // 2 < n // 2:
        nsteps += 1
        startr += U // 2
    for k in range(nsteps):
        g = oracle.step_geometry(numharm, zmax, U, float(rlo + k * (U // 2)))
        for h in range(1, numharm):
            assert g[h]["fftlen"] == ref[h]
            L = short[h]
            assert g[h]["numdata"] <= L // 2, (k, h, g[h]["numdata"], L)
            assert g[h]["numrs"] + 4 * g[h]["binoffset"] <= L, (k, h, g[h]["numrs"], g[h]["binoffset"], L)


def test_corr_variant_shapes_table():
    """Every correlation length's launch shapes (host tables, no GPU): at most
    FDAS_TUNE_MAXV per length, threads a power of two dividing the transform
    length with at least 2 values each, the split form only at powers of two,
    the wavefront-split form only at powers of two with its radix-W step
    dividing the values per thread (W = threads/64 >= 2), and an LDS ring of 2K
    regions only for K <= R transforms; out-of-range variants are rejected."""
    import ctypes as C
    import gpu_accelsearch_amd as fd
    from gpu_accelsearch_amd import _native as N
    lens = (256, 512, 1024, 2048, 4096, 8192, 16384, 768, 1536, 3072, 6144, 12288, 1280, 2560,
            5120, 10240, 7168)
    for L in lens:
        shapes = fd.corr_variant_shapes(L)
        assert 1 <= len(shapes) <= N.TUNE_MAXV, L
        r = next((q for q in (3, 5, 7) if L % q == 0), 1)
        m = L // r
        for sh in shapes:
            t, form = sh["threads"], sh["lds_form"]
            assert t & (t - 1) == 0 and m % t == 0 and m // t >= 2, (L, sh)
            if r == 1:
                assert form in (0, 1, 2, 8), (L, sh)
                assert sh["split"] == (form == 1)
                if form == 8:
                    w = t // 64
                    assert w >= 2 and (m // t) % w == 0, (L, sh)
            else:
                assert not sh["split"] and (form in (0, 1) or (form % 2 == 0 and form // 2 <= r)), (L, sh)
        z = C.c_int()
        assert fd.lib().fdas_corr_variant_shape(L, len(shapes), C.byref(z), C.byref(z), C.byref(z)) != 0
    assert fd.corr_variant_shapes(777) == []
