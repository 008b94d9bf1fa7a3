"""ctypes view of the CPU oracle (``liboracle.so``) and of ``_ref/libbpdref.so``.

TEST INFRASTRUCTURE ONLY: this is the checker.  It may be imported by ``tests/``,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``; the product
package ``gpu_accelsearch_amd`` never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libbpdref.so")

_lib = None
_ref = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


class OrParams(C.Structure):
    _fields_ = [("n", C.c_long), ("dt", C.c_float), ("zmax", C.c_int),
                ("numharm", C.c_int), ("fftlen", C.c_int), ("uselen", C.c_int)]


class OrHarmGeom(C.Structure):
    _fields_ = [("cur", C.c_int), ("max", C.c_int), ("rlo", C.c_int), ("rhi", C.c_double),
                ("zlo", C.c_int), ("numrs", C.c_int), ("numzs", C.c_int), ("fftlen", C.c_int),
                ("binoffset", C.c_int), ("lobin", C.c_int), ("numdata", C.c_int)]


def build() -> None:
    """Compile liboracle.so (and _ref when the reference sources are present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_nearest_int.restype = C.c_int
        L.or_nearest_int.argtypes = [C.c_double]
        L.or_calc_required_z.restype = C.c_int
        L.or_calc_required_z.argtypes = [C.c_double, C.c_double]
        L.or_calc_required_r.restype = C.c_double
        L.or_calc_required_r.argtypes = [C.c_double, C.c_double]
        L.or_index_from_r.restype = C.c_int
        L.or_index_from_r.argtypes = [C.c_double, C.c_double]
        L.or_index_from_z.restype = C.c_int
        L.or_index_from_z.argtypes = [C.c_double, C.c_double]
        L.or_z_response_halfwidth.restype = C.c_int
        L.or_z_response_halfwidth.argtypes = [C.c_double]
        L.or_next_pow2.restype = C.c_longlong
        L.or_next_pow2.argtypes = [C.c_longlong]
        L.or_fast_log2.restype = C.c_int
        L.or_fast_log2.argtypes = [C.c_int]
        L.or_total_harmonics.restype = C.c_int
        L.or_total_harmonics.argtypes = [C.c_int]
        L.or_fft_length.restype = C.c_int
        L.or_fft_length.argtypes = [C.c_int] * 4
        L.or_max_uselen.restype = C.c_int
        L.or_max_uselen.argtypes = [C.c_int, C.c_int]
        L.or_harmonic_order.restype = None
        L.or_harmonic_order.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_fresnel.restype = None
        L.or_fresnel.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_r_response.restype = None
        L.or_r_response.argtypes = [C.c_double, C.c_int, C.c_int, _dp]
        L.or_z_response.restype = None
        L.or_z_response.argtypes = [C.c_double, C.c_int, C.c_double, C.c_int, _dp]
        L.or_template_row.restype = None
        L.or_template_row.argtypes = [C.c_int, C.c_int, _dp]
        L.or_fft.restype = None
        L.or_fft.argtypes = [_dp, C.c_long, C.c_int]
        L.or_rfft.restype = None
        L.or_rfft.argtypes = [_dp, C.c_long, _dp]
        L.or_lower_median.restype = C.c_float
        L.or_lower_median.argtypes = [_fp, C.c_int]
        L.or_quake_rsqrt.restype = C.c_float
        L.or_quake_rsqrt.argtypes = [C.c_float]
        L.or_deredden.restype = None
        L.or_deredden.argtypes = [_fp, C.c_uint]
        L.or_step_geometry.restype = None
        L.or_step_geometry.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double,
                                       C.POINTER(OrHarmGeom)]
        L.or_num_steps.restype = C.c_int
        L.or_num_steps.argtypes = [C.c_long, C.c_int, C.c_int]
        L.or_segment_amplitudes.restype = C.c_float
        L.or_segment_amplitudes.argtypes = [_fp, C.c_long, C.c_int, C.c_int, C.c_int, _fp]
        L.or_row_powers_fft.restype = None
        L.or_row_powers_fft.argtypes = [_fp, _dp, C.c_int, C.c_int, C.c_int, _dp]
        L.or_row_powers_direct.restype = None
        L.or_row_powers_direct.argtypes = [_fp, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [_fp, C.POINTER(OrParams)]
        L.or_destroy.restype = None
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_ctx_info.restype = C.c_int
        L.or_ctx_info.argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 4 + [C.POINTER(C.c_long)]
        L.or_ctx_spectrum.restype = C.POINTER(C.c_float)
        L.or_ctx_spectrum.argtypes = [C.c_void_p]
        L.or_step_plane.restype = None
        L.or_step_plane.argtypes = [C.c_void_p, C.c_int, _dp, C.c_void_p]
        L.c32_fft.restype = None
        L.c32_fft.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.c32_create.restype = C.c_void_p
        L.c32_create.argtypes = [_fp, C.c_void_p]
        L.c32_destroy.restype = None
        L.c32_destroy.argtypes = [C.c_void_p]
        L.c32_info.restype = C.c_int
        L.c32_info.argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 3 + [C.POINTER(C.c_double)] * 3
        L.c32_step_plane.restype = None
        L.c32_step_plane.argtypes = [C.c_void_p, C.c_int, _fp]
        L.c32_run_steps.restype = C.c_double
        L.c32_run_steps.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                    C.POINTER(C.c_long), C.POINTER(C.c_double)]
        L.or_search_candidates.restype = C.c_long
        L.or_search_candidates.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_float),
                                           C.c_double, C.c_void_p, C.c_long]
        L.or_rz_response.restype = None
        L.or_rz_response.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
        L.or_rz_power.restype = C.c_double
        L.or_rz_power.argtypes = [_fp, C.c_long, C.c_double, C.c_double]
        _lib = L
    return _lib


def ref():
    """The reference's own bpd_math.c + bpd_fresnl.c, or None when not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        R.fast_median.restype = C.c_float
        R.fast_median.argtypes = [_fp, C.c_int]
        R.inverse_square_root.restype = C.c_float
        R.inverse_square_root.argtypes = [C.c_float]
        R.fresnl.restype = C.c_int
        R.fresnl.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        R.get2PowerNGreaterThan.restype = C.c_longlong
        R.get2PowerNGreaterThan.argtypes = [C.c_longlong]
        R.fast_log2.restype = C.c_int
        R.fast_log2.argtypes = [C.c_int]
        _ref = R
    return _ref


# ---------------------------------------------------------------- helpers
def fresnel(x: float):
    s, c = C.c_double(), C.c_double()
    lib().or_fresnel(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def harmonic_order(numharm: int):
    cur = (C.c_int * numharm)()
    mx = (C.c_int * numharm)()
    lib().or_harmonic_order(numharm, cur, mx)
    return list(zip(list(cur), list(mx)))


def z_response(z: float, numkern: int, roffset: float = 0.0, numbetween: int = 2):
    out = np.zeros(2 * numkern, np.float64)
    lib().or_z_response(roffset, numbetween, z, numkern, out)
    return out[0::2] + 1j * out[1::2]


def template_row(z: int, L: int):
    out = np.zeros(2 * L, np.float64)
    lib().or_template_row(int(z), int(L), out)
    return out[0::2] + 1j * out[1::2]


def fft(x: np.ndarray, sign: int = -1):
    a = np.ascontiguousarray(np.stack([x.real, x.imag], -1).reshape(-1), dtype=np.float64)
    lib().or_fft(a, len(x), sign)
    return a[0::2] + 1j * a[1::2]


def rfft(x: np.ndarray):
    xd = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(2 * (len(x) // 2 + 1), np.float64)
    lib().or_rfft(xd, len(x), out)
    return out[0::2] + 1j * out[1::2]


def deredden(spec_c64: np.ndarray, nbins: int | None = None):
    """Whiten a complex64 spectrum in place semantics; returns a new array."""
    a = np.ascontiguousarray(spec_c64.astype(np.complex64)).view(np.float32).copy()
    n = len(spec_c64) if nbins is None else nbins
    lib().or_deredden(a, n)
    return a.view(np.complex64)


def step_geometry(numharm: int, zmax: int, uselen: int, startr: float):
    g = (OrHarmGeom * numharm)()
    lib().or_step_geometry(numharm, zmax, uselen, float(startr), g)
    return [{f: getattr(x, f) for f, _ in OrHarmGeom._fields_} for x in g]


def segment_amplitudes(spec_c64: np.ndarray, lobin: int, numdata: int, L: int):
    s = np.ascontiguousarray(spec_c64.astype(np.complex64)).view(np.float32)
    out = np.zeros(L, np.float32)
    med = lib().or_segment_amplitudes(s, len(spec_c64), lobin, numdata, L, out)
    return out.view(np.complex64), med


OR_CAND_DTYPE = np.dtype([("step", np.int32), ("z_index", np.int32), ("r_index", np.int32),
                          ("stage", np.int32), ("power", np.float64), ("ambiguous", np.int32),
                          ("_pad", np.int32)])


class Search:
    """Whole-path oracle: spectrum, whitening, templates, per-step planes."""

    def __init__(self, series: np.ndarray, n: int, dt: float, zmax: int, numharm: int,
                 fftlen: int, uselen: int = 0):
        self.series = np.ascontiguousarray(series, dtype=np.float32)
        p = OrParams(n, dt, zmax, numharm, fftlen, uselen)
        self._p = p
        self.h = lib().or_create(self.series, C.byref(p))
        u, r, ns, z0 = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        nb = C.c_long()
        lib().or_ctx_info(self.h, C.byref(u), C.byref(r), C.byref(ns), C.byref(z0), C.byref(nb))
        self.uselen, self.rlo, self.nsteps, self.numz, self.nbins = (
            u.value, r.value, ns.value, z0.value, nb.value)
        self.numharm = numharm
        self.zmax = zmax

    def spectrum(self) -> np.ndarray:
        p = lib().or_ctx_spectrum(self.h)
        a = np.ctypeslib.as_array(p, shape=(2 * self.nbins,)).copy()
        return a.view(np.complex64)

    def step_plane(self, step: int, stages: bool = False):
        out = np.zeros(self.numz * self.uselen, np.float64)
        nst = int(np.log2(self.numharm)) + 1
        st = np.zeros(nst * self.numz * self.uselen, np.float64) if stages else None
        lib().or_step_plane(self.h, int(step), out,
                            st.ctypes.data_as(C.c_void_p) if stages else None)
        out = out.reshape(self.numz, self.uselen)
        if stages:
            return out, st.reshape(nst, self.numz, self.uselen)
        return out

    def candidates(self, powcut, margin: float, step0: int = 0, nsteps: int | None = None):
        """Cells above the stage cuts over steps [step0, step0 + nsteps) (OpenMP over
        steps): structured array step, z_index, r_index, stage, power, ambiguous
        (the stage sum lies within margin * max(P, 1) of its cut)."""
        ns = self.nsteps - step0 if nsteps is None else nsteps
        cut = (C.c_float * 5)(*[float(v) for v in list(powcut) + [0.0] * (5 - len(powcut))])
        n = lib().or_search_candidates(self.h, int(step0), int(ns), cut, float(margin), None, 0)
        if n < 0:
            raise MemoryError("or_search_candidates")
        out = np.zeros(n, OR_CAND_DTYPE)
        if n:
            m = lib().or_search_candidates(self.h, int(step0), int(ns), cut, float(margin),
                                           out.ctypes.data_as(C.c_void_p), n)
            assert m == n
        return out

    def close(self):
        if self.h:
            lib().or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Cpu32:
    """FP32 CPU form of the search (fdas_cpu32.c), the CPU baseline: R2C,
    whitening and template bank at construction (each timed), steps in parallel."""

    def __init__(self, series: np.ndarray, n: int, dt: float, zmax: int, numharm: int,
                 fftlen: int, uselen: int = 0):
        self.series = np.ascontiguousarray(series, dtype=np.float32)
        self._p = OrParams(n, dt, zmax, numharm, fftlen, uselen)
        self.h = lib().c32_create(self.series, C.byref(self._p))
        if not self.h:
            raise ValueError("c32_create: n/2 must be a power of two")
        u, ns, z0 = C.c_int(), C.c_int(), C.c_int()
        tr, tw, tt = C.c_double(), C.c_double(), C.c_double()
        lib().c32_info(self.h, C.byref(u), C.byref(ns), C.byref(z0), C.byref(tr), C.byref(tw),
                       C.byref(tt))
        self.uselen, self.nsteps, self.numz = u.value, ns.value, z0.value
        self.seconds = {"r2c": tr.value, "whiten": tw.value, "templates": tt.value}

    def step_plane(self, step: int) -> np.ndarray:
        out = np.zeros(self.numz * self.uselen, np.float32)
        lib().c32_step_plane(self.h, int(step), out)
        return out.reshape(self.numz, self.uselen)

    def run_steps(self, steps, powcut=(0.0,) * 5, threads: int = 0):
        """(wall seconds, cells above a stage cut, normalisation thread-seconds)."""
        st = np.ascontiguousarray(steps, dtype=np.int32)
        cut = np.array(list(powcut) + [0.0] * (5 - len(powcut)), np.float32)
        hits, tn = C.c_long(), C.c_double()
        wall = lib().c32_run_steps(self.h, st.ctypes.data, len(st), int(threads), cut.ctypes.data,
                                   C.byref(hits), C.byref(tn))
        return wall, hits.value, tn.value

    def close(self):
        if self.h:
            lib().c32_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fft32(x: np.ndarray, sign: int = -1) -> np.ndarray:
    """The FP32 CPU form's complex FFT (power-of-two length)."""
    a = np.ascontiguousarray(x, dtype=np.complex64).copy()
    lib().c32_fft(a.ctypes.data, len(a), int(sign))
    return a


def rz_response(q: float, z: float) -> complex:
    """int_0^1 exp(2 pi i (q u + z u^2/2)) du by Gauss-Legendre quadrature."""
    re, im = C.c_double(), C.c_double()
    lib().or_rz_response(float(q), float(z), C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def rz_power(spec: np.ndarray, r: float, z: float) -> float:
    """Matched power of a complex64 spectrum at (r, z) (no normalisation)."""
    s = np.ascontiguousarray(spec, dtype=np.complex64)
    return lib().or_rz_power(s.view(np.float32), len(s), float(r), float(z))
