"""ctypes binding of libfdas.so (include/fdas.h).

The library is the product: every compute call runs HIP kernels.  Loading fails
loudly when the library has not been built; there is no other code path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfdas.so")

MAX_HARM = 16
MAX_STAGES = 5
FDAS_OK = 0
ERRORS = {1: "invalid argument", 2: "HIP runtime error", 3: "out of device memory",
          4: "unsupported shape", 5: "candidate capacity exceeded"}
PREC = {"fp32": 0, "fp16": 1, "bf16": 2, "i16": 3, "i16mma": 4}
LENGTHS = {"shortest": 0, "reference": 1}


class FdasComplex(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class SubharmonicInfo(C.Structure):
    _fields_ = [("maxHarmonic", C.c_ushort), ("currentHarmonic", C.c_ushort),
                ("zmax", C.c_int), ("totalKernelsTemplates", C.c_ushort),
                ("kernelFFTLength", C.c_int), ("z", C.POINTER(C.c_int)),
                ("kernelHalfWidth", C.POINTER(C.c_ushort)),
                ("kernels_data", C.c_void_p)]


class FfdotInfo(C.Structure):
    _fields_ = [("currentHarmonic", C.c_ushort), ("maxHarmonic", C.c_ushort),
                ("numrs", C.c_int), ("numzs", C.c_int), ("rlo", C.c_int), ("zlo", C.c_int),
                ("rhi", C.c_double), ("fftlen", C.c_int)]


class FdasParams(C.Structure):
    _fields_ = [("n", C.c_long), ("dt", C.c_float), ("zmax", C.c_int), ("numharm", C.c_int),
                ("fftlen", C.c_int), ("uselen", C.c_int), ("precision", C.c_int),
                ("chunk_steps", C.c_int), ("max_candidates", C.c_int),
                ("powcut", C.c_float * MAX_STAGES), ("fft_lengths", C.c_int)]


class FdasPlanInfo(C.Structure):
    _fields_ = [("uselen", C.c_int), ("rlo", C.c_int), ("nsteps", C.c_int),
                ("numharm", C.c_int), ("numz", C.c_int),
                ("fftlen", C.c_int * MAX_HARM), ("numzs", C.c_int * MAX_HARM),
                ("harm_cur", C.c_int * MAX_HARM), ("harm_max", C.c_int * MAX_HARM),
                ("chunk_steps", C.c_int), ("nbins", C.c_long), ("cells", C.c_longlong),
                ("sub_cells", C.c_longlong * MAX_HARM), ("device_bytes", C.c_size_t),
                ("ntrials", C.c_int), ("template_bytes", C.c_size_t)]


class FdasTrialInfo(C.Structure):
    _fields_ = [("n", C.c_long), ("nbins", C.c_long), ("rlo", C.c_int), ("nsteps", C.c_int),
                ("first_step", C.c_int)]


class FdasTiming(C.Structure):
    _fields_ = [("launches", C.c_int * 8), ("ms", C.c_double * 8)]


KERNEL_CATEGORIES = ["r2c", "whiten", "seg_prep", "corr_sub", "corr_fund", "harmsum"]


class FdasCandidate(C.Structure):
    _fields_ = [("step", C.c_int), ("z_index", C.c_int), ("r_index", C.c_int),
                ("stage", C.c_int), ("power", C.c_float), ("trial", C.c_int)]


class FdasServerConfig(C.Structure):
    _fields_ = [("memory_bytes", C.c_size_t), ("max_active", C.c_int), ("chunk_steps", C.c_int),
                ("max_candidates", C.c_int)]


class FdasRequestResult(C.Structure):
    _fields_ = [("id", C.c_longlong), ("status", C.c_int), ("ncandidates", C.c_int),
                ("overflowed", C.c_int), ("nsteps", C.c_int), ("cells", C.c_longlong),
                ("device_bytes", C.c_size_t), ("queue_ms", C.c_double),
                ("service_ms", C.c_double), ("latency_ms", C.c_double),
                ("message", C.c_char * 256)]


class FdasServerStats(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong),
                ("failed", C.c_longlong), ("active", C.c_int), ("pending", C.c_int),
                ("max_active_seen", C.c_int), ("memory_bytes", C.c_size_t),
                ("memory_in_use", C.c_size_t), ("memory_peak", C.c_size_t),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
                ("busy_s", C.c_double), ("trials_per_s", C.c_double)]


class FdasAccelcand(C.Structure):
    _fields_ = [("trial", C.c_int), ("numharm", C.c_int), ("step", C.c_int), ("z_index", C.c_int),
                ("r_index", C.c_int), ("stage", C.c_int), ("power", C.c_float), ("sigma", C.c_float),
                ("r", C.c_double), ("z", C.c_double), ("opt_r", C.c_double), ("opt_z", C.c_double),
                ("opt_power", C.c_float), ("opt_sigma", C.c_float),
                ("harm_power", C.c_float * MAX_HARM), ("freq", C.c_double), ("fdot", C.c_double),
                ("period", C.c_double), ("pdot", C.c_double)]


class FdasSiftParams(C.Structure):
    _fields_ = [("sigma_min", C.c_double), ("r_tol", C.c_double), ("z_tol", C.c_double),
                ("max_ratio", C.c_int)]


class FdasSiftTrial(C.Structure):
    _fields_ = [("n", C.c_long), ("nbins", C.c_long), ("rlo", C.c_int), ("dt", C.c_float)]


class FdasInf(C.Structure):
    _fields_ = [("basename", C.c_char * 256), ("telescope", C.c_char * 64), ("n", C.c_long),
                ("dt", C.c_double), ("dm", C.c_double), ("epoch_mjd", C.c_double),
                ("barycentred", C.c_int)]


TUNE_MAX = 12
TUNE_MAXV = 16


class FdasTuneReport(C.Structure):
    _fields_ = [("fftlen", C.c_int), ("ncand", C.c_int),
                ("cand_fftlen", C.c_int * TUNE_MAX), ("cand_ms_per_trial", C.c_double * TUNE_MAX),
                ("ncorr", C.c_int), ("corr_len", C.c_int * TUNE_MAX),
                ("corr_variant", C.c_int * TUNE_MAX), ("corr_threads", C.c_int * TUNE_MAX),
                ("corr_steps", C.c_int * TUNE_MAX), ("corr_split", C.c_int * TUNE_MAX),
                ("corr_us", (C.c_double * TUNE_MAXV) * TUNE_MAX),
                ("harmsum_variant", C.c_int), ("harmsum_threads", C.c_int),
                ("harmsum_rows", C.c_int), ("harmsum_cols", C.c_int),
                ("harmsum_ms", C.c_double * TUNE_MAXV), ("corr_form", C.c_int * TUNE_MAX)]


vp = C.c_void_p
i32 = C.c_int
u16 = C.c_ushort
PLong = C.POINTER(C.c_long)

# (name, restype, argtypes) for every entry point declared in include/fdas.h
PROTOTYPES = [
    ("fdas_set_uselen", i32, [i32]),
    ("fdas_get_uselen", i32, []),
    ("createSubharmonicInfo", C.POINTER(SubharmonicInfo), [i32, i32]),
    ("freeSubharmonicInfo", None, [C.POINTER(SubharmonicInfo), i32]),
    ("calc_required_z", i32, [C.c_double, C.c_double]),
    ("calc_required_r", C.c_double, [C.c_double, C.c_double]),
    ("index_from_r", i32, [C.c_double, C.c_double]),
    ("index_from_z", i32, [C.c_double, C.c_double]),
    ("getTotalHarmonics", i32, [i32]),
    ("z_response_halfwidth", i32, [C.c_double]),
    ("calculateFFTLength", i32, [u16, u16, i32]),
    ("realToComplexFFT", C.c_uint, [vp, C.c_long, C.POINTER(vp), i32]),
    ("complexToComplexFFT", C.c_uint, [vp, C.c_uint, C.POINTER(vp), i32]),
    ("inverseComplexToComplexFFT", C.c_uint, [vp, C.c_uint, C.POINTER(vp), i32]),
    ("deredden", None, [vp, C.c_uint]),
    ("fdas_deredden_noalloc", i32, [vp, C.c_uint, vp]),
    ("get_fourier_amplitudes", vp, [i32, i32, vp, C.c_long, C.c_long]),
    ("spreadData", None, [vp, i32, i32, C.c_uint, vp]),
    ("generate_harmonics_ffdot", None, [u16, u16, C.POINTER(SubharmonicInfo), C.c_double,
                                        C.c_double, C.c_double, C.POINTER(FfdotInfo),
                                        C.POINTER(u16)]),
    ("generate_harmonic_ffdot_data", None, [u16, C.c_long, C.POINTER(FfdotInfo), C.POINTER(u16),
                                            PLong, vp, C.c_long, C.c_long, vp]),
    ("perform_fft_harmonics", None, [u16, vp, PLong, vp]),
    ("perform_convolution", None, [u16, vp, C.POINTER(FfdotInfo), C.POINTER(SubharmonicInfo),
                                   PLong, vp, PLong]),
    ("perform_inversefft", None, [u16, vp, PLong, vp]),
    ("generate_fundamental_powers", None, [u16, vp, PLong, C.POINTER(FfdotInfo),
                                           C.POINTER(SubharmonicInfo), vp]),
    ("perform_harmonic_summing", None, [u16, vp, PLong, C.POINTER(FfdotInfo),
                                        C.POINTER(SubharmonicInfo), vp, i32, vp, vp, i32]),
    ("initialize_cuda_plans", None, [u16, C.POINTER(SubharmonicInfo)]),
    ("free_cuda_plans", None, [u16]),
    ("fdas_plan_create", i32, [C.POINTER(FdasParams), C.POINTER(vp)]),
    ("fdas_plan_destroy", None, [vp]),
    ("fdas_plan_get_info", i32, [vp, C.POINTER(FdasPlanInfo)]),
    ("fdas_plan_set_stream", i32, [vp, vp]),
    ("fdas_prepare", i32, [vp, vp]),
    ("fdas_run_steps", i32, [vp, i32, i32, vp]),
    ("fdas_search", i32, [vp, vp]),
    ("fdas_get_spectrum", i32, [vp, vp]),
    ("fdas_get_candidates", i32, [vp, C.POINTER(FdasCandidate), i32, C.POINTER(i32),
                                  C.POINTER(i32)]),
    ("fdas_reset_candidates", i32, [vp]),
    ("fdas_synchronize", i32, [vp]),
    ("fdas_batch_create", i32, [C.POINTER(FdasParams), i32, PLong, C.POINTER(vp)]),
    ("fdas_batch_trial_info", i32, [vp, i32, C.POINTER(FdasTrialInfo)]),
    ("fdas_batch_series", C.POINTER(C.c_float), [vp, i32]),
    ("fdas_prepare_trial", i32, [vp, i32, vp]),
    ("fdas_batch_run", i32, [vp]),
    ("fdas_plan_enable_graph", i32, [vp, i32]),
    ("fdas_batch_spectrum", i32, [vp, i32, vp]),
    ("fdas_plan_enable_timing", i32, [vp, i32]),
    ("fdas_plan_get_timing", i32, [vp, C.POINTER(FdasTiming)]),
    ("readData", C.c_long, [C.c_char_p, C.POINTER(C.POINTER(C.c_float))]),
    ("fdas_free_data", None, [C.POINTER(C.c_float)]),
    ("fdas_read_inf", i32, [C.c_char_p, C.POINTER(FdasInf)]),
    ("fdas_candidate_sigma", C.c_double, [C.c_double, i32, C.c_double]),
    ("fdas_power_for_sigma", C.c_double, [C.c_double, i32, C.c_double]),
    ("fdas_numindep", C.c_double, [C.c_double, C.c_double, i32, i32]),
    ("fdas_stage_powcuts", i32, [vp, i32, C.c_double, C.POINTER(C.c_float)]),
    ("fdas_sift", i32, [vp, vp, i32, C.POINTER(FdasSiftParams), vp, i32, C.POINTER(i32)]),
    ("fdas_sift_raw", i32, [C.POINTER(FdasSiftTrial), i32, i32, i32, i32, vp, i32,
                            C.POINTER(FdasSiftParams), vp, i32, C.POINTER(i32)]),
    ("fdas_optimize", i32, [vp, vp, i32]),
    ("fdas_rz_power", i32, [vp, i32, vp, vp, i32, vp]),
    ("fdas_autotune", i32, [C.POINTER(FdasParams), C.POINTER(i32), i32, C.POINTER(FdasTuneReport)]),
    ("fdas_corr_variant_count", i32, [i32]),
    ("fdas_harmsum_variant_count", i32, []),
    ("fdas_get_harmsum_variant", i32, []),
    ("fdas_set_harmsum_variant", i32, [i32]),
    ("fdas_get_corr_variant", i32, [i32]),
    ("fdas_set_corr_variant", i32, [i32, i32]),
    ("fdas_mma_variant_count", i32, []),
    ("fdas_get_mma_variant", i32, []),
    ("fdas_set_mma_variant", i32, [i32]),
    ("fdas_mma_variant_units", i32, [i32]),
    ("fdas_mma_variant_steps", i32, [i32]),
    ("fdas_mma_variant_threads", i32, [i32]),
    ("fdas_corr_variant_shape", i32, [i32, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
    ("fdas_resolve_fftlen", i32, [C.POINTER(FdasParams)]),
    ("fdas_harmonic_fftlens", i32, [C.POINTER(FdasParams), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("fdas_server_create", i32, [C.POINTER(FdasServerConfig), C.POINTER(vp)]),
    ("fdas_server_destroy", None, [vp]),
    ("fdas_server_prewarm", i32, [vp, C.POINTER(FdasParams)]),
    ("fdas_server_submit", i32, [vp, C.POINTER(FdasParams), vp, C.POINTER(C.c_longlong)]),
    ("fdas_server_wait", i32, [vp, C.c_longlong, C.POINTER(FdasRequestResult),
                               C.POINTER(FdasCandidate), i32]),
    ("fdas_server_get_stats", i32, [vp, C.POINTER(FdasServerStats)]),
    ("fdas_server_reset_stats", i32, [vp]),
    ("fdas_malloc", i32, [C.POINTER(vp), C.c_size_t]),
    ("fdas_free", i32, [vp]),
    ("fdas_memcpy_htod", i32, [vp, vp, C.c_size_t]),
    ("fdas_memcpy_dtoh", i32, [vp, vp, C.c_size_t]),
    ("fdas_memcpy_dtod", i32, [vp, vp, C.c_size_t]),
    ("fdas_memset", i32, [vp, i32, C.c_size_t]),
    ("fdas_device_synchronize", i32, []),
    ("fdas_last_error", C.c_char_p, []),
    ("fdas_clear_error", None, []),
    ("fdas_version", C.c_char_p, []),
]

_lib = None


def lib():
    """Load libfdas.so; raises if it was not built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m gpu_accelsearch_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in PROTOTYPES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class FdasError(RuntimeError):
    pass


def last_error() -> str:
    m = lib().fdas_last_error()
    return m.decode() if m else ""


def check(rc: int, what: str = "") -> None:
    if rc != FDAS_OK:
        raise FdasError(f"{what}: {ERRORS.get(rc, rc)}: {last_error()}")
