"""MI355X-native Fourier-domain acceleration search (f-fdot) for DM-trial series.

The compute path is libfdas.so: hand-written HIP kernels for gfx950 behind the C
ABI in include/fdas.h.  Importing the package does not load the library; the
first call that needs it does, and raises if it has not been built.
"""
from ._native import FdasError, last_error, lib  # noqa: F401
from .accel import (ACCELCAND_DTYPE, AccelSearch, BatchSearch, autotune, candidate_sigma,
                    batch_from_files, load_dm_trials, numindep, power_for_sigma, readData,
                    read_inf, resolve_fftlen, harmonic_fftlens, corr_variant_shapes, sift_raw,
                    CANDIDATE_DTYPE, DeviceArray,
                    SearchParams,  # noqa: F401
                    Server, StagePipeline, SubharmonicBank, calc_required_r, calc_required_z,
                    calculateFFTLength, complexToComplexFFT, createSubharmonicInfo, deredden,
                    generate_harmonics_ffdot, getTotalHarmonics, index_from_r, index_from_z,
                    inverseComplexToComplexFFT, realToComplexFFT, stage_powcut, synchronize,
                    z_response_halfwidth)

__all__ = [
    "AccelSearch", "BatchSearch", "Server", "SearchParams", "DeviceArray", "StagePipeline", "SubharmonicBank",
    "createSubharmonicInfo", "realToComplexFFT", "complexToComplexFFT",
    "inverseComplexToComplexFFT", "deredden", "generate_harmonics_ffdot", "calc_required_z",
    "calc_required_r", "index_from_r", "index_from_z", "getTotalHarmonics",
    "z_response_halfwidth", "calculateFFTLength", "stage_powcut", "synchronize", "lib",
    "FdasError", "last_error", "CANDIDATE_DTYPE", "autotune", "resolve_fftlen", "harmonic_fftlens",
    "corr_variant_shapes",
    "ACCELCAND_DTYPE", "candidate_sigma", "power_for_sigma", "numindep", "sift_raw",
    "readData", "read_inf", "load_dm_trials", "batch_from_files",
]
