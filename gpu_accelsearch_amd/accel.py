"""Python face of the MI355X f-fdot acceleration search.

Mirrors the reference's host interface (AccelSearch.cu and the headers under
headers/) on top of libfdas.so.  Two layers:

* ``AccelSearch`` -- the plan API: one call whitens a device-resident series and
  runs every overlap-save step of the f-fdot search in fused HIP kernels.
* reference-named functions (``createSubharmonicInfo``, ``realToComplexFFT``,
  ``deredden``, ``generate_harmonics_ffdot``, ``perform_*`` ...) with the
  reference's argument meaning, for code that drives the stages one by one.

Device memory is plain ``DeviceArray`` allocations made by the library.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from ._native import check, lib


# ----------------------------------------------------------------- device memory
class DeviceArray:
    """A device allocation with a numpy dtype and shape."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().fdas_malloc(C.byref(p), max(self.nbytes, 1)), "fdas_malloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            check(lib().fdas_memcpy_htod(d.ptr, a.ctypes.data, a.nbytes), "fdas_memcpy_htod")
        return d

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().fdas_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes), "fdas_memcpy_dtoh")
        return out

    def zero(self) -> None:
        check(lib().fdas_memset(self.ptr, 0, self.nbytes), "fdas_memset")

    def free(self) -> None:
        if self.ptr:
            lib().fdas_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    check(lib().fdas_device_synchronize(), "fdas_device_synchronize")


# ----------------------------------------------------------------- plan API
@dataclass
class SearchParams:
    """Parameters of one DM-trial search (AccelSearch.cu:16-40, 52-58, 90-101)."""
    n: int
    dt: float = 6.4e-5
    zmax: int = 200
    numharm: int = 8
    fftlen: int = 8192
    uselen: int = 0
    precision: str = "fp32"
    chunk_steps: int = 0
    max_candidates: int = 0
    powcut: tuple = field(default_factory=lambda: (0.0,) * N.MAX_STAGES)
    fft_lengths: str = "shortest"   # sub-harmonic FFT lengths: "shortest" or "reference"

    def to_c(self) -> N.FdasParams:
        p = N.FdasParams()
        p.n, p.dt, p.zmax, p.numharm = self.n, self.dt, self.zmax, self.numharm
        p.fftlen, p.uselen, p.precision = self.fftlen, self.uselen, N.PREC[self.precision]
        p.chunk_steps, p.max_candidates = self.chunk_steps, self.max_candidates
        pc = list(self.powcut) + [0.0] * N.MAX_STAGES
        for i in range(N.MAX_STAGES):
            p.powcut[i] = float(pc[i])
        p.fft_lengths = N.LENGTHS[self.fft_lengths]
        return p


CANDIDATE_DTYPE = np.dtype([("step", np.int32), ("z_index", np.int32), ("r_index", np.int32),
                            ("stage", np.int32), ("power", np.float32), ("trial", np.int32)])


ACCELCAND_DTYPE = np.dtype([
    ("trial", np.int32), ("numharm", np.int32), ("step", np.int32), ("z_index", np.int32),
    ("r_index", np.int32), ("stage", np.int32), ("power", np.float32), ("sigma", np.float32),
    ("r", np.float64), ("z", np.float64), ("opt_r", np.float64), ("opt_z", np.float64),
    ("opt_power", np.float32), ("opt_sigma", np.float32), ("harm_power", np.float32, (N.MAX_HARM,)),
    ("freq", np.float64), ("fdot", np.float64), ("period", np.float64), ("pdot", np.float64)],
    align=True)
assert ACCELCAND_DTYPE.itemsize == C.sizeof(N.FdasAccelcand)


def candidate_sigma(power: float, numharm: int, numindep: float) -> float:
    """Equivalent Gaussian sigma of a summed power of ``numharm`` harmonics over
    ``numindep`` independent trials."""
    return lib().fdas_candidate_sigma(float(power), int(numharm), float(numindep))


def power_for_sigma(sigma: float, numharm: int, numindep: float) -> float:
    return lib().fdas_power_for_sigma(float(sigma), int(numharm), float(numindep))


def numindep(rlo: float, rhi: float, numz: int, stage: int) -> float:
    return lib().fdas_numindep(float(rlo), float(rhi), int(numz), int(stage))


def _sift_params(sigma_min, r_tol, z_tol, max_ratio):
    return N.FdasSiftParams(float(sigma_min), float(r_tol), float(z_tol), int(max_ratio))


def _run_sift(call, cands: np.ndarray) -> np.ndarray:
    cands = np.ascontiguousarray(cands, dtype=CANDIDATE_DTYPE)
    n_out = C.c_int()
    check(call(cands.ctypes.data, len(cands), None, 0, C.byref(n_out)), "fdas_sift")
    out = np.zeros(n_out.value, ACCELCAND_DTYPE)
    if n_out.value:
        check(call(cands.ctypes.data, len(cands), out.ctypes.data, n_out.value, C.byref(n_out)),
              "fdas_sift")
    return out


def sift_raw(trials, zmax: int, uselen: int, numz: int, cands: np.ndarray, sigma_min: float = 0.0,
             r_tol: float = 0.0, z_tol: float = 0.0, max_ratio: int = 0) -> np.ndarray:
    """Sift search candidates on explicit trial geometry: ``trials`` is a list of
    (n, nbins, rlo, dt)."""
    tr = (N.FdasSiftTrial * len(trials))(*[N.FdasSiftTrial(int(a), int(b), int(c), float(d))
                                            for a, b, c, d in trials])
    prm = _sift_params(sigma_min, r_tol, z_tol, max_ratio)
    return _run_sift(lambda ci, ni, co, cap, no: lib().fdas_sift_raw(
        tr, len(trials), zmax, uselen, numz, ci, ni, C.byref(prm), co, cap, no), cands)


class _CandidateOps:
    """Significance, sifting and optimisation on a plan's candidates."""

    def stage_powcuts(self, sigma: float, trial: int = 0):
        pc = (C.c_float * N.MAX_STAGES)()
        check(lib().fdas_stage_powcuts(self.handle, trial, float(sigma), pc), "fdas_stage_powcuts")
        return tuple(float(v) for v in pc)

    def sift(self, cands: np.ndarray | None = None, sigma_min: float = 0.0, r_tol: float = 0.0,
             z_tol: float = 0.0, max_ratio: int = 0) -> np.ndarray:
        if cands is None:
            cands = self.candidates()
        prm = _sift_params(sigma_min, r_tol, z_tol, max_ratio)
        return _run_sift(lambda ci, ni, co, cap, no: lib().fdas_sift(
            self.handle, ci, ni, C.byref(prm), co, cap, no), cands)

    def optimize(self, acc: np.ndarray) -> np.ndarray:
        """Optimised copy of sifted candidates (needs the spectra: run after a search)."""
        acc = np.ascontiguousarray(acc, dtype=ACCELCAND_DTYPE).copy()
        check(lib().fdas_optimize(self.handle, acc.ctypes.data, len(acc)), "fdas_optimize")
        return acc

    def rz_power(self, r, z, trial: int = 0) -> np.ndarray:
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.ascontiguousarray(z, dtype=np.float64)
        out = np.zeros(len(r), np.float32)
        check(lib().fdas_rz_power(self.handle, trial, r.ctypes.data, z.ctypes.data, len(r),
                                  out.ctypes.data), "fdas_rz_power")
        return out


class AccelSearch(_CandidateOps):
    """Fused f-fdot search of one series shape (templates stay on the device)."""

    def __init__(self, params: SearchParams):
        self.params = params
        h = C.c_void_p()
        cp = params.to_c()
        check(lib().fdas_plan_create(C.byref(cp), C.byref(h)), "fdas_plan_create")
        self.handle = h.value
        info = N.FdasPlanInfo()
        check(lib().fdas_plan_get_info(self.handle, C.byref(info)), "fdas_plan_get_info")
        self.info = info
        self.uselen, self.rlo, self.nsteps = info.uselen, info.rlo, info.nsteps
        self.numz, self.nbins, self.cells = info.numz, info.nbins, info.cells
        self.harmonics = [(info.harm_cur[h], info.harm_max[h], info.fftlen[h], info.numzs[h])
                          for h in range(info.numharm)]
        self._series = None

    # -- stages
    def prepare(self, series) -> None:
        """R2C FFT and whitening of a host array or DeviceArray of n float32."""
        if isinstance(series, DeviceArray):
            d = series
        else:
            x = np.ascontiguousarray(series, dtype=np.float32)
            if x.shape != (self.params.n,):
                raise ValueError(f"series must have shape ({self.params.n},), got {x.shape}")
            d = DeviceArray.from_host(x)
        self._series = d
        check(lib().fdas_prepare(self.handle, d.ptr), "fdas_prepare")

    def run_steps(self, step0: int, nsteps: int, out: DeviceArray | None = None):
        """Steps [step0, step0+nsteps); returns the summed planes on the host
        (shape [nsteps, numz, uselen]) unless ``out`` (a DeviceArray) is given."""
        dev = out if out is not None else DeviceArray((nsteps, self.numz, self.uselen), np.float32)
        check(lib().fdas_run_steps(self.handle, step0, nsteps, dev.ptr), "fdas_run_steps")
        self.synchronize()
        return dev if out is not None else dev.to_host()

    def search(self, series=None) -> None:
        """One DM trial: prepare + every step (planes stay on the device)."""
        if series is not None:
            self.prepare(series)
        check(lib().fdas_run_steps(self.handle, 0, self.nsteps, None), "fdas_run_steps")

    def synchronize(self) -> None:
        check(lib().fdas_synchronize(self.handle), "fdas_synchronize")

    def enable_timing(self, on: bool = True) -> None:
        check(lib().fdas_plan_enable_timing(self.handle, int(on)), "fdas_plan_enable_timing")

    def timing(self) -> dict:
        """{category: (launches, total ms)} from HIP events on the plan's stream."""
        t = N.FdasTiming()
        check(lib().fdas_plan_get_timing(self.handle, C.byref(t)), "fdas_plan_get_timing")
        return {name: (t.launches[i], t.ms[i]) for i, name in enumerate(N.KERNEL_CATEGORIES)}

    def spectrum(self) -> np.ndarray:
        d = DeviceArray((self.nbins,), np.complex64)
        check(lib().fdas_get_spectrum(self.handle, d.ptr), "fdas_get_spectrum")
        self.synchronize()
        return d.to_host()

    def candidates(self) -> np.ndarray:
        n, ov = C.c_int(), C.c_int()
        check(lib().fdas_get_candidates(self.handle, None, 0, C.byref(n), C.byref(ov)),
              "fdas_get_candidates")
        out = np.zeros(n.value, CANDIDATE_DTYPE)
        if n.value:
            check(lib().fdas_get_candidates(
                self.handle, out.ctypes.data_as(C.POINTER(N.FdasCandidate)), n.value,
                C.byref(n), C.byref(ov)), "fdas_get_candidates")
        self.overflowed = bool(ov.value)
        return out

    def candidate_rz(self, c: np.ndarray):
        """Fourier frequency r and drift z of candidates."""
        r = self.rlo + c["step"].astype(np.float64) * (self.uselen // 2) + 0.5 * c["r_index"]
        z = -self.params.zmax + 2.0 * c["z_index"]
        return r, z

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().fdas_plan_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchSearch(_CandidateOps):
    """Many DM trials (any even lengths, one zmax/numharm/fftlen) in one plan.

    The reference runs AccelSearch.cu's loop once per series; here every
    (trial, step) pair is an item of one chunked launch sequence, the plan owns
    one device input buffer per trial, and with ``graph=True`` the whole batch is
    recorded once into a hipGraph and replayed.  Template banks are shared by all
    live plans of the same configuration."""

    def __init__(self, params: SearchParams, lengths, graph: bool = False):
        self.params = params
        self.lengths = [int(n) for n in lengths]
        arr = (C.c_long * len(self.lengths))(*self.lengths)
        h = C.c_void_p()
        cp = params.to_c()
        check(lib().fdas_batch_create(C.byref(cp), len(self.lengths), arr, C.byref(h)),
              "fdas_batch_create")
        self.handle = h.value
        info = N.FdasPlanInfo()
        check(lib().fdas_plan_get_info(self.handle, C.byref(info)), "fdas_plan_get_info")
        self.info = info
        self.uselen, self.numz, self.cells = info.uselen, info.numz, info.cells
        self.trials = []
        for t in range(len(self.lengths)):
            ti = N.FdasTrialInfo()
            check(lib().fdas_batch_trial_info(self.handle, t, C.byref(ti)), "fdas_batch_trial_info")
            self.trials.append(ti)
        if graph:
            check(lib().fdas_plan_enable_graph(self.handle, 1), "fdas_plan_enable_graph")

    @property
    def ntrials(self) -> int:
        return len(self.lengths)

    def series_ptr(self, t: int) -> int:
        p = lib().fdas_batch_series(self.handle, t)
        if not p:
            raise N.FdasError(N.last_error())
        return C.cast(p, C.c_void_p).value

    def upload(self, t: int, series: np.ndarray) -> None:
        x = np.ascontiguousarray(series, dtype=np.float32)
        if x.shape != (self.lengths[t],):
            raise ValueError(f"trial {t}: expected ({self.lengths[t]},) floats, got {x.shape}")
        check(lib().fdas_memcpy_htod(self.series_ptr(t), x.ctypes.data, x.nbytes), "htod")

    def copy_in(self, t: int, d: DeviceArray) -> None:
        if d.nbytes != 4 * self.lengths[t]:
            raise ValueError(f"trial {t}: expected {4 * self.lengths[t]} bytes, got {d.nbytes}")
        check(lib().fdas_memcpy_dtod(self.series_ptr(t), d.ptr, d.nbytes), "dtod")

    def run(self) -> None:
        check(lib().fdas_batch_run(self.handle), "fdas_batch_run")

    def run_steps(self, step0: int, nsteps: int, out: DeviceArray | None = None):
        """Plan steps [step0, step0+nsteps) (numbered over all trials) of the
        prepared spectra; returns the summed planes unless ``out`` is given."""
        dev = out if out is not None else DeviceArray((nsteps, self.numz, self.uselen), np.float32)
        check(lib().fdas_run_steps(self.handle, step0, nsteps, dev.ptr), "fdas_run_steps")
        self.synchronize()
        return dev if out is not None else dev.to_host()

    def prepare(self) -> None:
        for t in range(self.ntrials):
            check(lib().fdas_prepare_trial(self.handle, t, None), "fdas_prepare_trial")
        check(lib().fdas_reset_candidates(self.handle), "fdas_reset_candidates")

    def synchronize(self) -> None:
        check(lib().fdas_synchronize(self.handle), "fdas_synchronize")

    def enable_timing(self, on: bool = True) -> None:
        check(lib().fdas_plan_enable_timing(self.handle, int(on)), "fdas_plan_enable_timing")

    def timing(self) -> dict:
        t = N.FdasTiming()
        check(lib().fdas_plan_get_timing(self.handle, C.byref(t)), "fdas_plan_get_timing")
        return {name: (t.launches[i], t.ms[i]) for i, name in enumerate(N.KERNEL_CATEGORIES)}

    def spectrum(self, t: int) -> np.ndarray:
        d = DeviceArray((self.trials[t].nbins,), np.complex64)
        check(lib().fdas_batch_spectrum(self.handle, t, d.ptr), "fdas_batch_spectrum")
        self.synchronize()
        return d.to_host()

    candidates = AccelSearch.candidates

    def candidate_rz(self, c: np.ndarray):
        rlo = np.array([tr.rlo for tr in self.trials], np.float64)[c["trial"]]
        r = rlo + c["step"].astype(np.float64) * (self.uselen // 2) + 0.5 * c["r_index"]
        z = -self.params.zmax + 2.0 * c["z_index"]
        return r, z

    close = AccelSearch.close
    __enter__ = AccelSearch.__enter__
    __exit__ = AccelSearch.__exit__
    __del__ = AccelSearch.__del__


class Server:
    """Continuous-batching request loop over the plan API (fdas_server_*).

    ``submit`` copies a host series and returns at once; a dispatcher thread in
    the library admits requests while a slot and device memory are free and runs
    them concurrently on separate streams.  ``wait`` returns the request's
    result record and candidates."""

    def __init__(self, memory_bytes: int = 0, max_active: int = 4, chunk_steps: int = 0,
                 max_candidates: int = 0):
        cfg = N.FdasServerConfig(memory_bytes, max_active, chunk_steps, max_candidates)
        h = C.c_void_p()
        check(lib().fdas_server_create(C.byref(cfg), C.byref(h)), "fdas_server_create")
        self.handle = h.value

    def prewarm(self, params: SearchParams) -> None:
        cp = params.to_c()
        check(lib().fdas_server_prewarm(self.handle, C.byref(cp)), "fdas_server_prewarm")

    def submit(self, params: SearchParams, series: np.ndarray) -> int:
        x = np.ascontiguousarray(series, dtype=np.float32)
        if params.n != len(x):
            params = SearchParams(**{**params.__dict__, "n": len(x)})
        cp = params.to_c()
        rid = C.c_longlong()
        check(lib().fdas_server_submit(self.handle, C.byref(cp), x.ctypes.data, C.byref(rid)),
              "fdas_server_submit")
        return rid.value

    def wait(self, rid: int, capacity: int = 1 << 20, raise_on_error: bool = True):
        """(result dict, candidates) of request ``rid``; raises if it failed, unless
        raise_on_error is False (then info["status"] holds the error code)."""
        res = N.FdasRequestResult()
        out = np.zeros(capacity, CANDIDATE_DTYPE)
        check(lib().fdas_server_wait(self.handle, rid, C.byref(res),
                                     out.ctypes.data_as(C.POINTER(N.FdasCandidate)), capacity),
              "fdas_server_wait")
        info = {k: getattr(res, k) for k, _ in N.FdasRequestResult._fields_ if k != "message"}
        if res.status != 0:
            if not raise_on_error:
                info["message"] = res.message.decode()
                return info, out[:0].copy()
            raise N.FdasError(f"request {rid}: {N.ERRORS.get(res.status, res.status)}: "
                              f"{res.message.decode()}")
        return info, out[:min(res.ncandidates, capacity)].copy()

    def stats(self) -> dict:
        st = N.FdasServerStats()
        check(lib().fdas_server_get_stats(self.handle, C.byref(st)), "fdas_server_get_stats")
        return {k: getattr(st, k) for k, _ in N.FdasServerStats._fields_}

    def reset_stats(self) -> None:
        check(lib().fdas_server_reset_stats(self.handle), "fdas_server_reset_stats")

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().fdas_server_destroy(self.handle)
            self.handle = None

    __enter__ = AccelSearch.__enter__
    __exit__ = AccelSearch.__exit__
    __del__ = AccelSearch.__del__


def autotune(params: SearchParams, fftlens=()) -> dict:
    """Measure correlation launch shapes and fundamental FFT lengths for series of
    ``params.n`` samples; the winners apply to later plans (fftlen 0)."""
    arr = (C.c_int * max(len(fftlens), 1))(*fftlens)
    rep = N.FdasTuneReport()
    cp = params.to_c()
    check(lib().fdas_autotune(C.byref(cp), arr, len(fftlens), C.byref(rep)), "fdas_autotune")
    return {
        "fftlen": rep.fftlen,
        "ms_per_trial": {rep.cand_fftlen[i]: (rep.cand_ms_per_trial[i] if rep.cand_ms_per_trial[i] >= 0
                                              else None) for i in range(rep.ncand)},
        "harmsum": {"variant": rep.harmsum_variant, "threads": rep.harmsum_threads,
                    "rows": rep.harmsum_rows, "cols": rep.harmsum_cols,
                    "ms_per_trial": [m for m in rep.harmsum_ms[:lib().fdas_harmsum_variant_count()]]},
        "corr": {rep.corr_len[i]: {"variant": rep.corr_variant[i], "threads": rep.corr_threads[i],
                                   "steps_per_wg": rep.corr_steps[i], "split": bool(rep.corr_split[i]),
                                   "lds_form": rep.corr_form[i],
                                   "us_per_variant": [u for u in rep.corr_us[i]
                                                      [:lib().fdas_corr_variant_count(rep.corr_len[i])]]}
                 for i in range(rep.ncorr)},
    }


def corr_variant_shapes(fftlen: int) -> list:
    """The correlation launch shapes selectable at one FFT length, variant 0 first:
    threads, steps per workgroup, LDS exchange form, and whether it is the split
    form (two L/2 transforms; powers of two only)."""
    out = []
    for v in range(lib().fdas_corr_variant_count(fftlen)):
        t, g, f = C.c_int(), C.c_int(), C.c_int()
        check(lib().fdas_corr_variant_shape(fftlen, v, C.byref(t), C.byref(g), C.byref(f)),
              "fdas_corr_variant_shape")
        out.append({"threads": t.value, "steps_per_wg": g.value, "lds_form": f.value,
                    "split": f.value == 1 and fftlen & (fftlen - 1) == 0})
    return out


def resolve_fftlen(params: SearchParams) -> int:
    cp = params.to_c()
    return lib().fdas_resolve_fftlen(C.byref(cp))


def harmonic_fftlens(params: SearchParams) -> tuple:
    """(uselen, [FFT length per harmonic]) a plan with ``params`` would use (host only)."""
    cp = params.to_c()
    u = C.c_int()
    arr = (C.c_int * N.MAX_HARM)()
    check(lib().fdas_harmonic_fftlens(C.byref(cp), C.byref(u), arr), "fdas_harmonic_fftlens")
    return u.value, list(arr[:params.numharm])


# ----------------------------------------------------------------- input files
PADDING = 2000   # zero floats around readData's samples (headers/bpd_common.h:35)


def readData(path: str) -> np.ndarray:
    """Samples of a PRESTO .dat file (float32 copy), via the library's readData."""
    p = C.POINTER(C.c_float)()
    n = lib().readData(str(path).encode(), C.byref(p))
    if n < 0:
        raise N.FdasError(f"readData: {N.last_error()}")
    try:
        return np.ctypeslib.as_array(p, shape=(n,)).copy()
    finally:
        lib().fdas_free_data(p)


def read_inf(path: str) -> dict:
    """Sample count, sample time, DM, epoch and names of a PRESTO .inf file."""
    inf = N.FdasInf()
    check(lib().fdas_read_inf(str(path).encode(), C.byref(inf)), "fdas_read_inf")
    return {"basename": inf.basename.decode(), "telescope": inf.telescope.decode(), "n": inf.n,
            "dt": inf.dt, "dm": inf.dm, "epoch_mjd": inf.epoch_mjd,
            "barycentred": bool(inf.barycentred)}


def load_dm_trials(basenames):
    """(series list, .inf dicts) of DM trials given as paths without suffix
    (``<base>.dat`` + ``<base>.inf``).  The .dat length must match the .inf."""
    series, infos = [], []
    for b in basenames:
        inf = read_inf(f"{b}.inf")
        x = readData(f"{b}.dat")
        if len(x) != inf["n"]:
            raise ValueError(f"{b}: .dat holds {len(x)} samples, .inf says {inf['n']}")
        series.append(x)
        infos.append(inf)
    return series, infos


def batch_from_files(params: SearchParams, basenames, graph: bool = False) -> "BatchSearch":
    """A BatchSearch over PRESTO DM-trial files (odd lengths lose their last
    sample: the search needs even lengths).  All trials must share dt, which sets
    params.dt."""
    series, infos = load_dm_trials(basenames)
    dts = {i["dt"] for i in infos}
    if len(dts) != 1:
        raise ValueError(f"DM trials with different sample times: {sorted(dts)}")
    params = SearchParams(**{**params.__dict__, "dt": dts.pop(), "n": 0})
    lengths = [len(x) - (len(x) & 1) for x in series]
    b = BatchSearch(params, lengths, graph=graph)
    for t, (x, n) in enumerate(zip(series, lengths)):
        b.upload(t, x[:n])
    b.infos = infos
    return b


def stage_powcut(sigma: float, numharm: int, numindep: float = 1.0) -> float:
    """Summed power whose false-alarm probability over ``numindep`` trials equals
    the Gaussian tail beyond ``sigma``.  Each normalised power is Exp(1) under
    noise, so the sum of ``numharm`` of them is Gamma(numharm, 1)."""
    from scipy.special import gammainccinv, ndtr
    p = (1.0 - ndtr(sigma)) / max(numindep, 1.0)
    return float(gammainccinv(numharm, p))


# ----------------------------------------------------------------- reference-named API
def fdas_set_uselen(u: int) -> None:
    if lib().fdas_set_uselen(int(u)) != 0:
        raise ValueError(N.last_error())


def calc_required_z(f: float, zfull: float) -> int:
    return lib().calc_required_z(f, zfull)


def calc_required_r(f: float, rfull: float) -> float:
    return lib().calc_required_r(f, rfull)


def index_from_r(r: float, rlo: float) -> int:
    return lib().index_from_r(r, rlo)


def index_from_z(z: float, zlo: float) -> int:
    return lib().index_from_z(z, zlo)


def getTotalHarmonics(stages: int) -> int:
    return lib().getTotalHarmonics(stages)


def z_response_halfwidth(z: float) -> int:
    return lib().z_response_halfwidth(z)


def calculateFFTLength(numharm: int, harmnum: int, max_zfull: int) -> int:
    return lib().calculateFFTLength(numharm, harmnum, max_zfull)


class SubharmonicBank:
    """createSubharmonicInfo() result: per-harmonic device template banks."""

    def __init__(self, stages: int, zhi: int):
        self.stages = stages
        self.total = getTotalHarmonics(stages)
        p = lib().createSubharmonicInfo(stages, zhi)
        if not p:
            raise N.FdasError(f"createSubharmonicInfo: {N.last_error()}")
        self.ptr = p
        self.entries = [p[h] for h in range(self.total)]

    def bank(self, h: int) -> np.ndarray:
        e = self.ptr[h]
        out = np.empty((e.totalKernelsTemplates, e.kernelFFTLength), np.complex64)
        check(lib().fdas_memcpy_dtoh(out.ctypes.data, e.kernels_data, out.nbytes), "bank copy")
        return out

    def free(self) -> None:
        if self.ptr:
            lib().freeSubharmonicInfo(self.ptr, self.total)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def createSubharmonicInfo(numHarmonicStages: int, zhi: int) -> SubharmonicBank:
    return SubharmonicBank(numHarmonicStages, zhi)


def realToComplexFFT(x: np.ndarray) -> np.ndarray:
    """R2C transform of a host float32 series -> n/2+1 complex64 bins."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = C.c_void_p()
    n = lib().realToComplexFFT(x.ctypes.data, len(x), C.byref(out), 1)
    if n == 0:
        raise N.FdasError(f"realToComplexFFT: {N.last_error()}")
    buf = (C.c_float * (2 * n)).from_address(out.value)
    res = np.frombuffer(buf, np.float32).copy().view(np.complex64)
    C.CDLL(None).free(C.c_void_p(out.value))
    return res


def _c2c(d_in: DeviceArray, n: int, batch: int, inverse: bool) -> DeviceArray:
    out = C.c_void_p()
    fn = lib().inverseComplexToComplexFFT if inverse else lib().complexToComplexFFT
    if fn(d_in.ptr, n, C.byref(out), batch) == 0:
        raise N.FdasError(N.last_error())
    d = DeviceArray.__new__(DeviceArray)
    d.shape, d.dtype, d.ptr = (batch, n), np.dtype(np.complex64), out.value
    d.nbytes = batch * n * 8
    return d


def complexToComplexFFT(d_in: DeviceArray, n: int, batch: int = 1) -> DeviceArray:
    return _c2c(d_in, n, batch, False)


def inverseComplexToComplexFFT(d_in: DeviceArray, n: int, batch: int = 1) -> DeviceArray:
    return _c2c(d_in, n, batch, True)


def deredden(spec: np.ndarray, numOfBins: int | None = None) -> np.ndarray:
    """Whitening of bins [0, numOfBins) of a complex64 spectrum (returns a copy)."""
    a = np.ascontiguousarray(spec, dtype=np.complex64).copy()
    nb = len(a) if numOfBins is None else int(numOfBins)
    lib().fdas_clear_error()
    lib().deredden(a.ctypes.data, nb)
    if N.last_error():
        raise N.FdasError(N.last_error())
    return a


def harmonic_fraction_order(stages: int):
    out = []
    for s in range(stages):
        mx = 1 if s == 0 else 1 << s
        out += [(c, mx) for c in range(1, mx + 1, 2)]
    return out


def generate_harmonics_ffdot(bank: SubharmonicBank, fullrlo: float, fullrhi: float, zlo: float):
    """Per-harmonic ffdotInfo of one step and the bin offsets (source/bpd_search.cu:72)."""
    H = bank.total
    info = (N.FfdotInfo * H)()
    bo = (C.c_ushort * H)()
    lib().generate_harmonics_ffdot(H, bank.stages, bank.ptr, fullrlo, fullrhi, zlo, info, bo)
    return info, bo


def step_bounds(startr: float, uselen: int):
    """startr -> (startr, lastr) as in AccelSearch.cu:211-212."""
    nextr = startr + uselen * 0.5
    return startr, nextr - 0.5


class StagePipeline:
    """The reference's per-step stage sequence (AccelSearch.cu:217-241) on device
    buffers laid out as there: harmonics concatenated at dataStartIdx /
    convolutionHarmonicIdx, and the fundamental power plane [numzs][numrs]."""

    def __init__(self, bank: SubharmonicBank):
        self.bank = bank
        H = bank.total
        self.H = H
        L = [bank.entries[h].kernelFFTLength for h in range(H)]
        Z = [bank.entries[h].totalKernelsTemplates for h in range(H)]
        self.dataStartIdx = (C.c_long * H)(*np.concatenate([[0], np.cumsum(L)[:-1]]).astype(int))
        self.convIdx = (C.c_long * H)(*np.concatenate(
            [[0], np.cumsum(np.array(L) * np.array(Z))[:-1]]).astype(int))
        self.total_fft = int(sum(L))
        self.total_conv = int(np.sum(np.array(L) * np.array(Z)))
        self.d_data = DeviceArray((self.total_fft,), np.complex64)
        self.d_fft = DeviceArray((self.total_fft,), np.complex64)
        self.d_conv = DeviceArray((self.total_conv,), np.complex64)
        self.d_ifft = DeviceArray((self.total_conv,), np.complex64)
        self.d_pow = None
        lib().initialize_cuda_plans(H, bank.ptr)

    def step(self, spectrum: np.ndarray, startr: float, zlo: float, uselen: int,
             nbins: int, stage: str = "summed", to_host: bool = True):
        """One step from a host whitened spectrum; returns the power plane (on
        the host, or the pipeline's device plane when to_host is False)."""
        H = self.H
        s0, s1 = step_bounds(startr, uselen)
        info, bo = generate_harmonics_ffdot(self.bank, s0, s1, zlo)
        host = np.zeros(self.total_fft, np.complex64)
        spec = np.ascontiguousarray(spectrum, dtype=np.complex64)
        lib().fdas_clear_error()
        lib().generate_harmonic_ffdot_data(H, self.total_fft, info, bo, self.dataStartIdx,
                                           spec.ctypes.data, nbins, 0, host.ctypes.data)
        if N.last_error():
            raise N.FdasError(N.last_error())
        check(lib().fdas_memcpy_htod(self.d_data.ptr, host.ctypes.data, host.nbytes), "htod")
        lib().perform_fft_harmonics(H, self.d_data.ptr, self.dataStartIdx, self.d_fft.ptr)
        lib().perform_convolution(H, self.d_fft.ptr, info, self.bank.ptr, self.dataStartIdx,
                                  self.d_conv.ptr, self.convIdx)
        lib().perform_inversefft(H, self.d_conv.ptr, self.convIdx, self.d_ifft.ptr)
        numz, numrs = info[0].numzs, info[0].numrs
        if self.d_pow is None or self.d_pow.shape != (numz, numrs):
            self.d_pow = DeviceArray((numz, numrs), np.float32)
        d_pow = self.d_pow
        lib().generate_fundamental_powers(H, self.d_ifft.ptr, self.convIdx, info, self.bank.ptr,
                                          d_pow.ptr)
        if stage == "summed":
            lib().perform_harmonic_summing(H, self.d_ifft.ptr, self.convIdx, info, self.bank.ptr,
                                           d_pow.ptr, int(startr), None, None, 0)
        if N.last_error():
            raise N.FdasError(N.last_error())
        if not to_host:
            return d_pow
        synchronize()
        return d_pow.to_host()

    def close(self) -> None:
        lib().free_cuda_plans(self.H)


def dm_trials_per_second(seconds_per_trial: float) -> float:
    return 1.0 / seconds_per_trial if seconds_per_trial > 0 else math.inf
