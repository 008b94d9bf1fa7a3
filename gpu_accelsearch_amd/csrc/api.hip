// api.hip -- the C ABI (include/fdas.h): reference-named entry points and the
// plan API of the fused search.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "fdas_internal.hpp"

namespace fdas {

// ------------------------------------------------------------------ errors
static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    const char* ab = getenv("FDAS_ABORT_ON_ERROR");
    if (ab && ab[0] == '1') {
        fprintf(stderr, "fdas: %s\n", g_err);
        abort();
    }
}

const char* last_error() { return g_err; }

int hip_fail(hipError_t e, const char* what, const char* file, int line)
{
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return e == hipErrorOutOfMemory ? FDAS_ERR_NOMEM : FDAS_ERR_HIP;
}

int check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return FDAS_ERR_HIP;
    }
    return FDAS_OK;
}

static int g_uselen = FDAS_DEFAULT_USELEN;

// Per-harmonic lengths recorded by initialize_cuda_plans (the reference keeps
// its cuFFT plans in globals, source/bpd_search.cu:10-13).
struct StagePlans {
    std::vector<int> L, Z;
};
static StagePlans g_plans;

}  // namespace fdas

using namespace fdas;

// ====================================================================== misc
extern "C" const char* fdas_last_error(void) { return last_error(); }
extern "C" void fdas_clear_error(void) { g_err[0] = 0; }
extern "C" const char* fdas_version(void) { return "fdas 0.1 (gfx950)"; }

extern "C" int fdas_set_uselen(int uselen)
{
    if (uselen <= 0 || (uselen & 1) || uselen >= 65536) {
        set_error("uselen %d must be even and in (0, 65536)", uselen);
        return -1;
    }
    g_uselen = uselen;
    return 0;
}
extern "C" int fdas_get_uselen(void) { return g_uselen; }

extern "C" int fdas_malloc(void** d_ptr, size_t bytes)
{
    FDAS_HIP(hipMalloc(d_ptr, bytes));
    return FDAS_OK;
}
extern "C" int fdas_free(void* d_ptr)
{
    FDAS_HIP(hipFree(d_ptr));
    return FDAS_OK;
}
extern "C" int fdas_memcpy_htod(void* d, const void* h, size_t bytes)
{
    FDAS_HIP(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
    return FDAS_OK;
}
extern "C" int fdas_memcpy_dtoh(void* h, const void* d, size_t bytes)
{
    FDAS_HIP(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
    return FDAS_OK;
}
extern "C" int fdas_memcpy_dtod(void* d, const void* src, size_t bytes)
{
    FDAS_HIP(hipMemcpy(d, src, bytes, hipMemcpyDeviceToDevice));
    return FDAS_OK;
}
extern "C" int fdas_memset(void* d, int value, size_t bytes)
{
    FDAS_HIP(hipMemset(d, value, bytes));
    return FDAS_OK;
}
extern "C" int fdas_device_synchronize(void)
{
    FDAS_HIP(hipDeviceSynchronize());
    return FDAS_OK;
}

// ====================================================================== geometry
extern "C" int calc_required_z(double f, double zfull) { return h_calc_required_z(f, zfull); }
extern "C" double calc_required_r(double f, double rfull) { return h_calc_required_r(f, rfull); }
extern "C" int index_from_r(double r, double rlo) { return (int)((r - rlo) * RDR + DBLCORRECT); }
extern "C" int index_from_z(double z, double zlo) { return (int)((z - zlo) * RDZ + DBLCORRECT); }
extern "C" int getTotalHarmonics(int stages) { return 1 << (stages - 1); }
extern "C" int z_response_halfwidth(double z) { return h_z_response_halfwidth(z); }
extern "C" int calculateFFTLength(unsigned short numharm, unsigned short harmnum, int max_zfull)
{
    return h_fft_length(g_uselen, numharm, harmnum, max_zfull);
}

// ====================================================================== templates
extern "C" SubharmonicInfo* createSubharmonicInfo(int numHarmonicStages, int zhi)
{
    if (numHarmonicStages < 1 || numHarmonicStages > 5 || zhi < 0 || zhi > 1200) {
        set_error("createSubharmonicInfo: stages %d / zhi %d out of range", numHarmonicStages, zhi);
        return nullptr;
    }
    const int H = 1 << (numHarmonicStages - 1);
    int cur[MAX_HARM], mx[MAX_HARM];
    h_harmonic_order(H, cur, mx);
    SubharmonicInfo* shi = (SubharmonicInfo*)calloc(H, sizeof(SubharmonicInfo));
    for (int h = 0; h < H; ++h) {
        SubharmonicInfo* s = &shi[h];
        const double f = (double)cur[h] / (double)mx[h];
        s->maxHarmonic = (unsigned short)mx[h];
        s->currentHarmonic = (unsigned short)cur[h];
        s->zmax = h_calc_required_z(f, zhi);
        s->totalKernelsTemplates = (unsigned short)((s->zmax / DZ) * 2 + 1);
        s->kernelFFTLength = h_fft_length(g_uselen, mx[h], cur[h], zhi);
        const int Z = s->totalKernelsTemplates, L = s->kernelFFTLength;
        s->z = (int*)malloc(sizeof(int) * Z);
        s->kernelHalfWidth = (unsigned short*)malloc(sizeof(unsigned short) * Z);
        for (int k = 0; k < Z; ++k) {
            s->z[k] = -s->zmax + k * DZ;
            s->kernelHalfWidth[k] = (unsigned short)h_z_response_halfwidth((double)s->z[k]);
        }
        cplx* bank = nullptr;
        if (hipMalloc(&bank, (size_t)Z * L * sizeof(cplx)) != hipSuccess ||
            build_template_bank(bank, s->zmax, Z, L, 0) != FDAS_OK) {
            if (!last_error()[0]) set_error("createSubharmonicInfo: bank allocation failed");
            (void)hipFree(bank);
            s->kernels_data = nullptr;
            freeSubharmonicInfo(shi, h + 1);
            return nullptr;
        }
        s->kernels_data = reinterpret_cast<fdas_complex*>(bank);
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        set_error("createSubharmonicInfo: template generation failed");
        freeSubharmonicInfo(shi, H);
        return nullptr;
    }
    return shi;
}

extern "C" void freeSubharmonicInfo(SubharmonicInfo* shi, int totalHarmonics)
{
    if (!shi) return;
    for (int h = 0; h < totalHarmonics; ++h) {
        free(shi[h].z);
        free(shi[h].kernelHalfWidth);
        (void)hipFree(shi[h].kernels_data);
    }
    free(shi);
}

// ====================================================================== FFTs
extern "C" unsigned realToComplexFFT(float* idata, long n, fdas_complex** odata, int batch)
{
    if (!idata || !odata || batch < 1) {
        set_error("realToComplexFFT: null argument");
        return 0;
    }
    *odata = nullptr;
    R2CPlan rp;
    if (r2c_plan_create(n, &rp, 0) != FDAS_OK) return 0;
    const long nout = n / 2 + 1;
    float* d_in = nullptr;
    cplx *d_out = nullptr, *work = nullptr;
    unsigned ret = 0;
    fdas_complex* host = (fdas_complex*)malloc(sizeof(fdas_complex) * nout * batch);
    if (hipMalloc(&d_in, sizeof(float) * n) == hipSuccess &&
        hipMalloc(&d_out, sizeof(cplx) * nout) == hipSuccess &&
        hipMalloc(&work, sizeof(cplx) * rp.work_elems) == hipSuccess) {
        ret = (unsigned)nout;
        for (int b = 0; b < batch && ret; ++b) {
            if (hipMemcpy(d_in, idata + (size_t)b * n, sizeof(float) * n, hipMemcpyHostToDevice) !=
                    hipSuccess ||
                fft_r2c_plan(rp, d_in, d_out, work, 0) != FDAS_OK ||
                hipMemcpy(host + (size_t)b * nout, d_out, sizeof(cplx) * nout,
                          hipMemcpyDeviceToHost) != hipSuccess)
                ret = 0;
        }
    } else {
        set_error("realToComplexFFT: device allocation failed");
    }
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    (void)hipFree(work);
    r2c_plan_destroy(&rp);
    if (!ret) {
        free(host);
        host = nullptr;
    }
    *odata = host;
    return ret;
}

static unsigned c2c_batched(fdas_complex* d_in, unsigned n, fdas_complex** d_out, int batch, int dir)
{
    if (!d_in || !d_out || batch < 1 || n < 64 || (n & (n - 1)) || n > (1u << 28)) {
        set_error("complex FFT: length %u (power of two in [64, 2^28] required)", n);
        return 0;
    }
    cplx* out = nullptr;
    if (hipMalloc(&out, sizeof(cplx) * (size_t)n * batch) != hipSuccess) {
        set_error("complex FFT: device allocation failed");
        return 0;
    }
    const cplx* in = reinterpret_cast<const cplx*>(d_in);
    int rc;
    if (n <= 16384) {
        rc = fft_rows(in, out, (int)n, batch, n, n, dir, 1.0f, 0, 0);
    } else {
        cplx* work = nullptr;
        rc = hipMalloc(&work, 2 * sizeof(cplx) * (size_t)n) == hipSuccess ? FDAS_OK : FDAS_ERR_NOMEM;
        for (int b = 0; b < batch && rc == FDAS_OK; ++b)
            rc = fft_c2c(in + (size_t)b * n, out + (size_t)b * n, n, dir, work, 0);
        (void)hipDeviceSynchronize();
        (void)hipFree(work);
    }
    if (rc == FDAS_OK && hipDeviceSynchronize() != hipSuccess) rc = FDAS_ERR_HIP;
    if (rc != FDAS_OK) {
        (void)hipFree(out);
        return 0;
    }
    *d_out = reinterpret_cast<fdas_complex*>(out);
    return n;
}

extern "C" unsigned complexToComplexFFT(fdas_complex* d_in, unsigned n, fdas_complex** d_out,
                                        int batch)
{
    return c2c_batched(d_in, n, d_out, batch, -1);
}

extern "C" unsigned inverseComplexToComplexFFT(fdas_complex* d_in, unsigned n,
                                               fdas_complex** d_out, int batch)
{
    return c2c_batched(d_in, n, d_out, batch, +1);
}

// ====================================================================== whitening
extern "C" void deredden(fdas_complex* fft, unsigned numOfBins)
{
    if (!fft) {
        set_error("deredden: null spectrum");
        return;
    }
    WhitenSchedule ws;
    if (whiten_schedule_create(numOfBins, &ws) != FDAS_OK) return;
    cplx* d = nullptr;
    if (hipMalloc(&d, sizeof(cplx) * numOfBins) == hipSuccess &&
        hipMemcpy(d, fft, sizeof(cplx) * numOfBins, hipMemcpyHostToDevice) == hipSuccess &&
        whiten(d, ws, 0) == FDAS_OK) {
        if (hipMemcpy(fft, d, sizeof(cplx) * numOfBins, hipMemcpyDeviceToHost) != hipSuccess)
            set_error("deredden: copy back failed");
    } else if (!last_error()[0]) {
        set_error("deredden: device failure");
    }
    (void)hipFree(d);
    whiten_schedule_destroy(&ws);
}

// ====================================================================== search stages
// source/bpd_search.cu:16-47 (host copy with zero padding outside the data).
extern "C" fdas_complex* get_fourier_amplitudes(int loBin, int numbins, fdas_complex* fft,
                                                long dataNumbins, long dataLoBin)
{
    fdas_complex* out = (fdas_complex*)malloc(sizeof(fdas_complex) * (numbins > 0 ? numbins : 1));
    for (int i = 0; i < numbins; ++i) {
        const long b = (long)loBin + i - dataLoBin;
        if (b >= 0 && b < dataNumbins) out[i] = fft[b];
        else out[i].x = out[i].y = 0.0f;
    }
    return out;
}

// source/bpd_search.cu:56-67
extern "C" void spreadData(fdas_complex* in, int inputElements, int interval, unsigned int startIdx,
                           fdas_complex* out)
{
    memset(out + startIdx, 0, sizeof(fdas_complex) * (size_t)inputElements * interval);
    for (int i = 0; i < inputElements; ++i) out[startIdx + (size_t)i * interval] = in[i];
}

// source/bpd_search.cu:72-127 (sequential harmonic order).
extern "C" void generate_harmonics_ffdot(unsigned short totalHarmonics, unsigned short stages,
                                         SubharmonicInfo* shi, double fullrlo, double fullrhi,
                                         double zlo, ffdotInfo* out, unsigned short* binoffset)
{
    int k = 0;
    for (int st = 0; st < stages && k < totalHarmonics; ++st) {
        const int mx = (st < 1) ? 1 : (1 << st);
        for (int c = 1; c <= mx && k < totalHarmonics; c += 2, ++k) {
            const double f = (double)c / (double)mx;
            const double drlo = h_calc_required_r(f, fullrlo);
            const double drhi = h_calc_required_r(f, fullrhi);
            out[k].currentHarmonic = (unsigned short)c;
            out[k].maxHarmonic = (unsigned short)mx;
            out[k].rlo = (int)std::floor(drlo);
            out[k].rhi = drhi;
            out[k].zlo = h_calc_required_z(f, zlo);
            if (c == 1 && mx == 1) {
                out[k].numrs = g_uselen;
            } else {
                int nr = (int)((std::ceil(drhi) - std::floor(drlo)) * RDR + DBLCORRECT) + 1;
                if (nr % RDR) nr = (nr / RDR + 1) * RDR;
                out[k].numrs = nr;
            }
            out[k].numzs = shi[k].totalKernelsTemplates;
            out[k].fftlen = shi[k].kernelFFTLength;
            binoffset[k] = shi[k].kernelHalfWidth[0];
        }
    }
}

// source/bpd_search.cu:130-184: fetch (host copy), then median normalisation and
// the 2x spread on the GPU; the spread rows land in output_data (host).
extern "C" void generate_harmonic_ffdot_data(unsigned short totalHarmonics, long totalDataLength,
                                             ffdotInfo* ffdot, unsigned short* binoffsetArray,
                                             long* dataStartIdx, fdas_complex* fft,
                                             long dataNumbins, long dataLoBin,
                                             fdas_complex* output_data)
{
    (void)totalDataLength;
    for (int h = 0; h < totalHarmonics; ++h) {
        const int bo = binoffsetArray[h];
        const int lobin = ffdot[h].rlo - bo;
        const int hibin = (int)std::ceil(ffdot[h].rhi) + bo;
        const int numdata = hibin - lobin + 1;
        const int nice = (int)h_next_pow2(numdata);
        const int L = ffdot[h].fftlen, nhalf = L / 2;
        if (nice < nhalf) {
            set_error("generate_harmonic_ffdot_data: %d fetched bins < fftlen/2 = %d", nice, nhalf);
            return;
        }
        fdas_complex* amps = get_fourier_amplitudes(lobin, nice, fft, dataNumbins, dataLoBin);
        cplx *d_amps = nullptr, *d_out = nullptr;
        bool ok = hipMalloc(&d_amps, sizeof(cplx) * nice) == hipSuccess &&
                  hipMalloc(&d_out, sizeof(cplx) * L) == hipSuccess &&
                  hipMemcpy(d_amps, amps, sizeof(cplx) * nice, hipMemcpyHostToDevice) == hipSuccess &&
                  launch_normalize_spread(d_amps, numdata, nhalf, d_out, nullptr, 0) == FDAS_OK &&
                  hipMemcpy(output_data + dataStartIdx[h], d_out, sizeof(cplx) * L,
                            hipMemcpyDeviceToHost) == hipSuccess;
        free(amps);
        (void)hipFree(d_amps);
        (void)hipFree(d_out);
        if (!ok) {
            if (!last_error()[0]) set_error("generate_harmonic_ffdot_data: device failure");
            return;
        }
    }
}

extern "C" void initialize_cuda_plans(unsigned short totalHarmonics, SubharmonicInfo* shi)
{
    g_plans.L.assign(totalHarmonics, 0);
    g_plans.Z.assign(totalHarmonics, 0);
    for (int h = 0; h < totalHarmonics; ++h) {
        g_plans.L[h] = shi[h].kernelFFTLength;
        g_plans.Z[h] = shi[h].totalKernelsTemplates;
    }
}

extern "C" void free_cuda_plans(unsigned short totalHarmonics)
{
    (void)totalHarmonics;
    g_plans = StagePlans();
}

static int row_fft_any(const cplx* in, cplx* out, int L, int rows, int dir)
{
    if (L <= 16384) return fft_rows(in, out, L, rows, L, L, dir, 1.0f, 0, 0);
    cplx* work = nullptr;
    FDAS_HIP(hipMalloc(&work, 2 * sizeof(cplx) * (size_t)L));
    int rc = FDAS_OK;
    for (int r = 0; r < rows && rc == FDAS_OK; ++r)
        rc = fft_c2c(in + (size_t)r * L, out + (size_t)r * L, L, dir, work, 0);
    (void)hipDeviceSynchronize();
    (void)hipFree(work);
    return rc;
}

extern "C" void perform_fft_harmonics(unsigned short totalHarmonics, fdas_complex* device_data,
                                      long* dataStartIdx, fdas_complex* output_device_fftdata)
{
    if ((size_t)totalHarmonics > g_plans.L.size()) {
        set_error("perform_fft_harmonics: initialize_cuda_plans was not called");
        return;
    }
    for (int h = 0; h < totalHarmonics; ++h) {
        const cplx* in = reinterpret_cast<const cplx*>(device_data + dataStartIdx[h]);
        cplx* out = reinterpret_cast<cplx*>(output_device_fftdata + dataStartIdx[h]);
        if (row_fft_any(in, out, g_plans.L[h], 1, -1) != FDAS_OK) return;
    }
}

extern "C" void perform_convolution(unsigned short totalHarmonics, fdas_complex* device_fft_data,
                                    ffdotInfo* harmonicInfo, SubharmonicInfo* shi,
                                    long* dataStartIdx, fdas_complex* out, long* convIdx)
{
    for (int h = 0; h < totalHarmonics; ++h) {
        if (launch_convolve(reinterpret_cast<const cplx*>(device_fft_data + dataStartIdx[h]),
                            reinterpret_cast<const cplx*>(shi[h].kernels_data),
                            harmonicInfo[h].numzs, harmonicInfo[h].fftlen,
                            reinterpret_cast<cplx*>(out + convIdx[h]), 0) != FDAS_OK)
            return;
    }
}

extern "C" void perform_inversefft(unsigned short totalHarmonics, fdas_complex* device_in,
                                   long* convIdx, fdas_complex* device_out)
{
    if ((size_t)totalHarmonics > g_plans.L.size()) {
        set_error("perform_inversefft: initialize_cuda_plans was not called");
        return;
    }
    for (int h = 0; h < totalHarmonics; ++h) {
        if (row_fft_any(reinterpret_cast<const cplx*>(device_in + convIdx[h]),
                        reinterpret_cast<cplx*>(device_out + convIdx[h]), g_plans.L[h],
                        g_plans.Z[h], +1) != FDAS_OK)
            return;
    }
}

extern "C" void generate_fundamental_powers(unsigned short totalHarmonics, fdas_complex* d_ifft,
                                            long* idx, ffdotInfo* hi, SubharmonicInfo* shi,
                                            float* d_powers)
{
    (void)totalHarmonics;
    const int off = shi[0].kernelHalfWidth[0] * NUMBETWEEN;
    const float scale = (float)(1.0 / (double)(hi[0].fftlen * hi[0].fftlen));
    launch_powers(reinterpret_cast<const cplx*>(d_ifft + idx[0] + off), hi[0].numzs, hi[0].numrs,
                  hi[0].fftlen, scale, d_powers, 0);
}

extern "C" void perform_harmonic_summing(unsigned short totalHarmonics, fdas_complex* d_ifft,
                                         long* idx, ffdotInfo* hi, SubharmonicInfo* shi,
                                         float* d_powers, int fullrlo, unsigned short* host_rinds,
                                         unsigned short* device_rinds, int totalRindsBins)
{
    (void)host_rinds;
    (void)device_rinds;
    (void)totalRindsBins;
    if (hipDeviceSynchronize() != hipSuccess) {
        set_error("perform_harmonic_summing: device error before summing");
        return;
    }
    for (int h = 1; h < totalHarmonics; ++h) {
        const double f = (double)hi[h].currentHarmonic / (double)hi[h].maxHarmonic;
        const int off = shi[h].kernelHalfWidth[0] * NUMBETWEEN;
        const float scale = (float)(1.0 / (double)(hi[h].fftlen * hi[h].fftlen));
        if (launch_add_powers(hi[0].numzs, hi[0].numrs, hi[h].fftlen, hi[0].zlo, hi[h].zlo,
                              hi[h].rlo, f, reinterpret_cast<const cplx*>(d_ifft + idx[h] + off),
                              scale, d_powers, fullrlo, 0) != FDAS_OK)
            return;
    }
}

// ====================================================================== plan API
struct fdas_plan {
    fdas_params p{};
    int uselen = 0, rlo = 0, nsteps = 0, H = 0, Z0 = 0, chunk = 0;
    long nbins = 0;
    int cur[MAX_HARM]{}, mx[MAX_HARM]{}, L[MAX_HARM]{}, Z[MAX_HARM]{}, zmax_h[MAX_HARM]{},
        off[MAX_HARM]{}, W[MAX_HARM]{}, zlo[MAX_HARM]{};
    cplx* bank[MAX_HARM]{};      // FP32 banks (freed after packing in reduced precision)
    void* bankq[MAX_HARM]{};     // bank in the storage format used by the kernels
    int* bank_exp[MAX_HARM]{};   // per-row exponents (reduced precision)
    int prec = FDAS_PREC_FP32;
    size_t cbytes = sizeof(cplx);   // bytes per stored complex value
    long long seg_off[MAX_HARM]{};
    long long seg_stride = 0;
    long long sub_total = 0;     // elements per step over all sub-harmonics
    long long sub_cells[MAX_HARM]{};  // sum over steps of Z_h * numrs_h
    std::vector<int> classes;    // distinct FFT lengths
    std::vector<std::vector<int>> class_harm;
    std::vector<SegDesc*> d_desc;  // per class: [nsteps][harmonics in class]
    int* d_fullrlo = nullptr;
    int* d_rlo_h = nullptr;      // [nsteps][MAX_HARM]
    cplx* d_spec = nullptr;
    cplx* d_r2c_work = nullptr;
    R2CPlan r2c;
    WhitenSchedule ws;
    void* d_seg = nullptr;       // [chunk][seg_stride] stored complex values
    int* d_seg_exp = nullptr;    // [chunk][H] segment exponents (reduced precision)
    void* d_sub = nullptr;       // sub-harmonic planes, float or FP16
    size_t sub_bytes = sizeof(float);
    float* d_plane = nullptr;
    fdas_candidate* d_cands = nullptr;
    int* d_ncand = nullptr;
    int cand_cap = 0;
    HarmSumArgs hs{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    size_t bytes = 0;
    // harmonic sum inside the fundamental's correlation kernel (FDAS_HARMSUM=fused)
    // instead of the tiled harmonic-sum kernel
    bool fused_harmsum = false;
    // event timing
    bool timing = false;
    struct Rec {
        int cat;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    size_t nrec = 0;
};

// Brackets one launch with events when timing is on.
struct TimedLaunch {
    fdas_plan* pl;
    size_t idx = (size_t)-1;
    TimedLaunch(fdas_plan* p, int cat) : pl(p)
    {
        if (!pl->timing) return;
        if (pl->nrec == pl->recs.size()) {
            fdas_plan::Rec r{cat, nullptr, nullptr};
            if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return;
            pl->recs.push_back(r);
        }
        idx = pl->nrec++;
        pl->recs[idx].cat = cat;
        (void)hipEventRecord(pl->recs[idx].a, pl->stream);
    }
    ~TimedLaunch()
    {
        if (idx != (size_t)-1) (void)hipEventRecord(pl->recs[idx].b, pl->stream);
    }
};

static int plan_alloc(fdas_plan* pl, void** ptr, size_t bytes)
{
    FDAS_HIP(hipMalloc(ptr, bytes));
    pl->bytes += bytes;
    return FDAS_OK;
}

static int plan_build(fdas_plan* pl)
{
    const fdas_params& p = pl->p;
    if (p.n < 256 || (p.n & 1) || p.n > (1L << 27)) {
        set_error("series length %ld: even and in [256, 2^27] required", p.n);
        return FDAS_ERR_UNSUPPORTED;
    }
    if (p.zmax < 0 || p.zmax > 1200 || (p.zmax & 1)) {
        set_error("zmax %d must be even and in [0, 1200]", p.zmax);
        return FDAS_ERR_INVALID;
    }
    if (p.numharm < 1 || p.numharm > 16 || (p.numharm & (p.numharm - 1))) {
        set_error("numharm %d must be 1, 2, 4, 8 or 16", p.numharm);
        return FDAS_ERR_INVALID;
    }
    if (p.fftlen < 1024 || p.fftlen > 16384 || (p.fftlen & (p.fftlen - 1))) {
        set_error("fftlen %d must be a power of two in [1024, 16384]", p.fftlen);
        return FDAS_ERR_UNSUPPORTED;
    }
    if (p.precision < FDAS_PREC_FP32 || p.precision > FDAS_PREC_BF16) {
        set_error("precision mode %d unknown", p.precision);
        return FDAS_ERR_INVALID;
    }
    if (p.precision != FDAS_PREC_FP32 && pl->fused_harmsum) {
        set_error("the fused harmonic-sum variant runs in FP32 only");
        return FDAS_ERR_UNSUPPORTED;
    }
    pl->prec = p.precision;
    pl->cbytes = p.precision == FDAS_PREC_FP32 ? sizeof(cplx) : 4;
    // power planes stay FP32: FP16 planes halve the bytes per staging load in the
    // harmonic sum and measured slower
    pl->sub_bytes = sizeof(float);
    pl->uselen = p.uselen ? p.uselen : h_max_uselen(p.fftlen, p.zmax);
    if (pl->uselen <= 0 || (pl->uselen & 1) || pl->uselen >= 65536 ||
        h_fft_length(pl->uselen, 1, 1, p.zmax) != p.fftlen) {
        set_error("uselen %d does not give fundamental FFT length %d at zmax %d", pl->uselen,
                  p.fftlen, p.zmax);
        return FDAS_ERR_INVALID;
    }
    const int U = pl->uselen;
    // AccelSearch.cu:101: rlo = floor(1.0 * (N * dt)) with a float product
    const float nd = (float)p.n * p.dt;
    pl->rlo = (int)std::floor(1.0 * nd);
    pl->nbins = p.n / 2;
    const double half = (double)(U / 2);
    {
        int ns = 0;
        double st = pl->rlo;
        while (st + half < (double)pl->nbins) {
            ++ns;
            st += half;
        }
        pl->nsteps = ns;
    }
    if (pl->nsteps <= 0) {
        set_error("no search steps: rlo %d + uselen/2 >= %ld bins", pl->rlo, pl->nbins);
        return FDAS_ERR_INVALID;
    }
    pl->H = p.numharm;
    h_harmonic_order(pl->H, pl->cur, pl->mx);
    long long so = 0, sub = 0;
    for (int h = 0; h < pl->H; ++h) {
        const double f = (double)pl->cur[h] / (double)pl->mx[h];
        pl->zmax_h[h] = h_calc_required_z(f, p.zmax);
        pl->Z[h] = (pl->zmax_h[h] / DZ) * 2 + 1;
        pl->L[h] = h_fft_length(U, pl->mx[h], pl->cur[h], p.zmax);
        pl->off[h] = NUMBETWEEN * h_z_response_halfwidth((double)-pl->zmax_h[h]);
        pl->zlo[h] = h_calc_required_z(f, -p.zmax);
        pl->seg_off[h] = so;
        so += pl->L[h] / 2;
        if (pl->L[h] < 256 || pl->L[h] > 16384) {
            set_error("harmonic %d/%d needs FFT length %d (fused kernels: 256..16384)",
                      pl->cur[h], pl->mx[h], pl->L[h]);
            return FDAS_ERR_UNSUPPORTED;
        }
    }
    pl->seg_stride = so;
    pl->Z0 = pl->Z[0];

    // per-step tables (source/bpd_search.cu:72-127, :139-143)
    const int S = pl->nsteps, H = pl->H;
    std::vector<int> fullrlo(S), rlo_h((size_t)S * MAX_HARM, 0);
    std::vector<int> lob((size_t)S * H), nd_((size_t)S * H);
    for (int h = 0; h < H; ++h) pl->W[h] = (h == 0) ? U : 0;
    for (int s = 0; s < S; ++s) {
        const double startr = pl->rlo + (double)s * half;
        const double lastr = startr + half - DR;
        fullrlo[s] = (int)startr;
        for (int h = 0; h < H; ++h) {
            const double f = (double)pl->cur[h] / (double)pl->mx[h];
            const double drlo = h_calc_required_r(f, startr);
            const double drhi = h_calc_required_r(f, lastr);
            const int rl = (int)std::floor(drlo);
            const int bo = pl->off[h] / NUMBETWEEN;
            int nr = U;
            if (h > 0) {
                nr = (int)((std::ceil(drhi) - std::floor(drlo)) * RDR + DBLCORRECT) + 1;
                if (nr % RDR) nr = (nr / RDR + 1) * RDR;
                // indices the harmonic sum reads at the first and last column
                const double s0 = (int)(RDR * (fullrlo[s] + 0 * DR) * f + 0.5) * DR;
                const double s1 = (int)(RDR * (fullrlo[s] + (U - 1) * DR) * f + 0.5) * DR;
                const int r0 = (int)((s0 - rl) * RDR), r1 = (int)((s1 - rl) * RDR);
                if (r0 < 0 || r1 >= nr) {
                    set_error("harmonic-sum index %d..%d outside [0,%d) at step %d", r0, r1, nr, s);
                    return FDAS_ERR_INVALID;
                }
            }
            pl->W[h] = std::max(pl->W[h], nr);
            pl->sub_cells[h] += (long long)pl->Z[h] * nr;
            rlo_h[(size_t)s * MAX_HARM + h] = rl;
            const int lobin = rl - bo;
            const int hibin = (int)std::ceil(drhi) + bo;
            const int numdata = hibin - lobin + 1;
            if (h_next_pow2(numdata) != pl->L[h] / 2) {
                set_error("step %d harmonic %d/%d: %d bins to normalise do not round to fftlen/2 = %d",
                          s, pl->cur[h], pl->mx[h], numdata, pl->L[h] / 2);
                return FDAS_ERR_UNSUPPORTED;
            }
            lob[(size_t)s * H + h] = lobin;
            nd_[(size_t)s * H + h] = numdata;
        }
    }
    for (int h = 1; h < H; ++h) {
        // rows the harmonic sum reads (monotone in the fundamental row)
        const int c = pl->cur[h], lm = h_fast_log2(pl->mx[h]);
        for (int j : {0, pl->Z[0] - 1}) {
            const int zz = pl->zlo[0] + DZ * j;
            const int zind = ((zz * c) / (1 << lm) - pl->zlo[h]) >> 1;
            if (zind < 0 || zind >= pl->Z[h]) {
                set_error("harmonic-sum row %d outside [0,%d) for harmonic %d", zind, pl->Z[h], h);
                return FDAS_ERR_INVALID;
            }
        }
    }
    for (int h = 0; h < H; ++h) {
        if (pl->off[h] + pl->W[h] > pl->L[h]) {
            set_error("harmonic %d: %d columns from offset %d exceed FFT length %d", h, pl->W[h],
                      pl->off[h], pl->L[h]);
            return FDAS_ERR_INVALID;
        }
        if (h > 0) sub += (long long)pl->Z[h] * pl->W[h];
    }
    pl->sub_total = sub;

    // size classes of the segment transform
    for (int h = 0; h < H; ++h) {
        auto it = std::find(pl->classes.begin(), pl->classes.end(), pl->L[h]);
        if (it == pl->classes.end()) {
            pl->classes.push_back(pl->L[h]);
            pl->class_harm.push_back({h});
        } else {
            pl->class_harm[it - pl->classes.begin()].push_back(h);
        }
    }

    // chunking
    const size_t per_step = (size_t)pl->seg_stride * sizeof(cplx) + (size_t)sub * sizeof(float) +
                            (size_t)pl->Z0 * U * sizeof(float);
    int chunk = p.chunk_steps;
    if (chunk <= 0) {
        const size_t budget = (size_t)16 << 30;
        chunk = (int)std::max<size_t>(1, budget / per_step);
    }
    pl->chunk = std::min(chunk, S);

    // device buffers (everything above is host-side validation and geometry)
    FDAS_HIP(hipStreamCreate(&pl->stream));
    pl->own_stream = true;
    for (int h = 0; h < H; ++h) {
        FDAS_TRY(plan_alloc(pl, (void**)&pl->bank[h], (size_t)pl->Z[h] * pl->L[h] * sizeof(cplx)));
        FDAS_TRY(build_template_bank(pl->bank[h], pl->zmax_h[h], pl->Z[h], pl->L[h], pl->stream));
        if (pl->prec == FDAS_PREC_FP32) {
            pl->bankq[h] = pl->bank[h];
        } else {
            FDAS_TRY(plan_alloc(pl, &pl->bankq[h], (size_t)pl->Z[h] * pl->L[h] * pl->cbytes));
            FDAS_TRY(plan_alloc(pl, (void**)&pl->bank_exp[h], (size_t)pl->Z[h] * sizeof(int)));
            FDAS_TRY(launch_bank_pack(pl->bank[h], pl->Z[h], pl->L[h], pl->prec, pl->bankq[h],
                                      pl->bank_exp[h], pl->stream));
            FDAS_HIP(hipStreamSynchronize(pl->stream));
            FDAS_HIP(hipFree(pl->bank[h]));
            pl->bytes -= (size_t)pl->Z[h] * pl->L[h] * sizeof(cplx);
            pl->bank[h] = nullptr;
        }
    }
    for (size_t c = 0; c < pl->classes.size(); ++c) {
        const auto& hl = pl->class_harm[c];
        std::vector<SegDesc> d((size_t)S * hl.size());
        for (int s = 0; s < S; ++s)
            for (size_t q = 0; q < hl.size(); ++q) {
                const int h = hl[q];
                SegDesc& e = d[(size_t)s * hl.size() + q];
                e.lobin = lob[(size_t)s * H + h];
                e.numdata = nd_[(size_t)s * H + h];
                e.out = (long long)s * pl->seg_stride + pl->seg_off[h];
                e.eslot = s * H + h;
                e.pad = 0;
            }
        SegDesc* dd = nullptr;
        FDAS_TRY(plan_alloc(pl, (void**)&dd, d.size() * sizeof(SegDesc)));
        pl->d_desc.push_back(dd);
        FDAS_HIP(hipMemcpy(dd, d.data(), d.size() * sizeof(SegDesc), hipMemcpyHostToDevice));
    }
    FDAS_TRY(plan_alloc(pl, (void**)&pl->d_fullrlo, (size_t)S * sizeof(int)));
    FDAS_TRY(plan_alloc(pl, (void**)&pl->d_rlo_h, rlo_h.size() * sizeof(int)));
    FDAS_HIP(hipMemcpy(pl->d_fullrlo, fullrlo.data(), (size_t)S * sizeof(int), hipMemcpyHostToDevice));
    FDAS_HIP(hipMemcpy(pl->d_rlo_h, rlo_h.data(), rlo_h.size() * sizeof(int), hipMemcpyHostToDevice));
    FDAS_TRY(plan_alloc(pl, (void**)&pl->d_spec, (size_t)(p.n / 2 + 1) * sizeof(cplx)));
    FDAS_TRY(r2c_plan_create(p.n, &pl->r2c, pl->stream));
    FDAS_TRY(plan_alloc(pl, (void**)&pl->d_r2c_work, pl->r2c.work_elems * sizeof(cplx)));
    FDAS_TRY(whiten_schedule_create((unsigned)pl->nbins, &pl->ws));
    FDAS_TRY(plan_alloc(pl, &pl->d_seg, (size_t)pl->chunk * pl->seg_stride * pl->cbytes));
    FDAS_TRY(plan_alloc(pl, (void**)&pl->d_seg_exp, (size_t)pl->chunk * H * sizeof(int)));
    if (sub > 0)
        FDAS_TRY(plan_alloc(pl, &pl->d_sub, (size_t)pl->chunk * sub * pl->sub_bytes));
    FDAS_TRY(plan_alloc(pl, (void**)&pl->d_plane, (size_t)pl->chunk * pl->Z0 * U * sizeof(float)));
    pl->cand_cap = p.max_candidates > 0 ? p.max_candidates : (1 << 20);
    FDAS_TRY(plan_alloc(pl, (void**)&pl->d_cands, (size_t)pl->cand_cap * sizeof(fdas_candidate)));
    FDAS_TRY(plan_alloc(pl, (void**)&pl->d_ncand, sizeof(int)));
    FDAS_HIP(hipMemsetAsync(pl->d_ncand, 0, sizeof(int), pl->stream));

    // harmonic-sum arguments
    HarmSumArgs& hs = pl->hs;
    hs.H = H;
    hs.U = U;
    hs.zlo0 = pl->zlo[0];
    long long acc = 0;
    for (int h = 0; h < H; ++h) {
        hs.c[h] = pl->cur[h];
        hs.lm[h] = h_fast_log2(pl->mx[h]);
        hs.zlo[h] = pl->zlo[h];
        hs.Z[h] = pl->Z[h];
        hs.W[h] = pl->W[h];
        if (h > 0) {
            hs.sub_off[h] = acc;
            hs.sub_stride[h] = (long long)pl->Z[h] * pl->W[h];
            acc += (long long)pl->chunk * hs.sub_stride[h];
        }
    }
    for (int st = 0; st < 5; ++st) hs.powcut[st] = p.powcut[st] > 0.f ? p.powcut[st] : INFINITY;
    hs.cand_cap = pl->cand_cap;
    hs.sub_half = pl->sub_bytes == 2;
    FDAS_HIP(hipStreamSynchronize(pl->stream));
    return FDAS_OK;
}

extern "C" void fdas_plan_destroy(fdas_plan* pl)
{
    if (!pl) return;
    if (pl->stream) (void)hipStreamSynchronize(pl->stream);
    for (int h = 0; h < MAX_HARM; ++h) {
        if (pl->bankq[h] != pl->bank[h]) (void)hipFree(pl->bankq[h]);
        (void)hipFree(pl->bank[h]);
        (void)hipFree(pl->bank_exp[h]);
    }
    for (SegDesc* d : pl->d_desc) (void)hipFree(d);
    (void)hipFree(pl->d_fullrlo);
    (void)hipFree(pl->d_rlo_h);
    (void)hipFree(pl->d_spec);
    (void)hipFree(pl->d_r2c_work);
    r2c_plan_destroy(&pl->r2c);
    if (pl->ws.d_start) whiten_schedule_destroy(&pl->ws);
    (void)hipFree(pl->d_seg);
    (void)hipFree(pl->d_seg_exp);
    (void)hipFree(pl->d_sub);
    (void)hipFree(pl->d_plane);
    (void)hipFree(pl->d_cands);
    (void)hipFree(pl->d_ncand);
    for (auto& r : pl->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    if (pl->own_stream) (void)hipStreamDestroy(pl->stream);
    delete pl;
}

extern "C" int fdas_plan_create(const fdas_params* params, fdas_plan** out)
{
    if (!params || !out) {
        set_error("fdas_plan_create: null argument");
        return FDAS_ERR_INVALID;
    }
    *out = nullptr;
    fdas_plan* pl = new fdas_plan();
    pl->p = *params;
    const char* hsenv = getenv("FDAS_HARMSUM");
    pl->fused_harmsum = hsenv && strcmp(hsenv, "fused") == 0;
    const int rc = plan_build(pl);
    if (rc != FDAS_OK) {
        fdas_plan_destroy(pl);
        return rc;
    }
    *out = pl;
    return FDAS_OK;
}

extern "C" int fdas_plan_get_info(const fdas_plan* pl, fdas_plan_info* info)
{
    if (!pl || !info) {
        set_error("fdas_plan_get_info: null argument");
        return FDAS_ERR_INVALID;
    }
    memset(info, 0, sizeof(*info));
    info->uselen = pl->uselen;
    info->rlo = pl->rlo;
    info->nsteps = pl->nsteps;
    info->numharm = pl->H;
    info->numz = pl->Z0;
    for (int h = 0; h < pl->H; ++h) {
        info->fftlen[h] = pl->L[h];
        info->numzs[h] = pl->Z[h];
        info->harm_cur[h] = pl->cur[h];
        info->harm_max[h] = pl->mx[h];
    }
    info->chunk_steps = pl->chunk;
    info->nbins = pl->nbins;
    info->cells = (long long)pl->nsteps * pl->uselen * pl->Z0;
    for (int h = 0; h < pl->H; ++h) info->sub_cells[h] = pl->sub_cells[h];
    info->device_bytes = pl->bytes;
    return FDAS_OK;
}

extern "C" int fdas_plan_set_stream(fdas_plan* pl, void* stream)
{
    if (!pl) {
        set_error("fdas_plan_set_stream: null plan");
        return FDAS_ERR_INVALID;
    }
    if (pl->own_stream) {
        (void)hipStreamSynchronize(pl->stream);
        (void)hipStreamDestroy(pl->stream);
        pl->own_stream = false;
    }
    if (stream) {
        pl->stream = (hipStream_t)stream;
    } else {
        FDAS_HIP(hipStreamCreate(&pl->stream));
        pl->own_stream = true;
    }
    return FDAS_OK;
}

extern "C" int fdas_prepare(fdas_plan* pl, const float* d_series)
{
    if (!pl || !d_series) {
        set_error("fdas_prepare: null argument");
        return FDAS_ERR_INVALID;
    }
    {
        TimedLaunch tl(pl, FDAS_K_R2C);
        FDAS_TRY(fft_r2c_plan(pl->r2c, d_series, pl->d_spec, pl->d_r2c_work, pl->stream));
    }
    {
        TimedLaunch tl(pl, FDAS_K_WHITEN);
        FDAS_TRY(whiten(pl->d_spec, pl->ws, pl->stream));
    }
    FDAS_HIP(hipMemsetAsync(pl->d_ncand, 0, sizeof(int), pl->stream));
    return FDAS_OK;
}

extern "C" int fdas_run_steps(fdas_plan* pl, int step0, int nsteps, float* d_planes)
{
    if (!pl || step0 < 0 || nsteps < 0 || step0 + nsteps > pl->nsteps) {
        set_error("fdas_run_steps: steps [%d, %d) outside [0, %d)", step0, step0 + nsteps,
                  pl ? pl->nsteps : 0);
        return FDAS_ERR_INVALID;
    }
    const int U = pl->uselen;
    for (int c0 = step0; c0 < step0 + nsteps; c0 += pl->chunk) {
        const int cs = std::min(pl->chunk, step0 + nsteps - c0);
        for (size_t c = 0; c < pl->classes.size(); ++c) {
            const int nh = (int)pl->class_harm[c].size();
            SegPrepArgs a{};
            a.spec = pl->d_spec;
            a.nbins = pl->nbins;
            a.desc = pl->d_desc[c] + (size_t)c0 * nh;
            a.nseg = cs * nh;
            a.seg = pl->d_seg;
            // descriptors hold offsets as if every step had a buffer slot
            a.out_base = (long long)c0 * pl->seg_stride;
            a.seg_exp = pl->d_seg_exp;
            a.eslot_base = c0 * pl->H;
            a.prec = pl->prec;
            TimedLaunch tl(pl, FDAS_K_SEGPREP);
            FDAS_TRY(launch_segment_prep(pl->classes[c], a, pl->stream));
        }
        auto corr_args = [&](int h, void* out, int W, int out_half) {
            CorrArgs a{};
            a.bank = pl->bankq[h];
            a.bank_exp = pl->bank_exp[h];
            a.Z = pl->Z[h];
            a.seg = pl->d_seg;
            a.seg_stride = pl->seg_stride;
            a.seg_off = pl->seg_off[h];
            a.seg_exp = pl->d_seg_exp;
            a.H = pl->H;
            a.h = h;
            a.nsteps = cs;
            a.off = pl->off[h];
            a.W = W;
            a.out = out;
            a.out_half = out_half;
            a.prec = pl->prec;
            return a;
        };
        for (int h = 1; h < pl->H; ++h) {
            TimedLaunch tl(pl, FDAS_K_CORR_SUB);
            void* out = static_cast<char*>(pl->d_sub) + (size_t)pl->hs.sub_off[h] * pl->sub_bytes;
            FDAS_TRY(launch_corr(pl->L[h], corr_args(h, out, pl->W[h], pl->sub_bytes == 2),
                                 pl->stream));
        }
        float* plane = d_planes ? d_planes + (size_t)(c0 - step0) * pl->Z0 * U : pl->d_plane;
        if (!pl->fused_harmsum) {
            {
                TimedLaunch tl(pl, FDAS_K_CORR_FUND);
                FDAS_TRY(launch_corr(pl->L[0], corr_args(0, plane, U, 0), pl->stream));
            }
            TimedLaunch tl(pl, FDAS_K_HARMSUM);
            FDAS_TRY(launch_harmsum(plane, pl->d_sub, pl->hs, c0, cs, pl->Z0, pl->d_fullrlo,
                                    pl->d_rlo_h, pl->d_cands, pl->d_ncand, pl->stream));
            continue;
        }
        TimedLaunch tl(pl, FDAS_K_CORR_FUND);
        FDAS_TRY(launch_corr_fund(pl->L[0], pl->bank[0], pl->Z0, static_cast<const cplx*>(pl->d_seg),
                                  pl->seg_stride, c0, cs, pl->off[0], pl->d_fullrlo, pl->d_rlo_h,
                                  static_cast<const float*>(pl->d_sub), pl->hs, plane,
                                  pl->d_cands, pl->d_ncand, pl->stream));
    }
    return FDAS_OK;
}

extern "C" int fdas_search(fdas_plan* pl, const float* d_series)
{
    FDAS_TRY(fdas_prepare(pl, d_series));
    return fdas_run_steps(pl, 0, pl->nsteps, nullptr);
}

extern "C" int fdas_get_spectrum(fdas_plan* pl, fdas_complex* d_out)
{
    if (!pl || !d_out) {
        set_error("fdas_get_spectrum: null argument");
        return FDAS_ERR_INVALID;
    }
    FDAS_HIP(hipMemcpyAsync(d_out, pl->d_spec, (size_t)pl->nbins * sizeof(cplx),
                            hipMemcpyDeviceToDevice, pl->stream));
    return FDAS_OK;
}

extern "C" int fdas_get_candidates(fdas_plan* pl, fdas_candidate* out, int capacity, int* count,
                                   int* overflowed)
{
    if (!pl || !count) {
        set_error("fdas_get_candidates: null argument");
        return FDAS_ERR_INVALID;
    }
    FDAS_HIP(hipStreamSynchronize(pl->stream));
    int n = 0;
    FDAS_HIP(hipMemcpy(&n, pl->d_ncand, sizeof(int), hipMemcpyDeviceToHost));
    const int kept = std::min(n, pl->cand_cap);
    if (overflowed) *overflowed = n > pl->cand_cap;
    std::vector<fdas_candidate> all(kept);
    if (kept > 0)
        FDAS_HIP(hipMemcpy(all.data(), pl->d_cands, (size_t)kept * sizeof(fdas_candidate),
                           hipMemcpyDeviceToHost));
    std::sort(all.begin(), all.end(), [](const fdas_candidate& a, const fdas_candidate& b) {
        if (a.step != b.step) return a.step < b.step;
        if (a.z_index != b.z_index) return a.z_index < b.z_index;
        if (a.r_index != b.r_index) return a.r_index < b.r_index;
        return a.stage < b.stage;
    });
    *count = kept;
    if (out && capacity > 0) memcpy(out, all.data(), sizeof(fdas_candidate) * std::min(kept, capacity));
    return FDAS_OK;
}

extern "C" int fdas_synchronize(fdas_plan* pl)
{
    if (!pl) {
        set_error("fdas_synchronize: null plan");
        return FDAS_ERR_INVALID;
    }
    FDAS_HIP(hipStreamSynchronize(pl->stream));
    return FDAS_OK;
}

extern "C" int fdas_plan_enable_timing(fdas_plan* pl, int enable)
{
    if (!pl) {
        set_error("fdas_plan_enable_timing: null plan");
        return FDAS_ERR_INVALID;
    }
    FDAS_HIP(hipStreamSynchronize(pl->stream));
    pl->timing = enable != 0;
    pl->nrec = 0;
    return FDAS_OK;
}

extern "C" int fdas_plan_get_timing(fdas_plan* pl, fdas_timing* out)
{
    if (!pl || !out) {
        set_error("fdas_plan_get_timing: null argument");
        return FDAS_ERR_INVALID;
    }
    FDAS_HIP(hipStreamSynchronize(pl->stream));
    memset(out, 0, sizeof(*out));
    for (size_t i = 0; i < pl->nrec; ++i) {
        float ms = 0.f;
        FDAS_HIP(hipEventElapsedTime(&ms, pl->recs[i].a, pl->recs[i].b));
        out->launches[pl->recs[i].cat] += 1;
        out->ms[pl->recs[i].cat] += ms;
    }
    return FDAS_OK;
}
