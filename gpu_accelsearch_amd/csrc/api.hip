// api.hip -- the C ABI (include/fdas.h): error state, memory plumbing and the
// reference-named entry points.  The plan API lives in plan.hip.
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

namespace {
struct HipAlloc final : DevAlloc {
    int alloc(void** p, size_t bytes) override
    {
        const hipError_t e = hipMalloc(p, bytes > 0 ? bytes : 1);
        if (e != hipSuccess) {
            *p = nullptr;
            return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
        }
        return FDAS_OK;
    }
    void release(void* p) override
    {
        if (p) (void)hipFree(p);
    }
};
}  // namespace

DevAlloc* hip_alloc()
{
    static HipAlloc a;
    return &a;
}

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
    if (r2c_plan_create(n, &rp) != FDAS_OK) return 0;
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

// Device staging buffers of the stage entry points, per device and purpose: kept
// for the life of the process and grown only when a call needs more, so calls
// in a loop allocate nothing.
enum { kStageAmps = 0, kStageSpread = 1, kStageFftWork = 2, kStageKinds = 3 };

static int stage_buffer(int kind, size_t bytes, void** out)
{
    static std::mutex mu;
    static std::map<std::pair<int, int>, std::pair<void*, size_t>> bufs;
    int dev = 0;
    FDAS_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto& b = bufs[std::make_pair(dev, kind)];
    if (b.second < bytes) {
        if (b.first) {
            FDAS_HIP(hipDeviceSynchronize());   // earlier calls may still read it
            (void)hipFree(b.first);
            b = {nullptr, 0};
        }
        FDAS_HIP(hipMalloc(&b.first, bytes));
        b.second = bytes;
    }
    *out = b.first;
    return FDAS_OK;
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
        void *d_amps = nullptr, *d_out = nullptr;
        bool ok = stage_buffer(kStageAmps, sizeof(cplx) * nice, &d_amps) == FDAS_OK &&
                  stage_buffer(kStageSpread, sizeof(cplx) * L, &d_out) == FDAS_OK &&
                  hipMemcpy(d_amps, amps, sizeof(cplx) * nice, hipMemcpyHostToDevice) == hipSuccess &&
                  launch_normalize_spread(static_cast<const cplx*>(d_amps), numdata, nhalf,
                                          static_cast<cplx*>(d_out), nullptr, 0) == FDAS_OK &&
                  hipMemcpy(output_data + dataStartIdx[h], d_out, sizeof(cplx) * L,
                            hipMemcpyDeviceToHost) == hipSuccess;
        free(amps);
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
    void* work = nullptr;
    FDAS_TRY(stage_buffer(kStageFftWork, 2 * sizeof(cplx) * (size_t)L, &work));
    int rc = FDAS_OK;
    for (int r = 0; r < rows && rc == FDAS_OK; ++r)
        rc = fft_c2c(in + (size_t)r * L, out + (size_t)r * L, L, dir, static_cast<cplx*>(work), 0);
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
