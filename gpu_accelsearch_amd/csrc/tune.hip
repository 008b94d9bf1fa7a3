// tune.hip -- measured autotuning of the search's shapes (fdas_autotune).
//
// Three choices are measured on the device, on synthetic data of the caller's
// shape:
//  * the launch shape of the fused correlation per FFT length (threads per
//    workgroup, which fixes the radix of the LDS passes, and steps per workgroup,
//    the template-block tile): each variant is timed on a block of 128
//    templates x 256 steps, and one more than 3% faster than the shape in use
//    replaces it process-wide;
//  * the tile shape of the harmonic sum: whole trials per shape (same margin);
//  * the fundamental FFT length (overlap-save segment length): whole DM trials
//    of the caller's series length are timed for every candidate length, and the
//    fastest is remembered for plans of the same (zmax, numharm, precision) that
//    leave fftlen at 0.
// Plans with fftlen 0 and no tuned entry use the smallest valid length whose
// template overhang 4*hw(zmax)+2 is at most 1/8 of the length.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "fdas_internal.hpp"
#include "precision.hpp"

using namespace fdas;

namespace fdas {

namespace {

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, int> g_tuned_fftlen;   // (zmax, numharm, prec) -> fftlen

__device__ __forceinline__ unsigned mix32(unsigned x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Roughly N(0,1) samples (sum of four uniforms, centred and scaled).
__global__ __launch_bounds__(256) void fill_noise_kernel(float* __restrict__ x, long n, unsigned seed)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned b = mix32((unsigned)i * 2654435761u ^ seed);
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) s += (float)(mix32(b + (unsigned)k * 0x9e3779b9u) >> 8) * (1.f / 16777216.f);
    x[i] = (s - 2.f) * 1.7320508f;
}

// Complex values in the given storage format.
template <int P>
__global__ __launch_bounds__(256) void fill_store_kernel(typename CStore<P>::T* __restrict__ x, long n,
                                                         unsigned seed)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned a = mix32((unsigned)i * 2246822519u ^ seed), b = mix32(a + 0x632be5abu);
    const cplx v = mk((float)(a >> 8) * (2.f / 16777216.f) - 1.f, (float)(b >> 8) * (2.f / 16777216.f) - 1.f);
    x[i] = CStore<P>::put(v);
}

unsigned blocks(long n) { return (unsigned)((n + 255) / 256); }

// A measured shape replaces the one in use only when it is faster by more than
// this factor's margin (run-to-run spread of a timing is about 2%).
constexpr double kTuneMargin = 0.97;

// Fundamental and harmonic FFT lengths of a configuration fit the kernels.
bool config_valid(int fftlen, int zmax, int numharm)
{
    if (fftlen < 1024 || fftlen > 16384 || (fftlen & (fftlen - 1))) return false;
    const int U = h_max_uselen(fftlen, zmax);
    if (U <= 0 || h_fft_length(U, 1, 1, zmax) != fftlen) return false;
    int cur[MAX_HARM], mx[MAX_HARM];
    h_harmonic_order(numharm, cur, mx);
    for (int h = 0; h < numharm; ++h) {
        const int L = h_fft_length(U, mx[h], cur[h], zmax);
        if (L < 256 || L > 16384) return false;
    }
    return true;
}

int fill_store(void* x, long n, int prec, unsigned seed, hipStream_t s)
{
    switch (prec) {
        case PREC_F16:
            hipLaunchKernelGGL(fill_store_kernel<PREC_F16>, dim3(blocks(n)), dim3(256), 0, s,
                               static_cast<unsigned*>(x), n, seed);
            break;
        case PREC_BF16:
            hipLaunchKernelGGL(fill_store_kernel<PREC_BF16>, dim3(blocks(n)), dim3(256), 0, s,
                               static_cast<unsigned*>(x), n, seed);
            break;
        case PREC_I16:
            hipLaunchKernelGGL(fill_store_kernel<PREC_I16>, dim3(blocks(n)), dim3(256), 0, s,
                               static_cast<unsigned*>(x), n, seed);
            break;
        default:
            hipLaunchKernelGGL(fill_store_kernel<PREC_F32>, dim3(blocks(n)), dim3(256), 0, s,
                               static_cast<cplx*>(x), n, seed);
    }
    return check_launch("fill_store_kernel");
}

// Microseconds per launch of each correlation variant at length L.
int time_corr_variants(int L, int prec, hipStream_t s, std::vector<double>* us)
{
    // 128 templates x 256 steps: enough workgroups that 16-step shapes are not starved
    const int Z = 128, S = 256, W = L - L / 8, off = L / 16;
    const size_t cb = cstore_bytes(prec);
    void *bank = nullptr, *seg = nullptr;
    int *bexp = nullptr, *sexp = nullptr;
    float* out = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    int rc = FDAS_OK;
    auto hip = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == FDAS_OK) rc = hip_fail(e, what, __FILE__, __LINE__);
    };
    hip(hipMalloc(&bank, (size_t)Z * L * cb), "hipMalloc");
    hip(hipMalloc(&seg, (size_t)S * (L / 2) * cb), "hipMalloc");
    hip(hipMalloc(&bexp, Z * sizeof(int)), "hipMalloc");
    hip(hipMalloc(&sexp, S * sizeof(int)), "hipMalloc");
    hip(hipMalloc(&out, (size_t)S * Z * W * sizeof(float)), "hipMalloc");
    hip(hipEventCreate(&a), "hipEventCreate");
    hip(hipEventCreate(&b), "hipEventCreate");
    if (rc == FDAS_OK) rc = fill_store(bank, (long)Z * L, prec, 11u, s);
    if (rc == FDAS_OK) rc = fill_store(seg, (long)S * (L / 2), prec, 12u, s);
    if (rc == FDAS_OK) {
        hip(hipMemsetAsync(bexp, 0, Z * sizeof(int), s), "hipMemsetAsync");
        hip(hipMemsetAsync(sexp, 0, S * sizeof(int), s), "hipMemsetAsync");
    }
    CorrArgs ca{};
    ca.bank = bank;
    ca.bank_exp = bexp;
    ca.bank_split = bank;   // same size and format; contents do not matter for timing
    ca.bank_split_exp = bexp;
    ca.Z = Z;
    ca.seg = seg;
    ca.seg_stride = L / 2;
    ca.seg_off = 0;
    ca.seg_exp = sexp;
    ca.H = 1;
    ca.h = 0;
    ca.nsteps = S;
    ca.off = off;
    ca.W = W;
    ca.out = out;
    ca.prec = prec;
    const int nv = corr_variant_count(L);
    us->assign(nv, 0.0);
    constexpr int REPS = 5;
    for (int v = 0; v < nv && rc == FDAS_OK; ++v) {
        for (int w = 0; w < 2 && rc == FDAS_OK; ++w) rc = launch_corr(L, ca, s, v);
        if (rc != FDAS_OK) break;
        hip(hipEventRecord(a, s), "hipEventRecord");
        for (int r = 0; r < REPS && rc == FDAS_OK; ++r) rc = launch_corr(L, ca, s, v);
        hip(hipEventRecord(b, s), "hipEventRecord");
        hip(hipEventSynchronize(b), "hipEventSynchronize");
        float ms = 0.f;
        hip(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        (*us)[v] = 1e3 * ms / REPS;
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(bank);
    (void)hipFree(seg);
    (void)hipFree(bexp);
    (void)hipFree(sexp);
    (void)hipFree(out);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    return rc;
}

// Milliseconds per DM trial (R2C, whitening and every step) at one fftlen, with
// harmonic-sum tile shape hs_variant (< 0: the process-wide one).
int time_trial(const fdas_params& base, int fftlen, hipStream_t s, double* ms_out,
               int hs_variant = -1)
{
    fdas_params p = base;
    p.fftlen = fftlen;
    p.uselen = 0;
    const long n = base.n;
    fdas_plan* pl = nullptr;
    FDAS_TRY(plan_create_ex(&p, 1, &n, true, s, nullptr, &pl));
    int rc = hs_variant >= 0 ? plan_set_harmsum_variant(pl, hs_variant) : FDAS_OK;
    float* x = fdas_batch_series(pl, 0);
    if (rc == FDAS_OK) {
        hipLaunchKernelGGL(fill_noise_kernel, dim3(blocks(n)), dim3(256), 0, s, x, n, 7u);
        rc = check_launch("fill_noise_kernel");
    }
    hipEvent_t a = nullptr, b = nullptr;
    if (rc == FDAS_OK && (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)) {
        set_error("autotune: event creation failed");
        rc = FDAS_ERR_HIP;
    }
    constexpr int REPS = 2;
    if (rc == FDAS_OK) rc = fdas_batch_run(pl);   // warm-up (twiddle tables, caches)
    if (rc == FDAS_OK && hipEventRecord(a, s) != hipSuccess) rc = FDAS_ERR_HIP;
    for (int r = 0; r < REPS && rc == FDAS_OK; ++r) rc = fdas_batch_run(pl);
    if (rc == FDAS_OK && hipEventRecord(b, s) != hipSuccess) rc = FDAS_ERR_HIP;
    if (rc == FDAS_OK && hipEventSynchronize(b) != hipSuccess) rc = FDAS_ERR_HIP;
    float ms = 0.f;
    if (rc == FDAS_OK && hipEventElapsedTime(&ms, a, b) != hipSuccess) rc = FDAS_ERR_HIP;
    *ms_out = ms / REPS;
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    fdas_plan_destroy(pl);
    return rc;
}

}  // namespace

int resolve_fftlen(const fdas_params& p)
{
    if (p.fftlen != 0) return p.fftlen;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned_fftlen.find({p.zmax, p.numharm, p.precision});
        if (it != g_tuned_fftlen.end()) return it->second;
    }
    const int over = 4 * h_z_response_halfwidth((double)p.zmax) + 2;
    int last = 0;
    for (int L = 1024; L <= 16384; L *= 2) {
        if (!config_valid(L, p.zmax, p.numharm)) continue;
        last = L;
        if (8 * over <= L) return L;
    }
    return last;
}

}  // namespace fdas

// ====================================================================== C ABI
extern "C" int fdas_corr_variant_count(int fftlen) { return corr_variant_count(fftlen); }

extern "C" int fdas_get_corr_variant(int fftlen) { return corr_variant(fftlen); }

extern "C" int fdas_set_corr_variant(int fftlen, int variant) { return set_corr_variant(fftlen, variant); }

extern "C" int fdas_mma_variant_count(void) { return mma_variant_count(); }

extern "C" int fdas_get_mma_variant(void) { return mma_variant(); }

extern "C" int fdas_set_mma_variant(int variant) { return set_mma_variant(variant); }

extern "C" int fdas_mma_variant_units(int variant) { return mma_variant_units(variant); }

extern "C" int fdas_mma_variant_steps(int variant) { return mma_variant_steps(variant); }

extern "C" int fdas_mma_variant_threads(int variant) { return mma_variant_threads(variant); }

extern "C" int fdas_corr_variant_shape(int fftlen, int variant, int* threads, int* steps_per_wg,
                                       int* lds_form)
{
    if (!threads || !steps_per_wg || !lds_form) {
        set_error("fdas_corr_variant_shape: null argument");
        return FDAS_ERR_INVALID;
    }
    if (variant < 0 || variant >= corr_variant_count(fftlen)) {
        set_error("correlation variant %d out of range at length %d", variant, fftlen);
        return FDAS_ERR_INVALID;
    }
    corr_variant_shape(fftlen, variant, threads, steps_per_wg, nullptr, lds_form);
    return FDAS_OK;
}

extern "C" int fdas_resolve_fftlen(const fdas_params* params)
{
    if (!params) {
        set_error("fdas_resolve_fftlen: null argument");
        return 0;
    }
    return resolve_fftlen(*params);
}

extern "C" int fdas_autotune(const fdas_params* base, const int* fftlens, int nfftlen,
                             fdas_tune_report* rep)
{
    if (!base || !rep || nfftlen < 0 || nfftlen > FDAS_TUNE_MAX || (nfftlen > 0 && !fftlens)) {
        set_error("fdas_autotune: bad argument");
        return FDAS_ERR_INVALID;
    }
    memset(rep, 0, sizeof(*rep));
    std::vector<int> cand;
    if (nfftlen == 0) {
        for (int L = 1024; L <= 16384; L *= 2) cand.push_back(L);
    } else {
        cand.assign(fftlens, fftlens + nfftlen);
    }
    hipStream_t s = nullptr;
    FDAS_HIP(hipStreamCreate(&s));
    int rc = FDAS_OK;
    // 1. correlation shapes of every FFT length the valid candidates use (the
    //    direct correlation of FDAS_PREC_I16_MMA has no FFT shapes)
    std::vector<int> Ls;
    for (int L0 : cand) {
        if (base->precision == FDAS_PREC_I16_MMA) break;
        if (!config_valid(L0, base->zmax, base->numharm)) continue;
        const int U = h_max_uselen(L0, base->zmax);
        int cur[MAX_HARM], mx[MAX_HARM];
        h_harmonic_order(base->numharm, cur, mx);
        for (int h = 0; h < base->numharm; ++h)
            Ls.push_back(h_harmonic_length(U, mx[h], cur[h], base->zmax, base->fft_lengths));
    }
    std::sort(Ls.begin(), Ls.end());
    Ls.erase(std::unique(Ls.begin(), Ls.end()), Ls.end());
    for (int L : Ls) {
        if (rep->ncorr == FDAS_TUNE_MAX) break;
        const int k = rep->ncorr++;
        rep->corr_len[k] = L;
        std::vector<double> us;
        if (corr_variant_count(L) > 1) {
            rc = time_corr_variants(L, base->precision, s, &us);
            if (rc != FDAS_OK) break;
            // the shape in use stays unless another is faster by the margin
            int best = corr_variant(L);
            for (int v = 0; v < (int)us.size(); ++v) {
                if (v < FDAS_TUNE_MAXV) rep->corr_us[k][v] = us[v];
                if (us[v] < kTuneMargin * us[best]) best = v;
            }
            set_corr_variant(L, best);
        }
        rep->corr_variant[k] = corr_variant(L);
        corr_variant_shape(L, rep->corr_variant[k], &rep->corr_threads[k], &rep->corr_steps[k],
                           &rep->corr_split[k], &rep->corr_form[k]);
    }
    // 2. harmonic-sum tile shape: whole trials at the first valid candidate length
    for (size_t i = 0; rc == FDAS_OK && i < cand.size() && base->numharm > 1; ++i) {
        if (!config_valid(cand[i], base->zmax, base->numharm)) continue;
        const int incumbent = harmsum_variant();
        int bestv = incumbent;
        bool timed = false;
        for (int v = 0; v < harmsum_variant_count() && v < FDAS_TUNE_MAXV; ++v) {
            double ms = 0.0;
            const int r = time_trial(*base, cand[i], s, &ms, v);
            if (r != FDAS_OK) {
                if (r != FDAS_ERR_INVALID && r != FDAS_ERR_UNSUPPORTED) rc = r;
                break;
            }
            rep->harmsum_ms[v] = ms;
            timed = true;
        }
        // the shape in use stays unless another is faster by the margin
        for (int v = 0; timed && v < harmsum_variant_count() && v < FDAS_TUNE_MAXV; ++v)
            if (rep->harmsum_ms[v] > 0.0 && rep->harmsum_ms[v] < kTuneMargin * rep->harmsum_ms[bestv]) bestv = v;
        set_harmsum_variant(bestv);
        rep->harmsum_variant = bestv;
        harmsum_variant_shape(bestv, &rep->harmsum_threads, &rep->harmsum_rows, &rep->harmsum_cols);
        break;
    }
    // 3. whole trials per candidate fundamental length
    int best = 0;
    double best_ms = 0.0;
    for (size_t i = 0; rc == FDAS_OK && i < cand.size(); ++i) {
        const int k = rep->ncand++;
        rep->cand_fftlen[k] = cand[i];
        rep->cand_ms_per_trial[k] = -1.0;
        if (!config_valid(cand[i], base->zmax, base->numharm)) continue;
        double ms = 0.0;
        const int r = time_trial(*base, cand[i], s, &ms);
        if (r == FDAS_ERR_INVALID || r == FDAS_ERR_UNSUPPORTED) continue;   // e.g. no steps
        if (r != FDAS_OK) {
            rc = r;
            break;
        }
        rep->cand_ms_per_trial[k] = ms;
        if (!best || ms < best_ms) {
            best = cand[i];
            best_ms = ms;
        }
    }
    (void)hipStreamDestroy(s);
    FDAS_TRY(rc);
    if (!best) {
        set_error("fdas_autotune: no candidate FFT length is valid for zmax %d, numharm %d, n %ld",
                  base->zmax, base->numharm, base->n);
        return FDAS_ERR_INVALID;
    }
    rep->fftlen = best;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned_fftlen[{base->zmax, base->numharm, base->precision}] = best;
    return FDAS_OK;
}

extern "C" int fdas_harmsum_variant_count(void) { return harmsum_variant_count(); }
extern "C" int fdas_get_harmsum_variant(void) { return harmsum_variant(); }
extern "C" int fdas_set_harmsum_variant(int variant) { return set_harmsum_variant(variant); }
