// whiten.hip -- red-noise whitening of the spectrum on the GPU.
//
// Restates deredden() (source/bpd_noise.c:12-92).  The block schedule (starts and
// lengths of the running-median blocks, and the bin ranges each normalisation
// segment covers) depends only on the bin count, so the host builds it once;
// the device then does (1) one wavefront per block: exact lower median of the
// block's powers, (2) one wavefront per segment: interpolated normalisation with
// the reference's float/double types and its fast inverse square root.
#include <cmath>
#include <mutex>

#include "fdas_internal.hpp"
#include "whiten_runs.hpp"

namespace fdas {

namespace {
constexpr double kLn2 = 0.69314718055994530942;  // log(2.0)
constexpr int kMaxBlock = 256;                   // longest median block (reference: 200)
}

// source/bpd_math.c:94-105 -- magic-constant guess + one Newton step, in float,
// each operation rounded separately.
__device__ __forceinline__ float quake_rsqrt(float x)
{
#pragma clang fp contract(off)
    int i = __float_as_int(x);
    i = 0x5f3759df - (i >> 1);
    const float y = __int_as_float(i);
    float h = 0.5f * x;
    h = h * y;
    h = h * y;
    return y * (1.5f - h);
}

__device__ __forceinline__ float power_of(cplx a)
{
#pragma clang fp contract(off)
    const float u = a.x * a.x;
    const float v = a.y * a.y;
    return u + v;
}

// Rank selection inside one wavefront: returns the element of 0-based rank k
// among vals[0..n) (n <= kMaxBlock), i.e. the value v with
// #{u < v} <= k < #{u <= v}.
__device__ float wave_select(const float* vals, int n, int k, int lane)
{
    float found = 0.0f;
    bool have = false;
    for (int c = lane; c < n; c += 64) {
        const float v = vals[c];
        int lt = 0, le = 0;
        for (int q = 0; q < n; ++q) {
            const float u = vals[q];
            lt += (u < v);
            le += (u <= v);
        }
        if (lt <= k && k < le) { found = v; have = true; }
    }
    const unsigned long long m = __ballot(have);
    const int src = __ffsll((long long)m) - 1;
    return __shfl(found, src < 0 ? 0 : src, 64);
}

__global__ __launch_bounds__(256) void whiten_median_kernel(const cplx* __restrict__ spec,
                                                            const int* __restrict__ start,
                                                            const int* __restrict__ len,
                                                            float* __restrict__ median,
                                                            int nblocks)
{
    __shared__ float buf[4][kMaxBlock];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int b = blockIdx.x * 4 + w;
    const bool valid = b < nblocks;   // uniform per wavefront
    const int s0 = valid ? start[b] : 0, n = valid ? len[b] : 0;
    for (int i = lane; i < n; i += 64) buf[w][i] = power_of(spec[s0 + i]);
    __syncthreads();
    if (valid) {
        const float m = wave_select(buf[w], n, (n - 1) / 2, lane);
        if (lane == 0) median[b] = m;
    }
}

// Segment j < nblocks covers [lo_j, hi_j): segment 0 uses the first block's mean
// without a slope; segment j >= 1 interpolates from block j-1 towards block j.
// Segment nblocks (the tail up to nbins) keeps the last slope (1.0 if the
// reference's loop never ran).
__global__ __launch_bounds__(256) void whiten_norm_kernel(cplx* __restrict__ spec,
                                                          const int* __restrict__ len,
                                                          const float* __restrict__ median,
                                                          const int* __restrict__ lo,
                                                          const int* __restrict__ hi,
                                                          int nblocks)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j = blockIdx.x * 4 + w;
    if (j > nblocks) return;
    float mean_old, dslope;
    bool flat = false;
    if (j == 0) {
        mean_old = (float)((double)median[0] / kLn2);
        dslope = 0.0f;
        flat = true;
        if (lane == 0) spec[0] = mk(1.0f, 0.0f);
    } else {
        const int jj = (j < nblocks) ? j : nblocks - 1;   // block whose slope applies
        if (jj == 0) {
            mean_old = (float)((double)median[0] / kLn2);
            dslope = 1.0f;
        } else {
            const float prev = (float)((double)median[jj - 1] / kLn2);
            const double mean_new = (double)median[jj] / kLn2;
            dslope = (float)((mean_new - (double)prev) / (0.5 * (double)(len[jj - 1] + len[jj])));
            mean_old = (j < nblocks) ? prev : (float)mean_new;
        }
    }
    const int a = lo[j], e = hi[j];
    const float norm_flat = quake_rsqrt(mean_old);
    for (int b = a + lane; b < e; b += 64) {
        float norm;
        if (flat) {
            norm = norm_flat;
        } else {
            const float t = dslope * (float)(b - a);
            norm = quake_rsqrt(mean_old + t);
        }
        cplx v = spec[b];
        v.x = v.x * norm;
        v.y = v.y * norm;
        spec[b] = v;
    }
}

// Host: the block and segment schedule of source/bpd_noise.c:33-88.
int whiten_schedule_create(unsigned nbins, WhitenSchedule* ws, DevAlloc* al, hipStream_t s)
{
    if (!al) al = hip_alloc();
    if (nbins < 16) {
        set_error("whitening needs at least 16 bins (got %u)", nbins);
        return FDAS_ERR_INVALID;
    }
    std::vector<int> start, len, lo, hi;
    int binnum = 1, buflen = 6;
    start.push_back(binnum);
    len.push_back(buflen);
    lo.push_back(1);
    hi.push_back(binnum + buflen / 2);
    binnum += buflen;
    buflen = (int)(6 * std::log((double)binnum));
    if (buflen > 200) buflen = 200;
    while ((unsigned)(binnum + buflen) < nbins) {
        start.push_back(binnum);
        len.push_back(buflen);
        lo.push_back(hi.back());
        hi.push_back(binnum + buflen / 2);
        binnum += buflen;
        buflen = (int)(6 * std::log((double)binnum));
        if (buflen > 200) buflen = 200;
    }
    lo.push_back(hi.back());
    hi.push_back((int)nbins);

    ws->nbins = nbins;
    ws->nblocks = (int)start.size();
    ws->nsegs = (int)lo.size();
    const size_t nb = start.size(), ns = lo.size();
    FDAS_TRY(al->alloc((void**)&ws->d_start, nb * sizeof(int)));
    FDAS_TRY(al->alloc((void**)&ws->d_len, nb * sizeof(int)));
    FDAS_TRY(al->alloc((void**)&ws->d_median, nb * sizeof(float)));
    FDAS_TRY(al->alloc((void**)&ws->d_seg_lo, ns * sizeof(int)));
    FDAS_TRY(al->alloc((void**)&ws->d_seg_hi, ns * sizeof(int)));
    FDAS_HIP(hipMemcpyAsync(ws->d_start, start.data(), nb * sizeof(int), hipMemcpyHostToDevice, s));
    FDAS_HIP(hipMemcpyAsync(ws->d_len, len.data(), nb * sizeof(int), hipMemcpyHostToDevice, s));
    FDAS_HIP(hipMemcpyAsync(ws->d_seg_lo, lo.data(), ns * sizeof(int), hipMemcpyHostToDevice, s));
    FDAS_HIP(hipMemcpyAsync(ws->d_seg_hi, hi.data(), ns * sizeof(int), hipMemcpyHostToDevice, s));
    FDAS_HIP(hipStreamSynchronize(s));
    return FDAS_OK;
}

void whiten_schedule_destroy(WhitenSchedule* ws, DevAlloc* al)
{
    if (!al) al = hip_alloc();
    al->release(ws->d_start);
    al->release(ws->d_len);
    al->release(ws->d_median);
    al->release(ws->d_seg_lo);
    al->release(ws->d_seg_hi);
    *ws = WhitenSchedule();
}

int whiten(cplx* d_spec, const WhitenSchedule& ws, hipStream_t s)
{
    hipLaunchKernelGGL(whiten_median_kernel, dim3((ws.nblocks + 3) / 4), dim3(256), 0, s, d_spec,
                       ws.d_start, ws.d_len, ws.d_median, ws.nblocks);
    FDAS_TRY(check_launch("whiten_median_kernel"));
    hipLaunchKernelGGL(whiten_norm_kernel, dim3((ws.nsegs + 3) / 4), dim3(256), 0, s, d_spec,
                       ws.d_len, ws.d_median, ws.d_seg_lo, ws.d_seg_hi, ws.nblocks);
    return check_launch("whiten_norm_kernel");
}

// ------------------------------------------------------------------ allocation-free form
// The block schedule as a run table (whiten_runs.hpp) passed by value in the kernel
// arguments, so whitening needs no device tables: only the block medians, which
// live in the file-scope array `scratch`.
namespace {
constexpr unsigned kNoallocMaxBins = 1u << 26;   // 663 030 median blocks
constexpr int kNoallocMaxBlocks = 1 << 20;       // 4 MiB of medians
constexpr int kMaxDevices = 64;

__global__ __launch_bounds__(256) void whiten_median_runs_kernel(const cplx* __restrict__ spec,
                                                                 const WhitenRuns w,
                                                                 float* __restrict__ median)
{
    __shared__ float buf[4][kMaxBlock];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int b = blockIdx.x * 4 + wv;
    const bool valid = b < w.nblocks;   // uniform per wavefront
    int s0 = 0, n = 0;
    if (valid) whiten_run_block(w, b, s0, n);
    for (int i = lane; i < n; i += 64) buf[wv][i] = power_of(spec[s0 + i]);
    __syncthreads();
    if (valid) {
        const float m = wave_select(buf[wv], n, (n - 1) / 2, lane);
        if (lane == 0) median[b] = m;
    }
}

// Segment j of whiten_norm_kernel with its bounds and block lengths from the runs:
// segment 0 is [1, 1 + len_0/2), segment j < nblocks is [hi_{j-1}, start_j + len_j/2),
// and the tail segment nblocks ends at nbins.
__global__ __launch_bounds__(256) void whiten_norm_runs_kernel(cplx* __restrict__ spec,
                                                               const WhitenRuns w,
                                                               const float* __restrict__ median)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int j = blockIdx.x * 4 + wv;
    const int nblocks = w.nblocks;
    if (j > nblocks) return;
    int sj = 0, lj = 0, sp = 0, lp = 0;
    if (j < nblocks) whiten_run_block(w, j, sj, lj);
    if (j > 0) whiten_run_block(w, j - 1, sp, lp);
    const int a = j == 0 ? 1 : sp + lp / 2;
    const int e = j == nblocks ? (int)w.nbins : sj + lj / 2;
    float mean_old, dslope;
    bool flat = false;
    if (j == 0) {
        mean_old = (float)((double)median[0] / kLn2);
        dslope = 0.0f;
        flat = true;
        if (lane == 0) spec[0] = mk(1.0f, 0.0f);
    } else {
        const int jj = (j < nblocks) ? j : nblocks - 1;   // block whose slope applies
        if (jj == 0) {
            mean_old = (float)((double)median[0] / kLn2);
            dslope = 1.0f;
        } else {
            int sq = 0, lq = 0, sr = 0, lr = 0;
            whiten_run_block(w, jj - 1, sq, lq);
            whiten_run_block(w, jj, sr, lr);
            const float prev = (float)((double)median[jj - 1] / kLn2);
            const double mean_new = (double)median[jj] / kLn2;
            dslope = (float)((mean_new - (double)prev) / (0.5 * (double)(lq + lr)));
            mean_old = (j < nblocks) ? prev : (float)mean_new;
        }
    }
    const float norm_flat = quake_rsqrt(mean_old);
    for (int b = a + lane; b < e; b += 64) {
        float norm;
        if (flat) {
            norm = norm_flat;
        } else {
            const float t = dslope * (float)(b - a);
            norm = quake_rsqrt(mean_old + t);
        }
        cplx v = spec[b];
        v.x = v.x * norm;
        v.y = v.y * norm;
        spec[b] = v;
    }
}

// Block medians of the allocation-free form.  Each GPU holds its own copy for the
// life of the process; every call on a GPU overwrites it (all medians a call reads
// are written by that call's median kernel first, so nothing needs clearing), and
// calls that overlap on one GPU race on it.
static __device__ __align__(256) float scratch[kNoallocMaxBlocks];

std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's `scratch` (looked up once per device; the host-side
// name is only a handle, never a device address).
int scratch_address(float** out)
{
    int dev = 0;
    FDAS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("fdas_deredden_noalloc: device %d beyond the %d supported", dev, kMaxDevices);
        return FDAS_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[dev];
    return FDAS_OK;
}
}  // namespace

}  // namespace fdas

using namespace fdas;

extern "C" int fdas_deredden_noalloc(fdas_complex* d_fft, unsigned numOfBins, void* hip_stream)
{
    if (!d_fft) {
        set_error("fdas_deredden_noalloc: null spectrum");
        return FDAS_ERR_INVALID;
    }
    if (numOfBins < 16 || numOfBins > kNoallocMaxBins) {
        set_error("fdas_deredden_noalloc: %u bins outside [16, %u]", numOfBins, kNoallocMaxBins);
        return FDAS_ERR_INVALID;
    }
    WhitenRuns w;
    if (whiten_runs(numOfBins, &w) != 0) {
        set_error("fdas_deredden_noalloc: more than %d block runs at %u bins", kWhitenMaxRuns, numOfBins);
        return FDAS_ERR_UNSUPPORTED;
    }
    if (w.nblocks > kNoallocMaxBlocks) {
        set_error("fdas_deredden_noalloc: %d median blocks exceed the %d of the scratch array",
                  w.nblocks, kNoallocMaxBlocks);
        return FDAS_ERR_UNSUPPORTED;
    }
    float* med = nullptr;
    FDAS_TRY(scratch_address(&med));
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    cplx* spec = reinterpret_cast<cplx*>(d_fft);
    hipLaunchKernelGGL(whiten_median_runs_kernel, dim3((w.nblocks + 3) / 4), dim3(256), 0, s, spec, w,
                       med);
    FDAS_TRY(check_launch("whiten_median_runs_kernel"));
    hipLaunchKernelGGL(whiten_norm_runs_kernel, dim3((w.nblocks + 1 + 3) / 4), dim3(256), 0, s, spec,
                       w, static_cast<const float*>(med));
    return check_launch("whiten_norm_runs_kernel");
}
