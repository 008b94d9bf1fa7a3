// whiten.hip -- red-noise whitening of the spectrum on the GPU.
//
// Restates deredden() (source/bpd_noise.c:12-92).  The block schedule (starts and
// lengths of the running-median blocks, and the bin ranges each normalisation
// segment covers) depends only on the bin count, so the host builds it once;
// the device then does (1) one wavefront per block: exact lower median of the
// block's powers, (2) one wavefront per segment: interpolated normalisation with
// the reference's float/double types and its fast inverse square root.
#include <cmath>

#include "fdas_internal.hpp"

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

}  // namespace fdas
