// segnorm.hpp -- the segment normalisation arithmetic of the reference
// (source/bpd_search.cu:157-177, source/bpd_math.c:20-105), shared by the
// FFT-correlation and direct-correlation segment kernels.
#pragma once

#include "fdas_common.hpp"

namespace fdas {

constexpr double kLn2 = 0.69314718055994530942;

__device__ __forceinline__ float quake_rsqrt_s(float x)
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

__device__ __forceinline__ float power_nc(cplx a)
{
#pragma clang fp contract(off)
    const float u = a.x * a.x;
    const float v = a.y * a.y;
    return u + v;
}

__device__ __forceinline__ cplx scale_nc(cplx a, float c)
{
#pragma clang fp contract(off)
    return mk(a.x * c, a.y * c);
}

// One 8-bit digit step of a workgroup radix select over non-negative float bit
// patterns.  Counts the candidates matching (bits & mask) == prefix by digit,
// then wavefront 0 finds the digit holding rank k.  Returns the digit and the
// rank within it through sel[].  hist has 256 entries.
template <int T>
__device__ __forceinline__ void select_digit(unsigned* hist, int* sel, int k, int t)
{
    __syncthreads();
    if (t < 64) {
        const unsigned c0 = hist[4 * t], c1 = hist[4 * t + 1], c2 = hist[4 * t + 2],
                       c3 = hist[4 * t + 3];
        const unsigned sum = c0 + c1 + c2 + c3;
        unsigned incl = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned y = __shfl_up(incl, o, 64);
            if (t >= o) incl += y;
        }
        const unsigned excl = incl - sum;
        if ((unsigned)k >= excl && (unsigned)k < incl) {
            unsigned r = (unsigned)k - excl;
            int d = 4 * t;
            if (r >= c0) {
                r -= c0; d++;
                if (r >= c1) {
                    r -= c1; d++;
                    if (r >= c2) { r -= c2; d++; }
                }
            }
            sel[0] = d;
            sel[1] = (int)r;
        }
    }
    __syncthreads();
}

}  // namespace fdas
