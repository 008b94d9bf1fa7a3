// precision.hpp -- storage formats of spectra, templates and power planes.
//
// FP32 mode stores complex values as two floats.  The reduced-precision modes
// store a complex value as two 16-bit halves packed in 32 bits (FP16 or BF16,
// round to nearest even; or int16 block floating point) together with one
// power-of-two exponent per template row
// or per segment, chosen so that the largest component maps below 2^15.  Scaling
// by a power of two is exact, so the exponents are folded into the power scale
// (|y|^2 scales with 2^(2 (e_template + e_segment))).  The reduced-precision modes
// also store the power planes in 16 bits (P16 below).  All FFT arithmetic and power
// accumulation stay in FP32.
#pragma once

#include <hip/hip_fp16.h>

#include <type_traits>

#include "fdas_common.hpp"

namespace fdas {

enum Prec : int { PREC_F32 = 0, PREC_F16 = 1, PREC_BF16 = 2, PREC_I16 = 3 };

__device__ __forceinline__ unsigned short f32_to_f16(float x) { return __half_as_ushort(__float2half_rn(x)); }
__device__ __forceinline__ float f16_to_f32(unsigned short b) { return __half2float(__ushort_as_half(b)); }
__device__ __forceinline__ unsigned short f32_to_bf16(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even (finite inputs)
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(unsigned short b) { return __uint_as_float(((unsigned)b) << 16); }

template <int P>
struct CStore;

template <>
struct CStore<PREC_F32> {
    using T = cplx;
    __device__ __forceinline__ static cplx get(T v) { return v; }
    __device__ __forceinline__ static T put(cplx v) { return v; }
};

template <>
struct CStore<PREC_F16> {
    using T = unsigned;
    __device__ __forceinline__ static cplx get(T v)
    {
        return mk(f16_to_f32((unsigned short)(v & 0xffffu)), f16_to_f32((unsigned short)(v >> 16)));
    }
    __device__ __forceinline__ static T put(cplx v)
    {
        return (unsigned)f32_to_f16(v.x) | ((unsigned)f32_to_f16(v.y) << 16);
    }
};

template <>
struct CStore<PREC_BF16> {
    using T = unsigned;
    __device__ __forceinline__ static cplx get(T v)
    {
        return mk(bf16_to_f32((unsigned short)(v & 0xffffu)), bf16_to_f32((unsigned short)(v >> 16)));
    }
    __device__ __forceinline__ static T put(cplx v)
    {
        return (unsigned)f32_to_bf16(v.x) | ((unsigned)f32_to_bf16(v.y) << 16);
    }
};

// 16-bit block floating point: two int16 with the row's / segment's exponent.
// The exponent puts the block's largest component in [2^14, 2^15), so values
// near the top of the block keep 15 significant bits (rounded to nearest,
// clamped to +-32767); smaller ones keep fewer.
__device__ __forceinline__ unsigned short f32_to_i16(float x)
{
    const float r = fminf(fmaxf(rintf(x), -32767.f), 32767.f);
    return (unsigned short)(short)r;
}
__device__ __forceinline__ float i16_to_f32(unsigned short b) { return (float)(short)b; }

template <>
struct CStore<PREC_I16> {
    using T = unsigned;
    __device__ __forceinline__ static cplx get(T v)
    {
        return mk(i16_to_f32((unsigned short)(v & 0xffffu)), i16_to_f32((unsigned short)(v >> 16)));
    }
    __device__ __forceinline__ static T put(cplx v)
    {
        return (unsigned)f32_to_i16(v.x) | ((unsigned)f32_to_i16(v.y) << 16);
    }
};

// Power-plane element: FP32 with FP32 storage; in the reduced-precision modes a
// 16-bit unsigned float ("P16": 6 exponent bits, 10 mantissa bits, no sign, since
// powers are non-negative) covering [2^-32, 2^32) with a fixed exponent bias, so no
// per-row exponent is needed.  A code c stands for the float whose bits are
// (c << 13) + (95 << 23): the top 16 bits of an FP32 value whose exponent is
// offset by 95 (2^-32).  Encoding rounds to nearest even (relative error at most
// 2^-11), sends values below 2^-32 to code 0 (read back as 2^-32) and saturates at
// 2^32 (1 - 2^-11).  Harmonic sums and thresholds stay FP32.
template <typename E>
struct PStore;
template <>
struct PStore<float> {
    __device__ __forceinline__ static float get(float v) { return v; }
    __device__ __forceinline__ static float put(float v) { return v; }
};
template <>
struct PStore<unsigned short> {
    static constexpr unsigned kBias = 95u;
    __device__ __forceinline__ static float get(unsigned short c)
    {
        return __uint_as_float(((unsigned)c << 13) + (kBias << 23));
    }
    __device__ __forceinline__ static unsigned short put(float v)
    {
        // Round to nearest even at bit 13 and remove the bias in one signed sum before
        // an arithmetic shift ((a - b 2^13) >> 13 == (a >> 13) - b exactly), so the
        // encoding is four instructions; the same code as shifting first for every
        // non-negative v, +inf and NaN included.
        const unsigned u = __float_as_uint(v);
        const int q = (int)(u + 0xfffu + ((u >> 13) & 1u) - (kBias << 23)) >> 13;
        return (unsigned short)min(max(q, 0), 65535);
    }
};

// Stores of one power-plane row of W elements through a buffer resource that covers
// exactly the row: the hardware's range check drops a store at an offset of W
// elements or more, so a kernel stores every one of its values without a branch, and
// the number of stores it has in flight is the same on every path (which lets the
// compiler's memory-wait counting keep loads issued before them in flight).  The
// byte offset of column col is col * sizeof(OutE): at least W elements for col >= W,
// and for a negative column (at least -2^16) a wrapped value near 2^32, far out of
// range too.  An offset must reach the store whole: the compiler may split one such
// as (base + constant) * size into a register part and an immediate, and a wrapped
// (negative) register part plus an immediate is not brought back into range by the
// hardware, which then drops in-range stores.  put() hides the offset's form behind
// an empty asm; put_linear() leaves it visible (measured 2-5% faster in the
// wavefront-split correlation's step loop), and tests/test_host_logic.py checks
// the built library for power stores with an immediate offset.  The descriptor's
// fourth word is the gfx9 one for raw (untyped) buffer access.
template <typename OutE>
struct RowStore {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ RowStore(OutE* row, int W)
        : r(__builtin_amdgcn_make_buffer_rsrc(row, (short)0, W * (int)sizeof(OutE), 0x00020000))
    {
    }
    __device__ __forceinline__ void put_at(unsigned off, float power) const
    {
        const OutE e = PStore<OutE>::put(power);
        if constexpr (sizeof(OutE) == 4)
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(e), r, (int)off, 0, 0);
        else
            __builtin_amdgcn_raw_buffer_store_b16(e, r, (int)off, 0, 0);
    }
    __device__ __forceinline__ void put(int col, float power) const
    {
        unsigned o = (unsigned)col * (unsigned)sizeof(OutE);
        asm("" : "+v"(o));   // no side effect: may be hoisted, never seen through
        put_at(o, power);
    }
    __device__ __forceinline__ void put_linear(int col, float power) const
    {
        put_at((unsigned)col * (unsigned)sizeof(OutE), power);
    }
};

// Plane element of storage mode P.
template <int P>
using PlaneT = typename std::conditional<P == PREC_F32, float, unsigned short>::type;
inline size_t plane_bytes(int prec) { return prec == PREC_F32 ? sizeof(float) : sizeof(unsigned short); }

// Exponent e with max_abs * 2^-e < 2^15 (0 for an all-zero block).
__device__ __forceinline__ int block_exponent(float max_abs)
{
    if (!(max_abs > 0.f)) return 0;
    int e2;
    (void)frexpf(max_abs, &e2);   // max_abs = m 2^e2, m in [0.5, 1)
    return e2 - 15;
}

inline size_t cstore_bytes(int prec) { return prec == PREC_F32 ? sizeof(cplx) : 4; }

}  // namespace fdas
