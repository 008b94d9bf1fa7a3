// fft_block.hpp -- workgroup-level Stockham FFT for CDNA4 (LDS-staged).
//
// One workgroup of T threads transforms one length-N sequence (N a power of two,
// N/T = E values per thread held in registers).  The transform is a sequence of
// radix-E passes (plus one smaller radix for the remainder); inside a pass each
// thread runs an in-register radix-R DFT with compile-time twiddles, and passes
// exchange data through LDS in Stockham (self-sorting) order, so the result comes
// out in natural order without a bit-reversal pass.
//
// Pass p with radix R and span Ns (product of the earlier radices):
//   butterfly j in [0, N/R) reads  x[j + r*N/R]              (r = 0..R-1)
//   multiplies by w^(r*k), w = exp(DIR 2 pi i / (Ns R)),  k = j mod Ns
//   runs DFT_R and writes          y[(j - k) R + k + r Ns]
// Thread t owns butterflies j = t + b*T, b < E/R.  Before pass 0 the caller
// places x[t + r*T] in v[r]; after the last pass v[b*R + r] holds
// y[t + b*T + r*(N/R)] (consecutive threads own consecutive outputs).
//
// Inter-pass twiddles come from a small table (per-pass blocks [r][k] of
// exp(+2 pi i r k / (Ns R)), built in double on the host) so every twiddle is
// correctly rounded; DIR < 0 (forward) uses the conjugate.  Consecutive threads
// have consecutive k, so each twiddle load of a wave reads one contiguous run.
//
// The phase functions are __host__ __device__ so the index arithmetic can be
// exercised by a host simulation of the workgroup (tests/test_fft_host_sim.py).
#pragma once

#ifndef __HIPCC__
#define __host__
#define __device__
#define __forceinline__ inline
#include <stdint.h>
#define FDAS_HOST_SIM 1
#endif
#include <math.h>

#include <type_traits>

#ifndef FDAS_HD
#define FDAS_HD __host__ __device__ __forceinline__
#endif

#ifdef __HIPCC__
#include "fdas_common.hpp"
#endif

namespace fdas {

#ifdef __HIPCC__
__device__ __forceinline__ void block_sync() { __syncthreads(); }
// Orders one wavefront's LDS accesses: the wavefront's lanes execute together and
// its LDS operations complete in issue order, so an exchange among the lanes of
// one wavefront only needs the compiler to keep every access on its side of the
// point (release and acquire fences at wavefront scope around a wave barrier).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
#elif defined(FDAS_HOST_BARRIER)
// A host test that runs each workgroup thread as an OS thread supplies the barriers
// (wave_sync: a barrier among the 64 threads of one wavefront).
void host_barrier();
void host_wave_barrier(int t);
inline void block_sync() { host_barrier(); }
#else
inline void block_sync() {}
#endif

#ifdef __HIPCC__
// Exchange of lane bit LB (3, 4 or 5) with a register bit, for a register pair x
// (bit clear) and y (bit set) of the lanes l and l ^ (1 << LB): the lane with bit LB
// set takes the partner's y into x, the lane with it clear the partner's x into y;
// the other register keeps its value.  Bits 4 and 5 are single instructions
// (v_permlane16_swap: 16-lane rows 1 and 3 of x with rows 0 and 2 of y;
// v_permlane32_swap: lanes 32..63 of x with lanes 0..31 of y); bit 3 sends
// through DPP row_ror:8, which reads lane l ^ 8 within each 16-lane row.
template <int LB>
__device__ __forceinline__ void lane_swap(float& x, float& y, int lane)
{
    static_assert(LB >= 3 && LB <= 5, "lane bits 3..5");
    const unsigned ux = __float_as_uint(x), uy = __float_as_uint(y);
    if constexpr (LB == 5) {
        const auto r = __builtin_amdgcn_permlane32_swap(ux, uy, false, false);
        x = __uint_as_float(r[0]);
        y = __uint_as_float(r[1]);
    } else if constexpr (LB == 4) {
        const auto r = __builtin_amdgcn_permlane16_swap(ux, uy, false, false);
        x = __uint_as_float(r[0]);
        y = __uint_as_float(r[1]);
    } else {
        const bool hi = lane & 8;
        const unsigned got = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(hi ? ux : uy), 0x128, 0xf, 0xf, false);
        if (hi) x = __uint_as_float(got);
        else y = __uint_as_float(got);
    }
}

// Lane-swap policy of BlockFFT::lane_exchange_last on the device: both
// components of a register pair through lane_swap.
struct LaneSwap {
    int lane;
    template <int LB>
    __device__ __forceinline__ void swap(cplx& x, cplx& y) const
    {
        lane_swap<LB>(x.x, y.x, lane);
        lane_swap<LB>(x.y, y.y, lane);
    }
};
#endif

#ifdef FDAS_HOST_SIM
struct alignas(8) cplx {
    float x, y;
};
FDAS_HD cplx mk(float x, float y) { cplx c; c.x = x; c.y = y; return c; }
FDAS_HD cplx operator+(cplx a, cplx b) { return mk(a.x + b.x, a.y + b.y); }
FDAS_HD cplx operator-(cplx a, cplx b) { return mk(a.x - b.x, a.y - b.y); }
FDAS_HD cplx cmul(cplx a, cplx b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }
#endif

#if defined(FDAS_HOST_SIM) && defined(FDAS_HOST_BARRIER)
// The same policy for a host test that runs each thread as an OS thread:
// host_lane_swap makes lane_swap's exchange for the calling thread's lane
// through memory shared by the threads of its wavefront.
void host_lane_swap(int lb, cplx& x, cplx& y);
struct LaneSwap {
    int lane;
    template <int LB>
    void swap(cplx& x, cplx& y) const
    {
        host_lane_swap(LB, x, y);
    }
};
#endif
#if defined(__HIPCC__) || defined(FDAS_HOST_BARRIER)
#define FDAS_LANE_SWAP 1
#endif

// cos and sin of 2*pi*m/64, m = 0..31 (float-rounded constants).
constexpr float kCos64[32] = {
    1.000000000e+00f, 9.951847196e-01f, 9.807852507e-01f, 9.569403529e-01f,
    9.238795042e-01f, 8.819212914e-01f, 8.314695954e-01f, 7.730104327e-01f,
    7.071067691e-01f, 6.343932748e-01f, 5.555702448e-01f, 4.713967443e-01f,
    3.826834261e-01f, 2.902846634e-01f, 1.950903237e-01f, 9.801714122e-02f,
    0.0f, -9.801714122e-02f, -1.950903237e-01f, -2.902846634e-01f,
    -3.826834261e-01f, -4.713967443e-01f, -5.555702448e-01f, -6.343932748e-01f,
    -7.071067691e-01f, -7.730104327e-01f, -8.314695954e-01f, -8.819212914e-01f,
    -9.238795042e-01f, -9.569403529e-01f, -9.807852507e-01f, -9.951847196e-01f};
constexpr float kSin64[32] = {
    0.0f, 9.801714122e-02f, 1.950903237e-01f, 2.902846634e-01f,
    3.826834261e-01f, 4.713967443e-01f, 5.555702448e-01f, 6.343932748e-01f,
    7.071067691e-01f, 7.730104327e-01f, 8.314695954e-01f, 8.819212914e-01f,
    9.238795042e-01f, 9.569403529e-01f, 9.807852507e-01f, 9.951847196e-01f,
    1.000000000e+00f, 9.951847196e-01f, 9.807852507e-01f, 9.569403529e-01f,
    9.238795042e-01f, 8.819212914e-01f, 8.314695954e-01f, 7.730104327e-01f,
    7.071067691e-01f, 6.343932748e-01f, 5.555702448e-01f, 4.713967443e-01f,
    3.826834261e-01f, 2.902846634e-01f, 1.950903237e-01f, 9.801714122e-02f};

constexpr int bitrev(int i, int bits)
{
    int r = 0;
    for (int b = 0; b < bits; ++b) r |= ((i >> b) & 1) << (bits - 1 - b);
    return r;
}

// b * exp(DIR * 2 pi i m / 64) for a compile-time m in [0, 32).
template <int DIR, int M>
FDAS_HD cplx twc(cplx b)
{
    constexpr float h = 7.071067691e-01f;
    if constexpr (M == 0) {
        return b;
    } else if constexpr (M == 16) {
        return DIR > 0 ? mk(-b.y, b.x) : mk(b.y, -b.x);
    } else if constexpr (M == 8) {
        return DIR > 0 ? mk((b.x - b.y) * h, (b.x + b.y) * h) : mk((b.x + b.y) * h, (b.y - b.x) * h);
    } else if constexpr (M == 24) {
        return DIR > 0 ? mk(-(b.x + b.y) * h, (b.x - b.y) * h) : mk((b.y - b.x) * h, -(b.x + b.y) * h);
    } else {
        constexpr float c = kCos64[M], s = DIR > 0 ? kSin64[M] : -kSin64[M];
        return mk(b.x * c - b.y * s, b.x * s + b.y * c);
    }
}

// Radix-2 butterfly (a, b) -> (a + w b, a - w b), w = exp(DIR 2 pi i M / 64) for a
// compile-time M, with the twiddle folded into fused multiply-adds: at M = 8 and
// 24 (w = (+-1 +- i)/sqrt 2) the sums b.x +- b.y are scaled inside the fused
// operations, and at a general M, a + w b takes four and a - w b = 2a - (a + w b)
// two, so a twiddled butterfly issues 6 instructions instead of 8.
template <int DIR, int M>
FDAS_HD void bfly(cplx& a, cplx& b)
{
    if constexpr (M == 0 || M == 16) {
        const cplx wb = twc<DIR, M>(b), t = a;
        a = t + wb;
        b = t - wb;
    } else if constexpr (M == 8 || M == 24) {
        constexpr float h = 7.071067691e-01f;
        // w b = h (px, py)
        float px, py;
        if constexpr (M == 8) {
            px = DIR > 0 ? b.x - b.y : b.x + b.y;
            py = DIR > 0 ? b.x + b.y : b.y - b.x;
        } else {
            px = DIR > 0 ? -(b.x + b.y) : b.y - b.x;
            py = DIR > 0 ? b.x - b.y : -(b.x + b.y);
        }
        const cplx t = a;
        a = mk(fmaf(px, h, t.x), fmaf(py, h, t.y));
        b = mk(fmaf(-px, h, t.x), fmaf(-py, h, t.y));
    } else {
        constexpr float c = kCos64[M], s = DIR > 0 ? kSin64[M] : -kSin64[M];
        const float y0x = fmaf(b.x, c, fmaf(-b.y, s, a.x));
        const float y0y = fmaf(b.x, s, fmaf(b.y, c, a.y));
        b = mk(fmaf(2.f, a.x, -y0x), fmaf(2.f, a.y, -y0y));
        a = mk(y0x, y0y);
    }
}

// In-register radix-2 stages (all indices compile-time, unrolled by recursion).
template <int DIR, int LEN, int I, int K>
FDAS_HD void dft_butterflies(cplx* v)
{
    if constexpr (K < LEN / 2) {
        bfly<DIR, K*(64 / LEN)>(v[I + K], v[I + K + LEN / 2]);
        dft_butterflies<DIR, LEN, I, K + 1>(v);
    }
}
template <int R, int DIR, int LEN, int I>
FDAS_HD void dft_groups(cplx* v)
{
    if constexpr (I < R) {
        dft_butterflies<DIR, LEN, I, 0>(v);
        dft_groups<R, DIR, LEN, I + LEN>(v);
    }
}
template <int R, int DIR, int LEN>
FDAS_HD void dft_stages(cplx* v)
{
    if constexpr (LEN <= R) {
        dft_groups<R, DIR, LEN, 0>(v);
        dft_stages<R, DIR, LEN * 2>(v);
    }
}
template <int R, int I>
FDAS_HD void bitrev_perm(cplx* v)
{
    if constexpr (I < R) {
        constexpr int J = bitrev(I, ilog2(R));
        if constexpr (J > I) {
            cplx t = v[I];
            v[I] = v[J];
            v[J] = t;
        }
        bitrev_perm<R, I + 1>(v);
    }
}

// Unnormalised DFT of R (<= 64) register values, sign DIR.
template <int R, int DIR>
FDAS_HD void dft_reg(cplx* v)
{
    bitrev_perm<R, 0>(v);
    dft_stages<R, DIR, 2>(v);
}

template <int N, int T, int DIR>
struct BlockFFT {
    static constexpr int E = N / T;
    static constexpr int LN = ilog2(N);
    static constexpr int LE = ilog2(E);
    static constexpr int NFULL = LN / LE;
    static constexpr int LREM = LN % LE;
    static constexpr int NPASS = NFULL + (LREM ? 1 : 0);
    static constexpr int RLAST = LREM ? (1 << LREM) : E;
    static constexpr int LDS_ELEMS = N + (N >> 5);
    static_assert((N & (N - 1)) == 0 && (T & (T - 1)) == 0 && E >= 2, "power-of-two sizes");

    static constexpr int radix(int p) { return p < NFULL ? E : RLAST; }
    static constexpr int span(int p)
    {
        int s = 1;
        for (int q = 0; q < p; ++q) s *= radix(q);
        return s;
    }
    static constexpr int twoff(int p)
    {
        int o = 0;
        for (int q = 1; q < p; ++q) o += span(q) * radix(q);
        return o;
    }
    static constexpr int TW_COUNT = twoff(NPASS);
    // LDS slot of exchange position i.  At 4 values per thread: the low 4 bits XORed
    // with bits 2..5 and 6..9, a unit-triangular map over GF(2), so a bijection on
    // each aligned 16-slot block; a 16-lane group of 8-byte accesses then covers
    // all 32 banks once in every store and load of every pass.  Otherwise one slot
    // of padding per 32: the same swizzle at 8 values per thread made the
    // 4096-point, 8-step shape 50% slower (more address arithmetic in a kernel at
    // its register limit), and the correlation kernels are VALU-bound.
    FDAS_HD static constexpr int pad(int i)
    {
        if constexpr (E == 4) return i ^ (((i >> 2) ^ (i >> 6)) & 15);
        else return i + (i >> 5);
    }
    static constexpr bool pad_is_injective()
    {
        bool seen[LDS_ELEMS] = {};
        for (int i = 0; i < N; ++i) {
            const int p = pad(i);
            if (p < 0 || p >= LDS_ELEMS || seen[p]) return false;
            seen[p] = true;
        }
        return true;
    }
    static_assert(pad_is_injective(), "exchange slot map must be one-to-one within the region");

    // Inter-pass twiddles of pass P for one thread, held in registers: the workgroup
    // exchange functions load them before the exchange that precedes the pass, so the
    // loads are in flight during its stores, barrier and loads (the compiler does not
    // move a load across a barrier).  Sub-harmonic correlations at 7168, 6144 and 5120
    // points: 7.62 -> 7.41 ms per trial (FP32).
    template <int P>
    struct Tw {
        static constexpr int R = radix(P), NB = E / R;
        static constexpr bool ANY = span(P) > 1;
        cplx w[ANY ? NB * (R - 1) : 1];
    };
    template <int P>
    FDAS_HD static void load_tw(Tw<P>& w, int t, const cplx* tw)
    {
        constexpr int R = radix(P), Ns = span(P), NB = E / R;
        if constexpr (Ns > 1) {
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int j = t + b * T;
                const int k = j & (Ns - 1);
                const cplx* tp = tw + twoff(P) + k;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    cplx c = tp[r * Ns];
                    if (DIR < 0) c.y = -c.y;
                    w.w[b * (R - 1) + r - 1] = c;
                }
            }
        }
    }
    template <int P>
    FDAS_HD static void compute_tw(cplx* v, const Tw<P>& w)
    {
        constexpr int R = radix(P), Ns = span(P), NB = E / R;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if constexpr (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) v[b * R + r] = cmul(v[b * R + r], w.w[b * (R - 1) + r - 1]);
            }
            dft_reg<R, DIR>(v + b * R);
        }
    }
    template <int P>
    FDAS_HD static void compute(cplx* v, int t, const cplx* tw)
    {
        Tw<P> w;
        load_tw<P>(w, t, tw);
        compute_tw<P>(v, w);
    }

    template <int P>
    FDAS_HD static void store(const cplx* v, cplx* lds, int t)
    {
        constexpr int R = radix(P), Ns = span(P), NB = E / R;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = t + b * T;
            const int k = j & (Ns - 1);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lds[pad(base + r * Ns)] = v[b * R + r];
        }
    }

    template <int P>
    FDAS_HD static void load(cplx* v, const cplx* lds, int t)
    {
        constexpr int R = radix(P), NB = E / R;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = t + b * T;
#pragma unroll
            for (int r = 0; r < R; ++r) v[b * R + r] = lds[pad(j + r * (N / R))];
        }
    }

    // Output position held in v[i] after the last pass.
    FDAS_HD static int out_index(int t, int i)
    {
        const int b = i / RLAST, r = i % RLAST;
        return t + b * T + r * (N / RLAST);
    }
    // Input position expected in v[i] before pass 0.
    FDAS_HD static int in_index(int t, int i) { return t + i * T; }

    // Last exchange of a wavefront transform (T == 64) in registers.  Pass NPASS-2
    // has radix E, one butterfly per lane (j = lane) and span Ns = 64 / RLAST; its
    // value r of lane Ns a + k lands at position Ns E a + Ns r + k, which the last
    // pass reads in lane Ns (r mod RLAST) + k, register (r / RLAST) RLAST + a.  In
    // each block of RLAST registers that is a transpose between the lane bits
    // [6 - log2 RLAST, 6) and the register bits [0, log2 RLAST): one lane-bit /
    // register-bit exchange per bit (sw.swap<lane bit>(x, y), see lane_swap) over
    // the register pairs that differ in that bit.
    static constexpr int XBITS = ilog2(RLAST);
    // Lane bits 4 and 5 only: at 4096 points and 512 threads (lane bits 3..5) the
    // DPP form of bit 3 (lane_swap) measured 3.5% slower than the LDS exchange.
    static constexpr bool LANE_LAST = T == 64 && NPASS >= 2 && XBITS >= 1 && XBITS <= 2;
    template <class SW, int I = 0>
    FDAS_HD static void lane_exchange_last(cplx* v, const SW& sw)
    {
        static_assert(T == 64 && NPASS >= 2 && XBITS >= 1 && XBITS <= 3, "lane bits 3..5 only");
        if constexpr (I < XBITS) {
#pragma unroll
            for (int b = 0; b < E / RLAST; ++b)
#pragma unroll
                for (int c = 0; c < RLAST; ++c)
                    if (!((c >> I) & 1))
                        sw.template swap<6 - XBITS + I>(v[b * RLAST + c], v[b * RLAST + c + (1 << I)]);
            lane_exchange_last<SW, I + 1>(v, sw);
        }
    }

    // Passes P..NPASS-1 of a transform run by one wavefront (T == 64) through an
    // LDS region of its own: the exchanges are ordered within the wavefront only.
    // SW (a lane-swap policy, see lane_exchange_last) makes the last exchange in
    // registers; without one every exchange goes through LDS.
    template <int P, class SW = void>
    FDAS_HD static void passes_from_wave(cplx* v, cplx* lds, const cplx* tw, int t, const SW* sw = nullptr)
    {
        static_assert(T == 64, "a wavefront transform has 64 threads");
        if constexpr (P < NPASS) {
            // (twiddles loaded at use: the correlation kernels' step loops hoist them)
            if constexpr (!std::is_void_v<SW> && P == NPASS - 1) {
                lane_exchange_last(v, *sw);
            } else {
                wave_barrier(t);
                store<P - 1>(v, lds, t);
                wave_barrier(t);
                load<P>(v, lds, t);
            }
            compute<P>(v, t, tw);
            passes_from_wave<P + 1, SW>(v, lds, tw, t, sw);
        }
    }
    FDAS_HD static void wave_barrier(int t)
    {
#ifdef __HIPCC__
        (void)t;
        wave_sync();
#elif defined(FDAS_HOST_BARRIER)
        host_wave_barrier(t);
#else
        (void)t;
#endif
    }

    // Passes P..NPASS-1 (each preceded by an LDS exchange).
    template <int P>
    FDAS_HD static void passes_from(cplx* v, cplx* lds, const cplx* tw, int t)
    {
        if constexpr (P < NPASS) {
            Tw<P> w;
            load_tw<P>(w, t, tw);
            block_sync();
            store<P - 1>(v, lds, t);
            block_sync();
            load<P>(v, lds, t);
            compute_tw<P>(v, w);
            passes_from<P + 1>(v, lds, tw, t);
        }
    }
    // Host: fill the inter-pass twiddle table (TW_COUNT entries, +sign).
    static void build_twiddles(cplx* tw)
    {
        for (int p = 1; p < NPASS; ++p) {
            const int R = radix(p), Ns = span(p);
            for (int k = 0; k < Ns; ++k)
                for (int r = 0; r < R; ++r) {
                    const double a = 6.283185307179586476925286766559 * (double)(r * k) / (double)(Ns * R);
                    tw[twoff(p) + r * Ns + k] = mk((float)cos(a), (float)sin(a));
                }
        }
    }
    // Full transform of the values in v (see the header comment for layouts).
    // Begins with a barrier, so lds may be reused between calls.
    FDAS_HD static void run(cplx* v, cplx* lds, const cplx* tw, int t)
    {
        compute<0>(v, t, tw);
        passes_from<1>(v, lds, tw, t);
    }

    // NA independent transforms in lock step, sharing each exchange's barriers;
    // transform a uses lds[a * LDS_ELEMS, (a + 1) * LDS_ELEMS).
    template <int NA, int P>
    FDAS_HD static void passes_multi(cplx (*v)[E], cplx* lds, const cplx* tw, int t)
    {
        if constexpr (P < NPASS) {
            Tw<P> w;
            load_tw<P>(w, t, tw);
            block_sync();
#pragma unroll
            for (int a = 0; a < NA; ++a) store<P - 1>(v[a], lds + a * LDS_ELEMS, t);
            block_sync();
#pragma unroll
            for (int a = 0; a < NA; ++a) load<P>(v[a], lds + a * LDS_ELEMS, t);
#pragma unroll
            for (int a = 0; a < NA; ++a) compute_tw<P>(v[a], w);
            passes_multi<NA, P + 1>(v, lds, tw, t);
        }
    }
    // The same NA transforms through one LDS region: each exchange is done for
    // one transform after another (more barriers, 1/NA of the LDS).
    template <int NA, int P>
    FDAS_HD static void passes_shared(cplx (*v)[E], cplx* lds, const cplx* tw, int t)
    {
        if constexpr (P < NPASS) {
            Tw<P> w;
            load_tw<P>(w, t, tw);
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                block_sync();
                store<P - 1>(v[a], lds, t);
                block_sync();
                load<P>(v[a], lds, t);
            }
#pragma unroll
            for (int a = 0; a < NA; ++a) compute_tw<P>(v[a], w);
            passes_shared<NA, P + 1>(v, lds, tw, t);
        }
    }
    template <int NA>
    FDAS_HD static void run_shared(cplx (*v)[E], cplx* lds, const cplx* tw, int t)
    {
#pragma unroll
        for (int a = 0; a < NA; ++a) compute<0>(v[a], t, tw);
        passes_shared<NA, 1>(v, lds, tw, t);
    }

    // The same through a ring of NR >= 2 slots of K LDS regions each: an
    // exchange moves a group of up to K transforms (all stores, one barrier, all
    // loads), and exchange c of a call uses slot c mod NR.  A thread's loads of
    // exchange c come before its stores of exchange c + 1 and so before barrier
    // c + 1, which every later store to any slot follows; so a store needs no
    // barrier of its own unless the same thread's previous exchange used the
    // same slot.  Within a call that never happens; across successive calls on
    // lds it happens only when the last exchange and the first share a slot,
    // and then a leading barrier is issued.  Between calls the caller must not
    // touch lds.
    static constexpr int NEXCH = NPASS - 1;
    template <int NA, int NR, int K>
    static constexpr bool ring_lead() { return (NEXCH * ((NA + K - 1) / K) - 1) % NR == 0; }
    template <int NA, int NR, int K, int P>
    FDAS_HD static void passes_ring(cplx (*v)[E], cplx* lds, const cplx* tw, int t)
    {
        if constexpr (P < NPASS) {
            constexpr int NG = (NA + K - 1) / K;
            Tw<P> w;
            load_tw<P>(w, t, tw);
#pragma unroll
            for (int gi = 0; gi < NG; ++gi) {
                cplx* slot = lds + (((P - 1) * NG + gi) % NR) * K * LDS_ELEMS;
                if (ring_lead<NA, NR, K>() && P == 1 && gi == 0) block_sync();
#pragma unroll
                for (int b = 0; b < K; ++b)
                    if (gi * K + b < NA) store<P - 1>(v[gi * K + b], slot + b * LDS_ELEMS, t);
                block_sync();
#pragma unroll
                for (int b = 0; b < K; ++b)
                    if (gi * K + b < NA) load<P>(v[gi * K + b], slot + b * LDS_ELEMS, t);
            }
#pragma unroll
            for (int a = 0; a < NA; ++a) compute_tw<P>(v[a], w);
            passes_ring<NA, NR, K, P + 1>(v, lds, tw, t);
        }
    }
    // lds holds NR * K * LDS_ELEMS.
    template <int NA, int NR = 2, int K = 1>
    FDAS_HD static void run_ring(cplx (*v)[E], cplx* lds, const cplx* tw, int t)
    {
        static_assert(NR >= 2 && K >= 1 && K <= NA, "a ring needs two slots of 1..NA regions");
#pragma unroll
        for (int a = 0; a < NA; ++a) compute<0>(v[a], t, tw);
        passes_ring<NA, NR, K, 1>(v, lds, tw, t);
    }
    // One transform (as run) through a ring of NR regions of LDS_ELEMS.
    template <int NR = 2>
    FDAS_HD static void run_ring1(cplx* v, cplx* lds, const cplx* tw, int t)
    {
        run_ring<1, NR, 1>(reinterpret_cast<cplx(*)[E]>(v), lds, tw, t);
    }

    template <int NA>
    FDAS_HD static void run_multi(cplx (*v)[E], cplx* lds, const cplx* tw, int t)
    {
#pragma unroll
        for (int a = 0; a < NA; ++a) compute<0>(v[a], t, tw);
        passes_multi<NA, 1>(v, lds, tw, t);
    }
};

// ------------------------------------------------------------------ wavefront split
// A length-N transform by T threads (W = T/64 wavefronts, E = N/T values each),
// decimated in time across the wavefronts.  With n = w + W m (w < W, m < NW =
// N/W) and k = k2 + NW k1 (k2 < NW, k1 < W), w = exp(DIR 2 pi i / N):
//   X[k2 + NW k1] = sum_w exp(DIR 2 pi i w k1 / W) (w^(w k2) F_w[k2]),
//   F_w = DFT_NW(x[w + W m], m < NW).
// Wavefront w computes F_w as a Stockham transform of its own (64 lanes x E
// values), whose exchanges go through an LDS band of its own and are ordered
// within the wavefront only, so one wavefront's LDS traffic overlaps the other
// wavefronts' arithmetic.  Then the twiddle w^(w k2), one exchange across the
// wavefronts (one workgroup barrier) that gives thread t the W values F_w[k2] of
// each of its columns k2 = t + T b (b < E/W), and a radix-W DFT in registers.
// Inputs come in the decimated order: v[i] holds x[in_index(t, i)] =
// x[w + W (lane + 64 i)], contiguous over the lanes once the input is stored as
// x'[w NW + m] = x[w + W m]; outputs come in natural order, contiguous over t.
template <int N, int T, int DIR>
struct WaveFFT {
    static constexpr int W = T / 64, E = N / T, NW = N / W, NB = E / W;
    using FW = BlockFFT<NW, 64, DIR>;
    static_assert(T % 64 == 0 && W >= 2 && FW::E == E && E % W == 0, "wavefront split shape");
    static constexpr int BAND = FW::LDS_ELEMS;      // one wavefront's region
    static constexpr int SET = W * BAND;            // the workgroup's regions
    // twiddle table: FW's inter-pass table, then exp(+2 pi i j / N), j < N
    static constexpr int TW_CROSS = FW::TW_COUNT;
    static constexpr int TW_COUNT = FW::TW_COUNT + N;

    FDAS_HD static int in_index(int t, int i) { return (t >> 6) + W * FW::in_index(t & 63, i); }
    FDAS_HD static int out_index(int t, int i) { return t + T * (i / W) + NW * (i % W); }

    static void build_twiddles(cplx* tw)
    {
        FW::build_twiddles(tw);
        for (int j = 0; j < N; ++j) {
            const double a = 6.283185307179586476925286766559 * (double)j / (double)N;
            tw[TW_CROSS + j] = mk((float)cos(a), (float)sin(a));
        }
    }

    // Wavefront part: F_w of the values in v (see in_index), left in v at
    // FW::out_index(lane, i), through band w of set.
    // The last of these exchanges is made in registers (LaneSwap) where the shape
    // allows it (FW::LANE_LAST).
    FDAS_HD static void wave_part(cplx* v, cplx* set, const cplx* tw, int t)
    {
        const int w = t >> 6, lane = t & 63;
        FW::template compute<0>(v, lane, tw);
#ifdef FDAS_LANE_SWAP
        if constexpr (FW::LANE_LAST) {
            const LaneSwap sw{lane};
            FW::template passes_from_wave<1>(v, set + w * BAND, tw, lane, &sw);
        } else {
            FW::template passes_from_wave<1>(v, set + w * BAND, tw, lane);
        }
#else
        FW::template passes_from_wave<1>(v, set + w * BAND, tw, lane);
#endif
    }
    // Cross part: twiddles, the exchange across wavefronts, radix-W DFTs.  The
    // exchange reuses band w of set for F_w (after the wavefront's own passes);
    // the caller alternates between two sets from call to call, so one barrier
    // per call suffices: a set is written again only two calls later, after the
    // barrier of the call in between, which every thread reaches after its loads.
    FDAS_HD static void cross_store(const cplx* v, cplx* set, const cplx* tw, int t)
    {
        const int w = t >> 6, lane = t & 63;
        cplx* band = set + w * BAND;
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int k2 = FW::out_index(lane, i);
            cplx c = tw[TW_CROSS + w * k2];
            if (DIR < 0) c.y = -c.y;
            band[FW::pad(k2)] = cmul(v[i], c);
        }
    }
    FDAS_HD static void cross_finish(cplx* v, const cplx* set, int t)
    {
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int ww = 0; ww < W; ++ww) v[b * W + ww] = set[ww * BAND + FW::pad(t + T * b)];
#pragma unroll
        for (int b = 0; b < NB; ++b) dft_reg<W, DIR>(v + b * W);
    }
    FDAS_HD static void cross_part(cplx* v, cplx* set, const cplx* tw, int t)
    {
        FW::wave_barrier(t & 63);   // the wavefront's last loads from its band are done
        cross_store(v, set, tw, t);
        block_sync();
        cross_finish(v, set, t);
    }
    FDAS_HD static void run(cplx* v, cplx* set, const cplx* tw, int t)
    {
        wave_part(v, set, tw, t);
        cross_part(v, set, tw, t);
    }
};

// ------------------------------------------------------------------ length R*M
// A transform of length N = R*M (R = 3, 5 or 7) is R length-M workgroup
// transforms plus one radix-R step that each thread applies to its own
// registers (no extra LDS exchange).  With w = exp(2 pi i / N):
//   forward:  x[n1 + M n2] (n1 = t + i T held in x[n2][i]) ->
//             b_r[n1] = w^(-r n1) DFT_R(x[n1 + M n2], n2 < R)[r],
//             FFT_M(b_r)[m] = X[R m + r]  ("decimated" order, r-major);
//   inverse:  V_r[n] = V[R n + r];  Y_r = IFFT_M(V_r);
//             y[k + M q] = sum_r exp(2 pi i r q / R) w^(r k) Y_r[k].
// tw[(R-1) k + r - 1] = w^(r k), r in [1, R), k in [0, M), rounded from double.

// cos and sin of 2 pi m / R (R = 3, 5, 7), float-rounded constants; after the
// callers' loops unroll every argument is a compile-time constant.
FDAS_HD constexpr float odd_cos(int R, int m)
{
    m %= R;
    if (m == 0) return 1.0f;
    if (m > R / 2) m = R - m;
    return R == 3 ? -5.000000000e-01f
         : R == 5 ? (m == 1 ? 3.090169944e-01f : -8.090169944e-01f)
         : (m == 1 ? 6.234898019e-01f : m == 2 ? -2.225209340e-01f : -9.009688679e-01f);
}
FDAS_HD constexpr float odd_sin(int R, int m)
{
    m %= R;
    if (m == 0) return 0.0f;
    const float sg = m > R / 2 ? -1.0f : 1.0f;
    if (m > R / 2) m = R - m;
    return sg * (R == 3 ? 8.660254038e-01f
               : R == 5 ? (m == 1 ? 9.510565163e-01f : 5.877852523e-01f)
               : (m == 1 ? 7.818314825e-01f : m == 2 ? 9.749279122e-01f : 4.338837391e-01f));
}

// DFT of R (odd) values in place, sign DIR, from the pair sums and differences
// s_p = x_p + x_(R-p), d_p = x_p - x_(R-p):
//   y_q, y_(R-q) = x_0 + sum_p cos(2 pi p q / R) s_p  +/- i DIR sum_p sin(2 pi p q / R) d_p.
template <int R, int DIR>
FDAS_HD void dft_odd(cplx* x)
{
    constexpr int H = R / 2;
    cplx sp[H], dp[H], y[R];
#pragma unroll
    for (int p = 1; p <= H; ++p) {
        sp[p - 1] = x[p] + x[R - p];
        dp[p - 1] = x[p] - x[R - p];
    }
    y[0] = x[0];
#pragma unroll
    for (int p = 0; p < H; ++p) y[0] = y[0] + sp[p];
#pragma unroll
    for (int q = 1; q <= H; ++q) {
        cplx a = x[0], b = mk(0.f, 0.f);
#pragma unroll
        for (int p = 1; p <= H; ++p) {
            const float c = odd_cos(R, p * q), sn = odd_sin(R, p * q);
            a = mk(a.x + c * sp[p - 1].x, a.y + c * sp[p - 1].y);
            b = mk(b.x + sn * dp[p - 1].x, b.y + sn * dp[p - 1].y);
        }
        const cplx ib = DIR > 0 ? mk(-b.y, b.x) : mk(b.y, -b.x);
        y[q] = a + ib;
        y[R - q] = a - ib;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = y[r];
}

inline void build_odd_twiddles(int R, int m, cplx* tw)
{
    for (int k = 0; k < m; ++k)
        for (int r = 1; r < R; ++r) {
            const double a = 6.283185307179586476925286766559 * (double)r * (double)k / ((double)R * (double)m);
            tw[(R - 1) * k + r - 1] = mk((float)cos(a), (float)sin(a));
        }
}

template <int R, int M, int T>
struct RadixOdd {
    static constexpr int N = R * M;
    static constexpr int E = M / T;
    using FI = BlockFFT<M, T, 1>;

    // Forward radix-R step before the R forward length-M transforms.
    FDAS_HD static void fwd_split(cplx (*x)[E], int t, const cplx* tw)
    {
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int n1 = FI::in_index(t, i);
            cplx u[R];
#pragma unroll
            for (int r = 0; r < R; ++r) u[r] = x[r][i];
            dft_odd<R, -1>(u);
            x[0][i] = u[0];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const cplx w = tw[(R - 1) * n1 + r - 1];
                x[r][i] = cmul(u[r], mk(w.x, -w.y));
            }
        }
    }
    // Inverse radix-R step after the R inverse length-M transforms: y[q][i]
    // becomes output out_index(t, i) + M q.
    FDAS_HD static void inv_combine(cplx (*y)[E], int t, const cplx* tw)
    {
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int k = FI::out_index(t, i);
            cplx u[R];
            u[0] = y[0][i];
#pragma unroll
            for (int r = 1; r < R; ++r) u[r] = cmul(y[r][i], tw[(R - 1) * k + r - 1]);
            dft_odd<R, 1>(u);
#pragma unroll
            for (int q = 0; q < R; ++q) y[q][i] = u[q];
        }
    }
    static void build_twiddles(cplx* tw) { build_odd_twiddles(R, M, tw); }
};

}  // namespace fdas
