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
// Inter-pass twiddles come from a small table (per-pass blocks [k][r] of
// exp(+2 pi i r k / (Ns R)), built in double on the host) so every twiddle is
// correctly rounded; DIR < 0 (forward) uses the conjugate.
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

#ifndef FDAS_HD
#define FDAS_HD __host__ __device__ __forceinline__
#endif

#ifdef __HIPCC__
#include "fdas_common.hpp"
#endif

namespace fdas {

#ifdef __HIPCC__
__device__ __forceinline__ void block_sync() { __syncthreads(); }
#else
inline void block_sync() {}
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

// In-register radix-2 stages (all indices compile-time, unrolled by recursion).
template <int DIR, int LEN, int I, int K>
FDAS_HD void dft_butterflies(cplx* v)
{
    if constexpr (K < LEN / 2) {
        cplx a = v[I + K];
        cplx b = twc<DIR, K*(64 / LEN)>(v[I + K + LEN / 2]);
        v[I + K] = a + b;
        v[I + K + LEN / 2] = a - b;
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
    FDAS_HD static int pad(int i) { return i + (i >> 5); }

    template <int P>
    FDAS_HD static void compute(cplx* v, int t, const cplx* tw)
    {
        constexpr int R = radix(P), Ns = span(P), NB = E / R;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if (Ns > 1) {
                const int j = t + b * T;
                const int k = j & (Ns - 1);
                const cplx* tp = tw + twoff(P) + k * R;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    cplx w = tp[r];
                    if (DIR < 0) w.y = -w.y;
                    v[b * R + r] = cmul(v[b * R + r], w);
                }
            }
            dft_reg<R, DIR>(v + b * R);
        }
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

    // Passes P..NPASS-1 (each preceded by an LDS exchange).
    template <int P>
    FDAS_HD static void passes_from(cplx* v, cplx* lds, const cplx* tw, int t)
    {
        if constexpr (P < NPASS) {
            block_sync();
            store<P - 1>(v, lds, t);
            block_sync();
            load<P>(v, lds, t);
            compute<P>(v, t, tw);
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
                    tw[twoff(p) + k * R + r] = mk((float)cos(a), (float)sin(a));
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
            block_sync();
#pragma unroll
            for (int a = 0; a < NA; ++a) store<P - 1>(v[a], lds + a * LDS_ELEMS, t);
            block_sync();
#pragma unroll
            for (int a = 0; a < NA; ++a) load<P>(v[a], lds + a * LDS_ELEMS, t);
#pragma unroll
            for (int a = 0; a < NA; ++a) compute<P>(v[a], t, tw);
            passes_multi<NA, P + 1>(v, lds, tw, t);
        }
    }
    template <int NA>
    FDAS_HD static void run_multi(cplx (*v)[E], cplx* lds, const cplx* tw, int t)
    {
#pragma unroll
        for (int a = 0; a < NA; ++a) compute<0>(v[a], t, tw);
        passes_multi<NA, 1>(v, lds, tw, t);
    }
};

// ------------------------------------------------------------------ length 3M
// A transform of length N = 3M is three length-M workgroup transforms plus one
// radix-3 step that each thread applies to its own registers (no extra LDS
// exchange).  With w = exp(2 pi i / N):
//   forward:  x[n1 + M n2] (n1 = t + i T held in x[n2][i]) ->
//             b_r[n1] = w^(-r n1) DFT3(x[n1], x[n1 + M], x[n1 + 2M])[r],
//             FFT_M(b_r)[m] = X[3m + r]  ("decimated" order, r-major);
//   inverse:  V_r[n] = V[3n + r];  Y_r = IFFT_M(V_r);
//             y[k + M q] = sum_r exp(2 pi i r q / 3) w^(r k) Y_r[k].
// tw3[2k] = w^k and tw3[2k + 1] = w^(2k) for k in [0, M), rounded from double.

// DFT of three values in place, sign DIR: a, b, c <- y0, y1, y2.
template <int DIR>
FDAS_HD void dft3(cplx& a, cplx& b, cplx& c)
{
    constexpr float h = 8.660254038e-01f;   // sqrt(3) / 2
    const cplx s = b + c, d = b - c;
    const cplx m = mk(a.x - 0.5f * s.x, a.y - 0.5f * s.y);
    const cplx u = DIR > 0 ? mk(-h * d.y, h * d.x) : mk(h * d.y, -h * d.x);
    a = a + s;
    b = m + u;
    c = m - u;
}

inline void build_radix3_twiddles(int m, cplx* tw3)
{
    for (int k = 0; k < m; ++k) {
        const double a = 6.283185307179586476925286766559 * (double)k / (3.0 * (double)m);
        tw3[2 * k] = mk((float)cos(a), (float)sin(a));
        tw3[2 * k + 1] = mk((float)cos(2.0 * a), (float)sin(2.0 * a));
    }
}

template <int M, int T>
struct Radix3 {
    static constexpr int N = 3 * M;
    static constexpr int E = M / T;
    using FI = BlockFFT<M, T, 1>;

    // Forward radix-3 step before the three forward length-M transforms.
    FDAS_HD static void fwd_split(cplx (*x)[E], int t, const cplx* tw3)
    {
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int n1 = FI::in_index(t, i);
            dft3<-1>(x[0][i], x[1][i], x[2][i]);
            const cplx w1 = tw3[2 * n1], w2 = tw3[2 * n1 + 1];
            x[1][i] = cmul(x[1][i], mk(w1.x, -w1.y));
            x[2][i] = cmul(x[2][i], mk(w2.x, -w2.y));
        }
    }
    // Inverse radix-3 step after the three inverse length-M transforms: y[q][i]
    // becomes output out_index(t, i) + M q.
    FDAS_HD static void inv_combine(cplx (*y)[E], int t, const cplx* tw3)
    {
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int k = FI::out_index(t, i);
            y[1][i] = cmul(y[1][i], tw3[2 * k]);
            y[2][i] = cmul(y[2][i], tw3[2 * k + 1]);
            dft3<1>(y[0][i], y[1][i], y[2][i]);
        }
    }
    static void build_twiddles(cplx* tw3) { build_radix3_twiddles(M, tw3); }
};


}  // namespace fdas
