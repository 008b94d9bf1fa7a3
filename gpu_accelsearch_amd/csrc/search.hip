// search.hip -- the overlap-save f-fdot loop on the GPU.
//
// Fused path (one launch per size class and harmonic, covering a batch of steps):
//   seg_prep_kernel   per (step, harmonic): fetch L/2 bins, exact lower median of
//                     the first numdata powers, scale by the fast inverse square
//                     root of median/ln2, forward FFT of length L/2.  The 2x spread
//                     of source/bpd_search.cu:56-67,181 makes the length-L forward
//                     transform two copies of this one, so only L/2 is stored.
//   corr_sub_kernel   per (template, step) of a sub-harmonic: multiply by the
//                     conjugate template, inverse FFT of length L in LDS, |y|^2/L^2
//                     of the kept columns -> sub-harmonic power plane.
//                     The fundamental uses the same kernel with its own plane;
//                     harmsum.hip then adds the sub-harmonic planes in place.
// The complex f-fdot plane is never written to memory.
//
// Stage path (reference data layout, source/bpd_search.cu:228-360): convolve,
// powers and add_powers kernels plus the normalise-and-spread kernel used by the
// reference-named entry points.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>

#include "fdas_internal.hpp"
#include "fft_block.hpp"
#include "precision.hpp"
#include "segnorm.hpp"

namespace fdas {

template <int N>
struct FftCfgS;
template <> struct FftCfgS<128>   { static constexpr int T = 64; };
template <> struct FftCfgS<256>   { static constexpr int T = 64; };
template <> struct FftCfgS<512>   { static constexpr int T = 64; };
template <> struct FftCfgS<1024>  { static constexpr int T = 64; };
template <> struct FftCfgS<2048>  { static constexpr int T = 128; };
template <> struct FftCfgS<4096>  { static constexpr int T = 256; };
template <> struct FftCfgS<8192>  { static constexpr int T = 256; };
template <> struct FftCfgS<16384> { static constexpr int T = 512; };

// ------------------------------------------------------------------ fused path
// Largest |re| or |im| over the workgroup's values (all threads get the result).
template <int T, int E>
__device__ __forceinline__ float block_max_abs(const cplx* v, float* red, int t)
{
    float m = 0.f;
#pragma unroll
    for (int i = 0; i < E; ++i) m = fmaxf(m, fmaxf(fabsf(v[i].x), fabsf(v[i].y)));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((t & 63) == 0) red[t >> 6] = m;
    __syncthreads();
    if (t == 0) {
        float r = red[0];
        for (int w = 1; w < T / 64; ++w) r = fmaxf(r, red[w]);
        red[T / 64] = r;
    }
    __syncthreads();
    return red[T / 64];
}

// Segment transforms of length R*NH (R = 1, or 3, 5, 7 for the lengths R*2^k):
// thread counts per power-of-two part.
template <int NH, int R>
constexpr int seg_threads()
{
    return R == 1 ? FftCfgS<NH>::T : R >= 5 ? (NH / 4 > 64 ? NH / 4 : 64) : (NH >= 1024 ? 128 : 64);
}

// Forward transform of one normalised segment of R*NH bins.  Bin n of the
// segment is held at n = in_index(t, i) + NH*r in v[r][i]; R > 1 adds the
// radix-R step, so the output comes in the decimated order A[R m + r] at
// r*NH + m (what the length-R*M correlation reads).
template <int NH, int R, int P>
__global__ __launch_bounds__((seg_threads<NH, R>())) void seg_prep_kernel(
    const cplx* __restrict__ spec_all, const SegDesc* __restrict__ desc,
    typename CStore<P>::T* __restrict__ seg, long long out_base, int* __restrict__ seg_exp,
    int eslot_base, const cplx* __restrict__ tw, const cplx* __restrict__ twr, int decim)
{
    constexpr int T = seg_threads<NH, R>();
    using F = BlockFFT<NH, T, -1>;
    constexpr int E = F::E;
    __shared__ cplx lds[R * F::LDS_ELEMS];
    __shared__ unsigned hist[256];
    __shared__ int sel[2];
    __shared__ float red[T / 64 + 1];
    const SegDesc d = desc[blockIdx.x];
    const cplx* __restrict__ spec = spec_all + d.spec_off;
    const int t = threadIdx.x;
    cplx v[R][E];
    unsigned pb[R][E];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int n = F::in_index(t, i) + NH * r;
            const long b = (long)d.lobin + n;
            v[r][i] = (b >= 0 && b < d.nbins) ? spec[b] : mk(0.f, 0.f);
            pb[r][i] = __float_as_uint(power_nc(v[r][i]));
        }
    // exact lower median: rank (numdata-1)/2 among the first numdata powers
    unsigned prefix = 0, mask = 0;
    int k = (d.numdata - 1) / 2;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int q = t; q < 256; q += T) hist[q] = 0;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int i = 0; i < E; ++i) {
                const int n = F::in_index(t, i) + NH * r;
                if (n < d.numdata && (pb[r][i] & mask) == prefix)
                    atomicAdd(&hist[(pb[r][i] >> shift) & 255u], 1u);
            }
        select_digit<T>(hist, sel, k, t);
        prefix |= (unsigned)sel[0] << shift;
        mask |= 255u << shift;
        k = sel[1];
    }
    const float med = __uint_as_float(prefix);
    const float c = quake_rsqrt_s((float)((double)med / kLn2));
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (F::in_index(t, i) + NH * r < d.numdata) v[r][i] = scale_nc(v[r][i], c);
    if constexpr (R > 1) RadixOdd<R, NH, T>::fwd_split(v, t, twr);
    F::template run_multi<R>(v, lds, tw, t);
    typename CStore<P>::T* o = seg + (d.out - out_base);
    // bin j of the transform at j (natural), or at (j mod W) NH/W + j/W (the
    // wavefront-decimated order of the wavefront-split correlation, R = 1)
    auto pos = [&](int r, int i) {
        const int j = F::out_index(t, i);
        if (R > 1 || decim == 0) return NH * r + j;
        return (j & ((1 << decim) - 1)) * (NH >> decim) + (j >> decim);
    };
    if constexpr (P == PREC_F32) {
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int i = 0; i < E; ++i) o[pos(r, i)] = v[r][i];
    } else {
        const int e = block_exponent(block_max_abs<T, R * E>(&v[0][0], red, t));
        if (t == 0) seg_exp[d.eslot - eslot_base] = e;
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int i = 0; i < E; ++i)
                o[pos(r, i)] = CStore<P>::put(mk(ldexpf(v[r][i].x, -e), ldexpf(v[r][i].y, -e)));
    }
}

// Template values of row z at this thread's first-pass positions (stored form),
// kept in registers while the workgroup walks its group of steps.
template <int L, int T = FftCfgS<L>::T, int P = PREC_F32>
__device__ __forceinline__ void load_template(const typename CStore<P>::T* __restrict__ K,
                                              typename CStore<P>::T* kr, int t)
{
    using F = BlockFFT<L, T, 1>;
#pragma unroll
    for (int i = 0; i < F::E; ++i) kr[i] = K[F::in_index(t, i)];
}

// y = IFFT_L(A[k mod L/2] * conj(K[k])), unnormalised.  Position t + i*T needs
// A[t + (i mod E/2)*T], so each thread loads E/2 segment values.
// NR: 1 one LDS region; >= 2 a ring of NR regions (BlockFFT::run_ring1).
template <int L, int T = FftCfgS<L>::T, int P = PREC_F32, int NR = 1>
__device__ __forceinline__ void corr_ifft(const typename CStore<P>::T* kr,
                                          const typename CStore<P>::T* __restrict__ A,
                                          const cplx* __restrict__ tw, cplx* lds, cplx* v, int t)
{
    using F = BlockFFT<L, T, 1>;
    constexpr int EH = F::E / 2;
    cplx a[EH];
#pragma unroll
    for (int i = 0; i < EH; ++i) a[i] = CStore<P>::get(A[t + i * T]);
#pragma unroll
    for (int i = 0; i < F::E; ++i) v[i] = cmulc(a[i % EH], CStore<P>::get(kr[i]));
    if constexpr (NR > 1) F::template run_ring1<NR>(v, lds, tw, t);
    else F::run(v, lds, tw, t);
}

// Workgroup -> (template row z, step group g) of the correlation kernels: rows
// fastest, over Z rounded up to a multiple of 8.  The hardware hands workgroups to
// the 8 XCDs round robin (id mod 8), so with the padding XCD k always gets the
// rows z = k mod 8 and its L2 keeps those template rows across steps (with Z odd
// the rows would rotate through every XCD and each step re-fetch the bank).  Ids
// of padding rows return before any barrier.  g0 offsets the step groups of one
// launch when a chunk needs several.
constexpr int kXcds = 8;

__device__ __forceinline__ bool corr_tile(int Z, int g0, int& z, int& g)
{
    const int zp = (Z + kXcds - 1) / kXcds * kXcds;
    const int id = blockIdx.x;
    z = id % zp;
    g = g0 + id / zp;
    return z < Z;
}

// Workgroup (z, g) (corr_tile) handles steps [g*G, min(g*G+G, nsteps))
// with the template row held in registers across them.  Reduced-precision
// inputs carry per-row / per-segment exponents, folded into the power scale.
// NR >= 2: the transform's exchanges go round a ring of NR LDS regions (one
// barrier each instead of two; with NR regions such that a step's last exchange
// and the next step's first use different ones, none between steps either).
template <int L, int G, int T, int P, typename OutE, int NR>
__global__ __launch_bounds__(T) void corr_sub_kernel(
    const typename CStore<P>::T* __restrict__ bank, const int* __restrict__ bank_exp, int Z,
    const typename CStore<P>::T* __restrict__ seg, long long seg_stride, long long seg_off,
    const int* __restrict__ seg_exp, int H, int h, int nsteps, int off, int W,
    OutE* __restrict__ out, const cplx* __restrict__ tw, float scale, int g0)
{
    using F = BlockFFT<L, T, 1>;
    using S = typename CStore<P>::T;
    __shared__ cplx lds[NR * F::LDS_ELEMS];
    const int t = threadIdx.x;
    int z, g;
    if (!corr_tile(Z, g0, z, g)) return;
    const int s0 = g * G;
    // G == 1: exactly one step (a loop would let the compiler hoist every twiddle)
    const int s1 = G > 1 ? min(s0 + G, nsteps) : s0 + 1;
    cplx v[F::E];
    [[maybe_unused]] S kr[G > 1 ? F::E : 1];
    if constexpr (G > 1) load_template<L, T, P>(bank + (size_t)z * L, kr, t);
    const int ek = P == PREC_F32 ? 0 : bank_exp[z];
    for (int s = s0; s < s1; ++s) {
        if constexpr (G > 1) {
            corr_ifft<L, T, P, NR>(kr, seg + s * seg_stride + seg_off, tw, lds, v, t);
        } else {
            S k1[F::E];
            load_template<L, T, P>(bank + (size_t)z * L, k1, t);
            corr_ifft<L, T, P, NR>(k1, seg + s * seg_stride + seg_off, tw, lds, v, t);
        }
        const float sc = P == PREC_F32 ? scale : ldexpf(scale, 2 * (ek + seg_exp[s * H + h]));
        const RowStore<OutE> o(out + ((size_t)s * Z + z) * W, W);
#pragma unroll
        for (int i = 0; i < F::E; ++i)
            o.put(F::out_index(t, i) - off, (v[i].x * v[i].x + v[i].y * v[i].y) * sc);
    }
}

// Wavefront-split form of the same correlation at a power of two L (WaveFFT):
// each of the W = T/64 wavefronts runs a length-L/W inverse transform with
// exchanges ordered inside the wavefront, then one exchange across wavefronts
// (one workgroup barrier per step; the steps alternate between two LDS sets).
// Segment and template rows come in the wavefront-decimated order
// x'[w L/W + m] = x[w + W m], so the inputs a thread needs, x[w + W (lane + 64 i)],
// are contiguous over the lanes: K'[w L/W + lane + 64 i], and, the product having
// period L/2, A'[w L/(2W) + lane + 64 (i mod E/2)].
template <int L, int G, int T, int P, typename OutE>
__global__ __launch_bounds__(T) void corr_wave_kernel(
    const typename CStore<P>::T* __restrict__ bank, const int* __restrict__ bank_exp, int Z,
    const typename CStore<P>::T* __restrict__ seg, long long seg_stride, long long seg_off,
    const int* __restrict__ seg_exp, int H, int h, int nsteps, int off, int W,
    OutE* __restrict__ out, const cplx* __restrict__ tw, float scale, int g0)
{
    using F = WaveFFT<L, T, 1>;
    using S = typename CStore<P>::T;
    constexpr int E = F::E, EH = E / 2, NW = F::NW;
    __shared__ cplx lds[2 * F::SET];
    const int t = threadIdx.x;
    int z, g;
    if (!corr_tile(Z, g0, z, g)) return;
    const int s0 = g * G;
    const int s1 = G > 1 ? min(s0 + G, nsteps) : s0 + 1;
    const int wv = t >> 6, lane = t & 63;
    const S* K = bank + (size_t)z * L + wv * NW + lane;
    S kr[E];
#pragma unroll
    for (int i = 0; i < E; ++i) kr[i] = K[64 * i];
    const int ek = P == PREC_F32 ? 0 : bank_exp[z];
    // A step's segment values are loaded one step ahead, during the previous step's
    // transform (the last step reloads its own, unused), and the powers go out
    // through row-bounded buffer stores (RowStore), so a step waits on no memory
    // access of its own.
    const S* Aw = seg + seg_off + wv * (NW / 2) + lane;
    S an[EH];
#pragma unroll
    for (int i = 0; i < EH; ++i) an[i] = Aw[s0 * seg_stride + 64 * i];
    for (int s = s0; s < s1; ++s) {
        cplx a[EH], v[E];
#pragma unroll
        for (int i = 0; i < EH; ++i) a[i] = CStore<P>::get(an[i]);
        const long long sn = min(s + 1, s1 - 1);
#pragma unroll
        for (int i = 0; i < EH; ++i) an[i] = Aw[sn * seg_stride + 64 * i];
#pragma unroll
        for (int i = 0; i < E; ++i) v[i] = cmulc(a[i % EH], CStore<P>::get(kr[i]));
        F::run(v, lds + ((s - s0) & 1) * F::SET, tw, t);
        const float sc = P == PREC_F32 ? scale : ldexpf(scale, 2 * (ek + seg_exp[s * H + h]));
        const RowStore<OutE> o(out + ((size_t)s * Z + z) * W, W);
#pragma unroll
        for (int i = 0; i < E; ++i)
            o.put_linear(F::out_index(t, i) - off, (v[i].x * v[i].x + v[i].y * v[i].y) * sc);
    }
}

// Length L = R*M form of the same correlation (R = 3, 5, 7).  The segment comes
// in the decimated order A'[r][m] = A[R m + r] (m < M/2) and the template row in
// the decimated order K'[r][n] = K[R n + r], so with X = [A, A] (period R M/2)
// the product at position R n + r is A'[r][n mod M/2] * conj K'[r][n]: R
// length-M inverse transforms in lock step, then the in-register radix-R step
// gives y[k + M q].
template <int R, int M, int G, int T, int P, typename OutE, int SH>
__global__ __launch_bounds__(T) void corr_odd_kernel(
    const typename CStore<P>::T* __restrict__ bank, const int* __restrict__ bank_exp, int Z,
    const typename CStore<P>::T* __restrict__ seg, long long seg_stride, long long seg_off,
    const int* __restrict__ seg_exp, int H, int h, int nsteps, int off, int W,
    OutE* __restrict__ out, const cplx* __restrict__ tw, const cplx* __restrict__ twr, float scale, int g0)
{
    using F = BlockFFT<M, T, 1>;
    using S = typename CStore<P>::T;
    constexpr int E = F::E, EH = E / 2, MH = M / 2;
    static_assert(SH <= 1 || (SH % 2 == 0 && SH / 2 <= R), "LDS form: 0, 1 or 2K with K <= R");
    __shared__ cplx lds[(SH == 0 ? R : SH) * F::LDS_ELEMS];
    const int t = threadIdx.x;
    int z, g;
    if (!corr_tile(Z, g0, z, g)) return;
    const int s0 = g * G;
    const int s1 = G > 1 ? min(s0 + G, nsteps) : s0 + 1;
    const S* K = bank + (size_t)z * R * M;
    [[maybe_unused]] S kr[G > 1 ? R : 1][G > 1 ? E : 1];
    if constexpr (G > 1) {
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int i = 0; i < E; ++i) kr[r][i] = K[M * r + F::in_index(t, i)];
    }
    const int ek = P == PREC_F32 ? 0 : bank_exp[z];
    for (int s = s0; s < s1; ++s) {
        const S* A = seg + s * seg_stride + seg_off;
        cplx v[R][E];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            cplx a[EH];
#pragma unroll
            for (int i = 0; i < EH; ++i) a[i] = CStore<P>::get(A[MH * r + t + i * T]);
#pragma unroll
            for (int i = 0; i < E; ++i) {
                const S kv = G > 1 ? kr[G > 1 ? r : 0][G > 1 ? i : 0] : K[M * r + F::in_index(t, i)];
                v[r][i] = cmulc(a[i % EH], CStore<P>::get(kv));
            }
        }
        if constexpr (SH == 1) F::template run_shared<R>(v, lds, tw, t);
        else if constexpr (SH >= 2) F::template run_ring<R, 2, SH / 2>(v, lds, tw, t);
        else F::template run_multi<R>(v, lds, tw, t);
        RadixOdd<R, M, T>::inv_combine(v, t, twr);
        const float sc = P == PREC_F32 ? scale : ldexpf(scale, 2 * (ek + seg_exp[s * H + h]));
        const RowStore<OutE> o(out + ((size_t)s * Z + z) * W, W);
#pragma unroll
        for (int q = 0; q < R; ++q)
#pragma unroll
            for (int i = 0; i < E; ++i)
                o.put(F::out_index(t, i) + M * q - off, (v[q][i].x * v[q][i].x + v[q][i].y * v[q][i].y) * sc);
    }
}

// Split form of the same correlation.  The product X * conj(K) has X = [A, A]
// (period L/2), so with k = k' + (L/2) j the length-L inverse transform splits
// into two length-L/2 transforms of A (no spread copy):
//   y[2m]   = IFFT_{L/2}(A * E)[m],  E[k'] = conj K[k'] + conj K[k' + L/2]
//   y[2m+1] = IFFT_{L/2}(A * O)[m],  O[k'] = (conj K[k'] - conj K[k' + L/2]) e^{2 pi i k'/L}
// E and O are precomputed per template row (split bank: row = [E | O]).  Half-
// length transforms need fewer registers per value, so more waves fit per SIMD.
template <int L, int G, int T, int P, typename OutE>
__global__ __launch_bounds__(T) void corr_split_kernel(
    const typename CStore<P>::T* __restrict__ bank, const int* __restrict__ bank_exp, int Z,
    const typename CStore<P>::T* __restrict__ seg, long long seg_stride, long long seg_off,
    const int* __restrict__ seg_exp, int H, int h, int nsteps, int off, int W,
    OutE* __restrict__ out, const cplx* __restrict__ tw, float scale, int g0)
{
    constexpr int LH = L / 2;
    using F = BlockFFT<LH, T, 1>;
    using S = typename CStore<P>::T;
    constexpr int E = F::E;
    __shared__ cplx lds[F::LDS_ELEMS];
    const int t = threadIdx.x;
    int z, g;
    if (!corr_tile(Z, g0, z, g)) return;
    const int s0 = g * G;
    const int s1 = G > 1 ? min(s0 + G, nsteps) : s0 + 1;
    const S* Krow = bank + (size_t)z * L;
    [[maybe_unused]] S ke[G > 1 ? E : 1], ko[G > 1 ? E : 1];
    if constexpr (G > 1) {
#pragma unroll
        for (int i = 0; i < E; ++i) {
            ke[i] = Krow[F::in_index(t, i)];
            ko[i] = Krow[LH + F::in_index(t, i)];
        }
    }
    const int ek = P == PREC_F32 ? 0 : bank_exp[z];
    for (int s = s0; s < s1; ++s) {
        const S* A = seg + s * seg_stride + seg_off;
        cplx a[E], v[E];
#pragma unroll
        for (int i = 0; i < E; ++i) a[i] = CStore<P>::get(A[F::in_index(t, i)]);
        const float sc = P == PREC_F32 ? scale : ldexpf(scale, 2 * (ek + seg_exp[s * H + h]));
        const RowStore<OutE> o(out + ((size_t)s * Z + z) * W, W);
#pragma unroll
        for (int half = 0; half < 2; ++half) {
#pragma unroll
            for (int i = 0; i < E; ++i) {
                S k;
                if constexpr (G > 1) k = half ? ko[i] : ke[i];
                else k = Krow[half * LH + F::in_index(t, i)];
                v[i] = cmul(a[i], CStore<P>::get(k));
            }
            F::run(v, lds, tw, t);
#pragma unroll
            for (int i = 0; i < E; ++i)
                o.put(2 * F::out_index(t, i) + half - off, (v[i].x * v[i].x + v[i].y * v[i].y) * sc);
        }
    }
}

// Split bank rows (see corr_split_kernel) from full rows of length L.
__global__ __launch_bounds__(256) void bank_split_kernel(const cplx* __restrict__ in, int L,
                                                         cplx* __restrict__ out, PhaseTable ph)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    const int LH = L / 2;
    if (k >= LH) return;
    const cplx* row = in + (size_t)z * L;
    const cplx lo = mk(row[k].x, -row[k].y), hi = mk(row[k + LH].x, -row[k + LH].y);
    const cplx d = mk(lo.x - hi.x, lo.y - hi.y);
    out[(size_t)z * L + k] = mk(lo.x + hi.x, lo.y + hi.y);
    out[(size_t)z * L + LH + k] = cmul(d, phase_at(ph, k));   // exp(2 pi i k / L)
}

// Reduced-precision copy of a template bank: per row the exponent and the packed
// values (one workgroup per row).
template <int P>
__global__ __launch_bounds__(256) void bank_pack_kernel(const cplx* __restrict__ in, int L,
                                                        typename CStore<P>::T* __restrict__ out,
                                                        int* __restrict__ row_exp)
{
    __shared__ float red[5];
    const int z = blockIdx.x, t = threadIdx.x;
    const cplx* row = in + (size_t)z * L;
    float m = 0.f;
    for (int i = t; i < L; i += 256) m = fmaxf(m, fmaxf(fabsf(row[i].x), fabsf(row[i].y)));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((t & 63) == 0) red[t >> 6] = m;
    __syncthreads();
    if (t == 0) red[4] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    __syncthreads();
    const int e = block_exponent(red[4]);
    if (t == 0) row_exp[z] = e;
    for (int i = t; i < L; i += 256)
        out[(size_t)z * L + i] = CStore<P>::put(mk(ldexpf(row[i].x, -e), ldexpf(row[i].y, -e)));
}

// ------------------------------------------------------------------ launchers
template <int NH, int R, int P>
static int seg_prep_t(const SegPrepArgs& a, hipStream_t s)
{
    constexpr int T = seg_threads<NH, R>();
    int err = FDAS_OK;
    const cplx* tw = fft_twiddles_t(NH, T, &err);
    if (!tw) return err;
    const cplx* twr = nullptr;
    if constexpr (R > 1) {
        twr = fft_odd_twiddles(R, NH, &err);
        if (!twr) return err;
    }
    hipLaunchKernelGGL((seg_prep_kernel<NH, R, P>), dim3(a.nseg), dim3(T), 0, s, a.spec, a.desc,
                       static_cast<typename CStore<P>::T*>(a.seg), a.out_base, a.seg_exp,
                       a.eslot_base, tw, twr, a.decim);
    return check_launch("seg_prep_kernel");
}

template <int NH, int R = 1>
static int seg_prep_p(const SegPrepArgs& a, hipStream_t s)
{
    switch (a.prec) {
        case PREC_F16: return seg_prep_t<NH, R, PREC_F16>(a, s);
        case PREC_BF16: return seg_prep_t<NH, R, PREC_BF16>(a, s);
        case PREC_I16: return seg_prep_t<NH, R, PREC_I16>(a, s);
        default: return seg_prep_t<NH, R, PREC_F32>(a, s);
    }
}

int launch_segment_prep(int L, const SegPrepArgs& a, hipStream_t s)
{
    if (a.nseg <= 0) return FDAS_OK;
    switch (L / 2) {
        case 128: return seg_prep_p<128>(a, s);
        case 256: return seg_prep_p<256>(a, s);
        case 512: return seg_prep_p<512>(a, s);
        case 1024: return seg_prep_p<1024>(a, s);
        case 2048: return seg_prep_p<2048>(a, s);
        case 4096: return seg_prep_p<4096>(a, s);
        case 8192: return seg_prep_p<8192>(a, s);
        case 384: return seg_prep_p<128, 3>(a, s);
        case 768: return seg_prep_p<256, 3>(a, s);
        case 1536: return seg_prep_p<512, 3>(a, s);
        case 3072: return seg_prep_p<1024, 3>(a, s);
        case 6144: return seg_prep_p<2048, 3>(a, s);
        case 640: return seg_prep_p<128, 5>(a, s);
        case 1280: return seg_prep_p<256, 5>(a, s);
        case 2560: return seg_prep_p<512, 5>(a, s);
        case 5120: return seg_prep_p<1024, 5>(a, s);
        case 3584: return seg_prep_p<512, 7>(a, s);
        default:
            set_error("segment transform length %d unsupported (FFT length %d)", L / 2, L);
            return FDAS_ERR_UNSUPPORTED;
    }
}

// Launch shapes of the fused correlation per FFT length: (threads, steps per
// workgroup, LDS form).  Variant 0 is the default; the autotuner may select
// another per length.  More threads give more waves per SIMD and fewer values
// per thread (a smaller radix, so one more LDS pass).  More steps per workgroup
// load the template row and the hoisted twiddles fewer times (with one LDS
// region 16 measured 3-5% faster than 8 at 4096 and 8192) but leave fewer
// workgroups for short trials.  At 16384 the registers are needed for the
// transform itself, so every step reloads its template row.  Two alternating
// LDS regions (S = 2) halve the barriers of each exchange: 7% faster at 4096
// and 2% at 8192 (512 threads, one workgroup per CU, where every barrier idles
// the CU), slower with 1024 threads; with them 8 steps per workgroup beat 16 by
// 3% at 4096.  Three regions (S = 3) would also drop the barrier between steps
// (3 exchanges per transform), but at 4096 they ran 1.55x slower: 101 KB of
// LDS leaves one workgroup per CU instead of two.  At 8192 they do not fit.
namespace {
struct CorrShape {
    int T, G, S;   // threads, steps per workgroup; S at powers of two: 1 split form
                   // (two L/2 transforms), 2 a ring of two LDS regions, 8 the
                   // wavefront-split form (corr_wave_kernel); at lengths
                   // R*2^k: 0 one region per transform, 1 one region shared by the
                   // R transforms, 2K a ring of two slots of K regions (K
                   // transforms exchanged per barrier)
};
template <int L>
struct Shapes {
    static constexpr int n = 0;
};
#define FDAS_SHAPES(NN, ...)                                               \
    template <>                                                            \
    struct Shapes<NN> {                                                    \
        static constexpr CorrShape v[] = {__VA_ARGS__};                    \
        static constexpr int n = (int)(sizeof(v) / sizeof(v[0]));          \
    };
FDAS_SHAPES(256, {64, 8, 0})
FDAS_SHAPES(512, {64, 8, 0})
FDAS_SHAPES(1024, {64, 8, 0}, {128, 8, 0})
FDAS_SHAPES(2048, {256, 8, 0}, {128, 8, 0}, {128, 8, 1}, {128, 16, 0}, {256, 8, 2}, {128, 8, 8})
// The wavefront-split form leads at 4096 and 8192: 10-11% faster than the
// two-region ring at 4096 (FP32 and I16 storage) and 7% at 8192 (128 templates x
// 256 steps per launch); slower at 2048.
FDAS_SHAPES(4096, {512, 16, 8}, {512, 8, 2}, {512, 16, 2}, {512, 16, 0}, {512, 8, 0}, {256, 8, 0},
            {512, 4, 0}, {256, 8, 1}, {128, 8, 1}, {512, 8, 8}, {256, 8, 8})
// At 8192, 32 steps per workgroup in the wavefront-split form (last variant) measured 2-3%
// faster on a 1084-step trial, but a short trial gets half the workgroups (135 steps of
// 56 rows: 280 workgroups, 1.1 per CU), so it is left to the tuner.
FDAS_SHAPES(8192, {512, 16, 8}, {512, 16, 2}, {512, 16, 0}, {512, 8, 0}, {512, 32, 0}, {1024, 8, 0},
            {512, 8, 1}, {256, 8, 1}, {512, 8, 2}, {512, 8, 8}, {512, 32, 8})
FDAS_SHAPES(16384, {512, 1, 0}, {1024, 1, 0}, {512, 1, 1}, {1024, 4, 1})
// Lengths R*2^k (R = 3, 5, 7): R transforms of L/R per step and no split form; S = 1
// runs their LDS exchanges one transform after another through one region (1/R
// of the LDS, so more workgroups per CU: 9-15% faster at 2560, 5120 and 6144;
// 1536 fastest with its own region per transform).  S = 2 does the same through
// two alternating regions, one barrier per exchange: faster again by 10% at
// 10240, 8% at 7168 and 6144, 6% at 12288, 1-2% at 3072 and 5120; slower at 2560.
// S = 2K (K > 1) exchanges K transforms per barrier through 2K regions: 3%
// faster at 7168 with K = 2 (the default there), but 1.2-1.4x slower at 5120 and 6144 (the LDS
// cuts the workgroups per CU), and slower with K = 3 or 4.  At 3072, 128
// threads (8 values per transform) beat 256 by 5%; at 6144 and 12288 fewer
// threads than the defaults were 13-20% slower.
// One step per workgroup measured fastest here (the R transforms' values and
// twiddles already take the registers a cached template row would need).
FDAS_SHAPES(768, {64, 1, 1}, {64, 1, 0}, {64, 8, 0})
FDAS_SHAPES(1536, {64, 1, 0}, {128, 8, 0}, {64, 1, 1})
FDAS_SHAPES(3072, {128, 1, 2}, {256, 1, 2}, {256, 1, 1}, {256, 1, 0}, {128, 1, 0}, {256, 8, 0})
FDAS_SHAPES(6144, {256, 1, 2}, {256, 1, 1}, {512, 1, 1}, {256, 1, 0}, {512, 1, 0}, {512, 8, 0})
FDAS_SHAPES(12288, {512, 1, 2}, {512, 1, 1}, {512, 1, 0}, {256, 1, 0})
FDAS_SHAPES(1280, {64, 1, 1}, {64, 1, 0})
FDAS_SHAPES(2560, {128, 1, 1}, {128, 1, 0}, {64, 1, 0})
FDAS_SHAPES(5120, {128, 1, 2}, {128, 1, 1}, {256, 1, 1}, {256, 1, 0}, {128, 1, 0})
FDAS_SHAPES(10240, {256, 1, 2}, {256, 1, 1}, {256, 1, 0}, {512, 1, 0})
FDAS_SHAPES(7168, {256, 1, 4}, {256, 1, 2}, {128, 1, 2}, {256, 1, 1}, {128, 1, 1}, {256, 1, 0})
#undef FDAS_SHAPES

struct ShapeEntry {
    int L;
    const CorrShape* v;
    int n;
};
#define FDAS_ENTRY(NN) {NN, Shapes<NN>::v, Shapes<NN>::n}
constexpr ShapeEntry kShapeTable[] = {
    FDAS_ENTRY(256),   FDAS_ENTRY(512),   FDAS_ENTRY(1024),  FDAS_ENTRY(2048),  FDAS_ENTRY(4096),
    FDAS_ENTRY(8192),  FDAS_ENTRY(16384), FDAS_ENTRY(768),   FDAS_ENTRY(1536),  FDAS_ENTRY(3072),
    FDAS_ENTRY(6144),  FDAS_ENTRY(12288), FDAS_ENTRY(1280),  FDAS_ENTRY(2560),  FDAS_ENTRY(5120),
    FDAS_ENTRY(10240), FDAS_ENTRY(7168)};
#undef FDAS_ENTRY
constexpr int kNumLengths = (int)(sizeof(kShapeTable) / sizeof(kShapeTable[0]));

int shape_slot(int L)
{
    for (int k = 0; k < kNumLengths; ++k)
        if (kShapeTable[k].L == L) return k;
    return -1;
}

const CorrShape* shapes_of(int L, int* count)
{
    const int k = shape_slot(L);
    *count = k < 0 ? 0 : kShapeTable[k].n;
    return k < 0 ? nullptr : kShapeTable[k].v;
}

std::atomic<int> g_shape[kNumLengths];
}  // namespace

int corr_variant_count(int L)
{
    int c = 0;
    (void)shapes_of(L, &c);
    return c;
}

int corr_variant(int L)
{
    const int k = shape_slot(L);
    return k < 0 ? -1 : g_shape[k].load();
}

int set_corr_variant(int L, int v)
{
    const int k = shape_slot(L);
    if (k < 0 || v < 0 || v >= corr_variant_count(L)) {
        set_error("no correlation variant %d at FFT length %d", v, L);
        return FDAS_ERR_INVALID;
    }
    g_shape[k].store(v);
    return FDAS_OK;
}

int corr_variant_decim(int L, int v)
{
    int c = 0;
    const CorrShape* sh = shapes_of(L, &c);
    if (!sh || v < 0 || v >= c || odd_factor(L) != 1 || sh[v].S != 8) return 0;
    return ilog2(sh[v].T / 64);
}

void corr_variant_shape(int L, int v, int* threads, int* steps, int* split, int* form)
{
    int c = 0;
    const CorrShape* sh = shapes_of(L, &c);
    const bool ok = sh && v >= 0 && v < c;
    *threads = ok ? sh[v].T : 0;
    *steps = ok ? sh[v].G : 0;
    if (split) *split = ok && odd_factor(L) == 1 && sh[v].S == 1;
    if (form) *form = ok ? sh[v].S : 0;
}

// Launches of one correlation: step groups [g0, g0 + n) per launch, n chosen so
// that (Z rounded up to a multiple of 8) * n workgroups of T threads stay below
// 2^31 work-items.
template <int T, int G, typename Launch>
static int corr_launches(const CorrArgs& a, Launch&& launch)
{
    const long groups = (a.nsteps + G - 1) / G;
    const long zr = (a.Z + kXcds - 1) / kXcds;
    const long per = std::max(1L, (2147483647L / T) / (kXcds * zr));
    for (long g0 = 0; g0 < groups; g0 += per) {
        const long n = std::min(per, groups - g0);
        const int rc = launch(dim3((unsigned)(kXcds * zr * n)), (int)g0);
        if (rc != FDAS_OK) return rc;
    }
    return FDAS_OK;
}

template <int L, int T, int G, int SPLIT, int P, typename OutE>
static int corr_t(const CorrArgs& a, hipStream_t s)
{
    using S = typename CStore<P>::T;
    int err = FDAS_OK;
    const cplx* tw = SPLIT == 8 ? fft_wave_twiddles(L, T, &err)
                                : fft_twiddles_t(SPLIT == 1 ? L / 2 : L, T, &err);
    if (!tw) return err;
    const float scale = (float)(1.0 / ((double)L * (double)L));
    if constexpr (SPLIT == 8) {
        return corr_launches<T, G>(a, [&](dim3 grid, int g0) {
            hipLaunchKernelGGL((corr_wave_kernel<L, G, T, P, OutE>), grid, dim3(T), 0, s,
                               static_cast<const S*>(a.bank), a.bank_exp, a.Z,
                               static_cast<const S*>(a.seg), a.seg_stride, a.seg_off, a.seg_exp,
                               a.H, a.h, a.nsteps, a.off, a.W, static_cast<OutE*>(a.out), tw, scale,
                               g0);
            return check_launch("corr_wave_kernel");
        });
    } else if constexpr (SPLIT == 1) {
        if (!a.bank_split) {
            set_error("split correlation at length %d without a split bank", L);
            return FDAS_ERR_INVALID;
        }
        return corr_launches<T, G>(a, [&](dim3 grid, int g0) {
            hipLaunchKernelGGL((corr_split_kernel<L, G, T, P, OutE>), grid, dim3(T), 0, s,
                               static_cast<const S*>(a.bank_split), a.bank_split_exp, a.Z,
                               static_cast<const S*>(a.seg), a.seg_stride, a.seg_off, a.seg_exp,
                               a.H, a.h, a.nsteps, a.off, a.W, static_cast<OutE*>(a.out), tw, scale,
                               g0);
            return check_launch("corr_split_kernel");
        });
    } else {
        return corr_launches<T, G>(a, [&](dim3 grid, int g0) {
            hipLaunchKernelGGL((corr_sub_kernel<L, G, T, P, OutE, SPLIT == 0 ? 1 : SPLIT>), grid, dim3(T), 0, s,
                               static_cast<const S*>(a.bank), a.bank_exp, a.Z,
                               static_cast<const S*>(a.seg), a.seg_stride, a.seg_off, a.seg_exp,
                               a.H, a.h, a.nsteps, a.off, a.W, static_cast<OutE*>(a.out), tw, scale,
                               g0);
            return check_launch("corr_sub_kernel");
        });
    }
}

// Power planes are FP32 with FP32 storage and P16 in the reduced-precision modes.
template <int L, int T, int G, int SPLIT>
static int corr_p(const CorrArgs& a, hipStream_t s)
{
    switch (a.prec) {
        case PREC_F16: return corr_t<L, T, G, SPLIT, PREC_F16, PlaneT<PREC_F16>>(a, s);
        case PREC_BF16: return corr_t<L, T, G, SPLIT, PREC_BF16, PlaneT<PREC_BF16>>(a, s);
        case PREC_I16: return corr_t<L, T, G, SPLIT, PREC_I16, PlaneT<PREC_I16>>(a, s);
        default: return corr_t<L, T, G, SPLIT, PREC_F32, PlaneT<PREC_F32>>(a, s);
    }
}

template <int L, int V = 0>
static int corr_l(const CorrArgs& a, hipStream_t s, int v)
{
    if constexpr (V < Shapes<L>::n) {
        constexpr CorrShape sh = Shapes<L>::v[V];
        if (v == V) return corr_p<L, sh.T, sh.G, sh.S>(a, s);
        return corr_l<L, V + 1>(a, s, v);
    } else {
        set_error("correlation variant %d out of range at length %d", v, L);
        return FDAS_ERR_INVALID;
    }
}

template <int R, int M, int T, int G, int SH, int P>
static int corr_odd_t(const CorrArgs& a, hipStream_t s)
{
    using S = typename CStore<P>::T;
    int err = FDAS_OK;
    const cplx* tw = fft_twiddles_t(M, T, &err);
    if (!tw) return err;
    const cplx* twr = fft_odd_twiddles(R, M, &err);
    if (!twr) return err;
    const double L = (double)R * M;
    const float scale = (float)(1.0 / (L * L));
    return corr_launches<T, G>(a, [&](dim3 grid, int g0) {
        hipLaunchKernelGGL((corr_odd_kernel<R, M, G, T, P, PlaneT<P>, SH>), grid, dim3(T), 0, s,
                           static_cast<const S*>(a.bank), a.bank_exp, a.Z, static_cast<const S*>(a.seg),
                           a.seg_stride, a.seg_off, a.seg_exp, a.H, a.h, a.nsteps, a.off, a.W,
                           static_cast<PlaneT<P>*>(a.out), tw, twr, scale, g0);
        return check_launch("corr_odd_kernel");
    });
}

template <int R, int M, int T, int G, int SH>
static int corr_odd_p(const CorrArgs& a, hipStream_t s)
{
    switch (a.prec) {
        case PREC_F16: return corr_odd_t<R, M, T, G, SH, PREC_F16>(a, s);
        case PREC_BF16: return corr_odd_t<R, M, T, G, SH, PREC_BF16>(a, s);
        case PREC_I16: return corr_odd_t<R, M, T, G, SH, PREC_I16>(a, s);
        default: return corr_odd_t<R, M, T, G, SH, PREC_F32>(a, s);
    }
}

template <int L, int V = 0>
static int corr_odd_l(const CorrArgs& a, hipStream_t s, int v)
{
    constexpr int R = L % 3 == 0 ? 3 : L % 5 == 0 ? 5 : 7;
    if constexpr (V < Shapes<L>::n) {
        constexpr CorrShape sh = Shapes<L>::v[V];
        if (v == V) return corr_odd_p<R, L / R, sh.T, sh.G, sh.S>(a, s);
        return corr_odd_l<L, V + 1>(a, s, v);
    } else {
        set_error("correlation variant %d out of range at length %d", v, L);
        return FDAS_ERR_INVALID;
    }
}

int launch_corr(int L, const CorrArgs& a, hipStream_t s, int variant)
{
    if (a.nsteps <= 0) return FDAS_OK;

    const int v = variant >= 0 ? variant : corr_variant(L);
    switch (L) {
        case 256: return corr_l<256>(a, s, v);
        case 512: return corr_l<512>(a, s, v);
        case 1024: return corr_l<1024>(a, s, v);
        case 2048: return corr_l<2048>(a, s, v);
        case 4096: return corr_l<4096>(a, s, v);
        case 8192: return corr_l<8192>(a, s, v);
        case 16384: return corr_l<16384>(a, s, v);
        case 768: return corr_odd_l<768>(a, s, v);
        case 1536: return corr_odd_l<1536>(a, s, v);
        case 3072: return corr_odd_l<3072>(a, s, v);
        case 6144: return corr_odd_l<6144>(a, s, v);
        case 12288: return corr_odd_l<12288>(a, s, v);
        case 1280: return corr_odd_l<1280>(a, s, v);
        case 2560: return corr_odd_l<2560>(a, s, v);
        case 5120: return corr_odd_l<5120>(a, s, v);
        case 10240: return corr_odd_l<10240>(a, s, v);
        case 7168: return corr_odd_l<7168>(a, s, v);
        default:
            set_error("fused correlation length %d unsupported", L);
            return FDAS_ERR_UNSUPPORTED;
    }
}

int launch_bank_pack(const cplx* d_bank, int Z, int L, int prec, void* d_out, int* d_row_exp,
                     hipStream_t s)
{
    if (prec == PREC_F16)
        hipLaunchKernelGGL(bank_pack_kernel<PREC_F16>, dim3(Z), dim3(256), 0, s, d_bank, L,
                           static_cast<unsigned*>(d_out), d_row_exp);
    else if (prec == PREC_BF16)
        hipLaunchKernelGGL(bank_pack_kernel<PREC_BF16>, dim3(Z), dim3(256), 0, s, d_bank, L,
                           static_cast<unsigned*>(d_out), d_row_exp);
    else if (prec == PREC_I16)
        hipLaunchKernelGGL(bank_pack_kernel<PREC_I16>, dim3(Z), dim3(256), 0, s, d_bank, L,
                           static_cast<unsigned*>(d_out), d_row_exp);
    else {
        set_error("bank packing needs a reduced-precision mode");
        return FDAS_ERR_INVALID;
    }
    return check_launch("bank_pack_kernel");
}

int launch_bank_split(const cplx* d_bank, int Z, int L, cplx* d_out, hipStream_t s)
{
    PhaseTable ph;
    FDAS_TRY(phase_table(L, &ph));
    hipLaunchKernelGGL(bank_split_kernel, dim3((L / 2 + 255) / 256, Z), dim3(256), 0, s, d_bank, L, d_out, ph);
    return check_launch("bank_split_kernel");
}

// ------------------------------------------------------------------ stage path
// source/bpd_gpu_kernels.cu:39-81: out[z][k] = X[k] * conj(K_z[k])
__global__ __launch_bounds__(256) void convolve_kernel(const cplx* __restrict__ x,
                                                       const cplx* __restrict__ bank, int L,
                                                       cplx* __restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (k >= L) return;
    const size_t idx = (size_t)z * L + k;
    out[idx] = cmulc(x[k], bank[idx]);
}

// source/bpd_gpu_kernels.cu:84-116: P[z][i] = |y[z*L + i]|^2 * scale, i < numrs
__global__ __launch_bounds__(256) void powers_kernel(const cplx* __restrict__ y, int numrs, int L,
                                                     float scale, float* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (i >= numrs) return;
    const cplx a = y[(size_t)z * L + i];
    out[(size_t)z * numrs + i] = (a.x * a.x + a.y * a.y) * scale;
}

// source/bpd_gpu_kernels.cu:119-170 with its FP64 index arithmetic.
__global__ __launch_bounds__(256) void add_powers_kernel(int Z0, int numrs0, int Lh, int zlo0,
                                                         int zlo_h, int rlo_h, double f,
                                                         const cplx* __restrict__ y, float scale,
                                                         float* __restrict__ P, int fullrlo)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= numrs0) return;
    const double rr = fullrlo + i * DR;
    const double subr = (int)(RDR * rr * f + 0.5) * DR;
    const unsigned short rind = (unsigned short)((subr - rlo_h) * RDR);
    const int zz = zlo0 + j * DZ;
    const int subz = (int)(zz * f);
    const int zind = (int)((subz - zlo_h) * RDZ);
    const cplx a = y[(size_t)zind * Lh + rind];
    P[(size_t)j * numrs0 + i] += (a.x * a.x + a.y * a.y) * scale;
}

// source/bpd_search.cu:150-181 for one harmonic: median of the first numdata
// powers, scale those amplitudes, spread into out[2i] (odd entries zero),
// i < nhalf.  Single workgroup; amps holds max(numdata, nhalf) values.
__global__ __launch_bounds__(1024) void normalize_spread_kernel(const cplx* __restrict__ amps,
                                                                int numdata, int nhalf,
                                                                cplx* __restrict__ out,
                                                                float* __restrict__ med_out)
{
    constexpr int T = 1024;
    __shared__ unsigned hist[256];
    __shared__ int sel[2];
    const int t = threadIdx.x;
    unsigned prefix = 0, mask = 0;
    int k = (numdata - 1) / 2;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int q = t; q < 256; q += T) hist[q] = 0;
        __syncthreads();
        for (int n = t; n < numdata; n += T) {
            const unsigned b = __float_as_uint(power_nc(amps[n]));
            if ((b & mask) == prefix) atomicAdd(&hist[(b >> shift) & 255u], 1u);
        }
        select_digit<T>(hist, sel, k, t);
        prefix |= (unsigned)sel[0] << shift;
        mask |= 255u << shift;
        k = sel[1];
    }
    const float med = __uint_as_float(prefix);
    const float c = quake_rsqrt_s((float)((double)med / kLn2));
    for (int n = t; n < nhalf; n += T) {
        cplx a = amps[n];
        if (n < numdata) a = scale_nc(a, c);
        out[2 * n] = a;
        out[2 * n + 1] = mk(0.f, 0.f);
    }
    if (t == 0 && med_out) *med_out = med;
}

int launch_convolve(const cplx* d_fft, const cplx* d_bank, int Z, int L, cplx* d_out,
                    hipStream_t s)
{
    hipLaunchKernelGGL(convolve_kernel, dim3((L + 255) / 256, Z), dim3(256), 0, s, d_fft, d_bank,
                       L, d_out);
    return check_launch("convolve_kernel");
}

int launch_powers(const cplx* d_ifft, int Z, int numrs, int L, float scale, float* d_out,
                  hipStream_t s)
{
    hipLaunchKernelGGL(powers_kernel, dim3((numrs + 255) / 256, Z), dim3(256), 0, s, d_ifft,
                       numrs, L, scale, d_out);
    return check_launch("powers_kernel");
}

int launch_add_powers(int Z0, int numrs0, int Lh, int zlo0, int zlo_h, int rlo_h, double f,
                      const cplx* d_ifft_h, float scale, float* d_powers, int fullrlo,
                      hipStream_t s)
{
    hipLaunchKernelGGL(add_powers_kernel, dim3((numrs0 + 255) / 256, Z0), dim3(256), 0, s, Z0,
                       numrs0, Lh, zlo0, zlo_h, rlo_h, f, d_ifft_h, scale, d_powers, fullrlo);
    return check_launch("add_powers_kernel");
}

int launch_normalize_spread(const cplx* d_amps, int numdata, int nhalf, cplx* d_out,
                            float* d_median_out, hipStream_t s)
{
    hipLaunchKernelGGL(normalize_spread_kernel, dim3(1), dim3(1024), 0, s, d_amps, numdata, nhalf,
                       d_out, d_median_out);
    return check_launch("normalize_spread_kernel");
}

}  // namespace fdas
