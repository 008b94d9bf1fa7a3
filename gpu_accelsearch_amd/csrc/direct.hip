// direct.hip -- precision mode FDAS_PREC_I16_MMA: the correlation of every
// (template, step) computed directly in the time domain, exactly, in integers on
// the matrix cores, instead of by FFT (mma_layout.hpp gives the matrix form).
//
//   mma_taps_kernel   the time-domain taps of each template row of a harmonic
//                     (FP64 Fresnel form, rounded to FP32) and the row's
//                     quantisation scale kQMax / max |component|.
//   mma_pack_kernel   the rows quantised to integers, split into two signed bytes
//                     and packed as the MFMA B fragments of each template tile.
//   seg_mma_kernel    per (step, harmonic): fetch, exact lower median, fast
//                     inverse square root normalisation (as seg_prep_kernel,
//                     source/bpd_search.cu:130-184), then quantisation of the
//                     segment to integers with a per-segment scale, split into
//                     two byte arrays of 16-byte chunks.  No FFT.
//   corr_mma_kernel   per (template tile, group of steps): 65536 HH + 256 (HL +
//                     LH) + LL by v_mfma_i32_32x32x32_i8 with 32-bit integer
//                     accumulation (exact for the quantised inputs), then the
//                     power |c|^2 scaled back in FP32 and stored as P16.
//
// The quantisation (per segment and per template row, about 15 significant bits
// at the top of each block) is the only rounding before the FP32 power; the power
// planes and the harmonic sum are those of the other reduced-precision modes.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <vector>

#include "fdas_internal.hpp"
#include "mma_layout.hpp"
#include "precision.hpp"
#include "segnorm.hpp"
#include "template_tap.hpp"

namespace fdas {

namespace {
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kMmaXcds = 8;        // workgroups reach the XCDs round robin

__device__ __forceinline__ int quantise(float x)
{
    return (int)fminf(fmaxf(rintf(x), (float)-kQMax), (float)kQMax);
}
// The two bytes (re, im) of one half of a quantised complex value.
__device__ __forceinline__ unsigned pair_bytes(int re, int im)
{
    return (unsigned)(re & 255) | ((unsigned)(im & 255) << 8);
}

// Taps of template row k (z = -zmax_h + 2k) into taps[k][0, stride) (zero past
// its 4 hw taps) and the row's scale and inverse scale.  One workgroup per row.
__global__ __launch_bounds__(256) void mma_taps_kernel(cplx* __restrict__ taps,
                                                       float* __restrict__ row_scale,
                                                       float* __restrict__ row_inv, int zmax_h,
                                                       int stride)
{
    __shared__ float red[4];
    const int k = blockIdx.x;
    const int z = -zmax_h + DZ * k;
    const int numkern = 2 * NUMBETWEEN * d_z_response_halfwidth((double)z);
    float m = 0.f;
    for (int i = threadIdx.x; i < stride; i += 256) {
        const cplx v = i < numkern ? template_tap(z, numkern, i) : mk(0.f, 0.f);
        taps[(size_t)k * stride + i] = v;
        m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float r = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        row_scale[k] = r > 0.f ? (float)kQMax / r : 1.f;
        row_inv[k] = r > 0.f ? r / (float)kQMax : 1.f;
    }
}

// B fragments of tile blockIdx.x, K step blockIdx.y, lane threadIdx.x: four
// 16-byte values (re output hi / lo bytes, im output hi / lo bytes).  With
// a conj(T) = (ar Tr + ai Ti) + i (ai Tr - ar Ti), the re column holds (Tr, Ti)
// and the im column (-Ti, Tr) for each segment pair (re, im).
__global__ __launch_bounds__(64) void mma_pack_kernel(v4i* __restrict__ frags,
                                                      const MmaTile* __restrict__ tiles,
                                                      const cplx* __restrict__ taps, int stride,
                                                      const float* __restrict__ row_scale,
                                                      int zmax_h, int bo)
{
    const MmaTile tl = tiles[blockIdx.x];
    const int k = blockIdx.y;
    if (k >= tl.nk) return;
    const int l = threadIdx.x, n = l & 31, g = l >> 5;
    const int row = mma_col_template(n) ? tl.row[1] : tl.row[0];
    unsigned w[4][4] = {};
    if (row >= 0) {
        const int hw = d_z_response_halfwidth((double)(-zmax_h + DZ * row));
        const float sc = row_scale[row];
#pragma unroll
        for (int sg = 0; sg < 8; ++sg) {
            const int ti = mma_b_tap(bo, hw, tl.jlo, n, g, k, sg);
            int xr = 0, xi = 0;
            if (ti >= 0) {
                const cplx v = taps[(size_t)row * stride + ti];
                xr = quantise(v.x * sc);
                xi = quantise(v.y * sc);
            }
            const int word = sg >> 1, sh = 16 * (sg & 1);
            w[0][word] |= pair_bytes(mma_split_hi(xr), mma_split_hi(xi)) << sh;
            w[1][word] |= pair_bytes(mma_split_lo(xr), mma_split_lo(xi)) << sh;
            w[2][word] |= pair_bytes(mma_split_hi(-xi), mma_split_hi(xr)) << sh;
            w[3][word] |= pair_bytes(mma_split_lo(-xi), mma_split_lo(xr)) << sh;
        }
    }
    v4i* o = frags + tl.boff + (size_t)k * 256;
#pragma unroll
    for (int f = 0; f < 4; ++f) o[f * 64 + l] = v4i{(int)w[f][0], (int)w[f][1], (int)w[f][2], (int)w[f][3]};
}

// Segment d = desc[blockIdx.x]: NH = L/2 values, value n at n = t + i T.  The
// first numdata are fetched, normalised exactly as seg_prep_kernel does, and
// quantised with the segment's scale; the rest of the 8 nch staged values are
// zero (no kept column reads past numdata).  Bytes of value n: hi array
// (re, im) at 2n, lo array the same, nch chunks further on.
template <int NH>
__global__ __launch_bounds__(NH < 256 ? NH : 256) void seg_mma_kernel(
    const cplx* __restrict__ spec_all, const SegDesc* __restrict__ desc, v4i* __restrict__ seg,
    long long out_base, float* __restrict__ seg_inv, int eslot_base)
{
    constexpr int T = NH < 256 ? NH : 256, E = NH / T;
    static_assert(T >= 64 && NH % T == 0, "segment shape");
    __shared__ unsigned hist[256];
    __shared__ int sel[2];
    __shared__ float red[T / 64];
    const SegDesc d = desc[blockIdx.x];
    const cplx* __restrict__ spec = spec_all + d.spec_off;
    const int t = threadIdx.x;
    cplx v[E];
    unsigned pb[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int n = t + i * T;
        const long b = (long)d.lobin + n;
        v[i] = (n < d.numdata && b >= 0 && b < d.nbins) ? spec[b] : mk(0.f, 0.f);
        pb[i] = __float_as_uint(power_nc(v[i]));
    }
    // exact lower median: rank (numdata-1)/2 among the first numdata powers
    unsigned prefix = 0, mask = 0;
    int k = (d.numdata - 1) / 2;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int q = t; q < 256; q += T) hist[q] = 0;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (t + i * T < d.numdata && (pb[i] & mask) == prefix)
                atomicAdd(&hist[(pb[i] >> shift) & 255u], 1u);
        select_digit<T>(hist, sel, k, t);
        prefix |= (unsigned)sel[0] << shift;
        mask |= 255u << shift;
        k = sel[1];
    }
    const float med = __uint_as_float(prefix);
    const float c = quake_rsqrt_s((float)((double)med / kLn2));
    float mx = 0.f;
#pragma unroll
    for (int i = 0; i < E; ++i) {
        if (t + i * T < d.numdata) v[i] = scale_nc(v[i], c);
        mx = fmaxf(mx, fmaxf(fabsf(v[i].x), fabsf(v[i].y)));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((t & 63) == 0) red[t >> 6] = mx;
    __syncthreads();
    float r = red[0];
#pragma unroll
    for (int q = 1; q < T / 64; ++q) r = fmaxf(r, red[q]);
    const float sa = r > 0.f ? (float)kQMax / r : 1.f;
    if (t == 0) seg_inv[d.eslot - eslot_base] = r > 0.f ? r / (float)kQMax : 1.f;
    v4i* base = seg + (d.out - out_base);
    unsigned short* hi = reinterpret_cast<unsigned short*>(base);
    unsigned short* lo = reinterpret_cast<unsigned short*>(base + d.nch);
    const int nv = 8 * d.nch;
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int n = t + i * T;
        if (n < nv) {
            const int xr = quantise(v[i].x * sa), xi = quantise(v[i].y * sa);
            hi[n] = (unsigned short)pair_bytes(mma_split_hi(xr), mma_split_hi(xi));
            lo[n] = (unsigned short)pair_bytes(mma_split_lo(xr), mma_split_lo(xi));
        }
    }
    for (int n = NH + t; n < nv; n += T) {
        hi[n] = 0;
        lo[n] = 0;
    }
}

// Workgroup (template tile, group of G steps): the tile's B fragments are staged
// in LDS once; then each wavefront independently takes (step, M tile) units
// round robin and runs the tile's K steps for them, reading its A fragments
// (aligned 16-byte chunks, contiguous over the lanes and mostly re-read by the
// neighbouring lanes and K steps, so served by the L1) straight from the segment
// buffer.  No barrier after the staging.  8 MFMAs per unit and K step: HH, HL,
// LH, LL for the re and im columns, the two middle ones into one accumulator.
// Tiles fastest over a count padded to a multiple of 8, so XCD k always gets the
// same tiles and keeps their fragments in its L2 across step groups.  Padding ids
// return before the barrier.
template <int MT, int NT>
__global__ __launch_bounds__(NT, MT == 1 ? (NT > 512 ? 3 : 2) : 1) void corr_mma_kernel(
    const v4i* __restrict__ frags, const MmaTile* __restrict__ tiles, int ntiles,
    const float* __restrict__ row_inv, const v4i* __restrict__ seg, long long seg_stride,
    long long seg_off, int nch, const float* __restrict__ seg_inv, int H, int h, int nsteps, int W,
    int Z, int nm, int G, unsigned short* __restrict__ out, int g0)
{
    extern __shared__ v4i ldsB[];
    const int tp = (ntiles + kMmaXcds - 1) / kMmaXcds * kMmaXcds;
    const int tile = (int)(blockIdx.x % (unsigned)tp);
    if (tile >= ntiles) return;
    const int s0 = (g0 + (int)(blockIdx.x / (unsigned)tp)) * G;
    const int s1 = min(s0 + G, nsteps);
    const MmaTile tl = tiles[tile];
    const int tid = threadIdx.x, l = tid & 63, wv = tid >> 6, x = l & 31, g = l >> 5;
    for (int q = tid; q < tl.nk * 256; q += NT) ldsB[q] = frags[tl.boff + q];
    const int row0 = tl.row[0], row1 = tl.row[1];
    const float rinv0 = row_inv[row0], rinv1 = row1 >= 0 ? row_inv[row1] : 0.f;
    __syncthreads();
    const int units = (s1 - s0) * nm;
    constexpr int UW = (NT / 64) * MT;   // units per round of the workgroup
    // Units are pipelined: the next unit's first two K steps of fragments (and its
    // segment scale) are loaded before the current unit's stores, so no load waits
    // behind those stores (loads and stores share one in-order counter).
    struct Frags {
        v4i bh, bl, ch, cl, ah[MT], al[MT];
    };
    const v4i* ap[MT];
    float sa[MT];
    auto locate = [&](int u0) {
#pragma unroll
        for (int j = 0; j < MT; ++j) {
            const int u = min(u0 + j, units - 1);
            const int st = u / nm, m = u - st * nm;
            ap[j] = seg + (long long)(s0 + st) * seg_stride + seg_off + mma_a_chunk(m, x, g, tl.jlo, 0);
            sa[j] = seg_inv[(s0 + st) * H + h];
        }
    };
    auto load = [&](Frags& f, int k) {
        const v4i* bp = ldsB + k * 256 + l;
        f.bh = bp[0];
        f.bl = bp[64];
        f.ch = bp[128];
        f.cl = bp[192];
#pragma unroll
        for (int j = 0; j < MT; ++j) {
            f.ah[j] = ap[j][2 * k];
            f.al[j] = ap[j][nch + 2 * k];
        }
    };
    Frags f0, f1;
    if (wv * MT < units) {
        locate(wv * MT);
        load(f0, 0);
        load(f1, min(1, tl.nk - 1));
    }
    for (int u0 = wv * MT; u0 < units; u0 += UW) {
        v16i acc[MT][6];
        // K steps: the next step's fragments load while the current step's MFMAs
        // run (two register sets; loads past the last step are clamped to it).
        // The first step's MFMAs start the sums from zero.
        auto mma = [&](const Frags& f) {
#pragma unroll
            for (int j = 0; j < MT; ++j) {
                acc[j][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.bh, f.ah[j], acc[j][0], 0, 0, 0);
                acc[j][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.bl, f.ah[j], acc[j][1], 0, 0, 0);
                acc[j][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.bh, f.al[j], acc[j][1], 0, 0, 0);
                acc[j][2] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.bl, f.al[j], acc[j][2], 0, 0, 0);
                acc[j][3] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.ch, f.ah[j], acc[j][3], 0, 0, 0);
                acc[j][4] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.cl, f.ah[j], acc[j][4], 0, 0, 0);
                acc[j][4] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.ch, f.al[j], acc[j][4], 0, 0, 0);
                acc[j][5] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.cl, f.al[j], acc[j][5], 0, 0, 0);
            }
        };
        auto mma0 = [&](const Frags& f) {
            const v16i z{};
#pragma unroll
            for (int j = 0; j < MT; ++j) {
                acc[j][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.bh, f.ah[j], z, 0, 0, 0);
                acc[j][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.bl, f.ah[j], z, 0, 0, 0);
                acc[j][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.bh, f.al[j], acc[j][1], 0, 0, 0);
                acc[j][2] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.bl, f.al[j], z, 0, 0, 0);
                acc[j][3] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.ch, f.ah[j], z, 0, 0, 0);
                acc[j][4] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.cl, f.ah[j], z, 0, 0, 0);
                acc[j][4] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.ch, f.al[j], acc[j][4], 0, 0, 0);
                acc[j][5] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.cl, f.al[j], z, 0, 0, 0);
            }
        };
        mma0(f0);
        int k = 1;
        for (; k + 1 < tl.nk; k += 2) {
            load(f0, k + 1);
            mma(f1);
            load(f1, min(k + 2, tl.nk - 1));
            mma(f0);
        }
        if (k < tl.nk) mma(f1);
        // this unit's outputs, then the next unit's first fragments
        float sa_cur[MT];
        int mcur[MT], scur[MT];
#pragma unroll
        for (int j = 0; j < MT; ++j) {
            sa_cur[j] = sa[j];
            const int u = min(u0 + j, units - 1);
            scur[j] = u / nm;
            mcur[j] = u - scur[j] * nm;
        }
        if (u0 + UW < units) {
            locate(u0 + UW);
            load(f0, 0);
            load(f1, min(1, tl.nk - 1));
        }
        // power = |re + i im|^2 / (segment scale * row scale)^2.  Registers 4q ..
        // 4q + 3 hold plane columns 16 P + 8 (q % 2) + 4 g + 0..3 of template q / 2
        // (P = 32 m + x): stored as two pairs; the column test is needed only in a
        // row's last M tile.
#pragma unroll
        for (int j = 0; j < MT; ++j) {
            if (u0 + j >= units) continue;
            const int m = mcur[j], s = s0 + scur[j];
            const bool full = 512 * (m + 1) <= W;
            const int cb = 16 * (32 * m + x) + 4 * g;
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int row = t ? row1 : row0;
                if (row < 0) continue;
                const float sc = sa_cur[j] * (t ? rinv1 : rinv0), sc2 = sc * sc;
                unsigned* o = reinterpret_cast<unsigned*>(out + ((size_t)s * Z + row) * W);
                auto put = [&](int r) {
                    const float re = fmaf((float)acc[j][0][r], 65536.f,
                                          fmaf((float)acc[j][1][r], 256.f, (float)acc[j][2][r]));
                    const float im = fmaf((float)acc[j][3][r], 65536.f,
                                          fmaf((float)acc[j][4][r], 256.f, (float)acc[j][5][r]));
                    return (unsigned)PStore<unsigned short>::put(fmaf(re, re, im * im) * sc2);
                };
#pragma unroll
                for (int q = 2 * t; q < 2 * t + 2; ++q)
#pragma unroll
                    for (int pr = 0; pr < 2; ++pr) {
                        const int r = 4 * q + 2 * pr;
                        const int col = cb + 8 * (q & 1) + 2 * pr;   // even: W is even
                        if (full || col < W) o[col >> 1] = put(r) | (put(r + 1) << 16);
                    }
            }
        }
    }
}

template <int NH>
int seg_mma_p(const SegPrepArgs& a, hipStream_t s)
{
    constexpr int T = NH < 256 ? NH : 256;
    hipLaunchKernelGGL(seg_mma_kernel<NH>, dim3((unsigned)a.nseg), dim3(T), 0, s, a.spec, a.desc,
                       static_cast<v4i*>(a.seg), a.out_base, reinterpret_cast<float*>(a.seg_exp),
                       a.eslot_base);
    return check_launch("seg_mma_kernel");
}
}  // namespace

// ------------------------------------------------------------------ host side
void free_mma_bank(MmaBank* b)
{
    (void)hipFree(b->frags);
    (void)hipFree(b->tiles);
    (void)hipFree(b->row_inv);
    *b = MmaBank{};
}

// Tiles pair the rows by decreasing |z| (z and -z, then the next |z|), so the two
// templates of a tile have equal or neighbouring half-widths and the tile's K
// range is that of the larger.
int build_mma_bank(int zmax_h, int Z, int bo, hipStream_t s, MmaBank* out)
{
    *out = MmaBank{};
    std::vector<int> rows;
    for (int az = zmax_h; az >= 0; az -= DZ) {
        rows.push_back((zmax_h + az) / DZ);
        if (az > 0) rows.push_back((zmax_h - az) / DZ);
    }
    if ((int)rows.size() != Z) {
        set_error("direct bank: %d rows for zmax %d, expected %d", (int)rows.size(), zmax_h, Z);
        return FDAS_ERR_INVALID;
    }
    std::vector<MmaTile> tiles;
    long long boff = 0;
    int nk_max = 0;
    for (size_t i = 0; i < rows.size(); i += 2) {
        MmaTile t{};
        t.row[0] = rows[i];
        t.row[1] = i + 1 < rows.size() ? rows[i + 1] : -1;
        int hw = 0;
        for (int q = 0; q < 2; ++q)
            if (t.row[q] >= 0) hw = std::max(hw, h_z_response_halfwidth((double)(-zmax_h + DZ * t.row[q])));
        mma_tile_range(bo, hw, &t.jlo, &t.nk);
        t.boff = boff;
        boff += (long long)t.nk * 256;
        nk_max = std::max(nk_max, t.nk);
        tiles.push_back(t);
    }
    if (nk_max > kMmaMaxK) {
        set_error("direct correlation: templates up to half-width %d need %d K steps (at most %d, "
                  "zmax up to 300)", bo, nk_max, kMmaMaxK);
        return FDAS_ERR_UNSUPPORTED;
    }
    const int stride = 2 * NUMBETWEEN * bo;
    cplx* taps = nullptr;
    float* scale = nullptr;
    int rc = FDAS_OK;
    auto fail = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == FDAS_OK) rc = hip_fail(e, what, __FILE__, __LINE__);
    };
    fail(hipMalloc(&out->frags, (size_t)boff * 16), "hipMalloc frags");
    fail(hipMalloc(&out->tiles, tiles.size() * sizeof(MmaTile)), "hipMalloc tiles");
    fail(hipMalloc(&out->row_inv, (size_t)Z * sizeof(float)), "hipMalloc row_inv");
    fail(hipMalloc(&taps, (size_t)Z * stride * sizeof(cplx)), "hipMalloc taps");
    fail(hipMalloc(&scale, (size_t)Z * sizeof(float)), "hipMalloc scale");
    if (rc == FDAS_OK)
        fail(hipMemcpyAsync(out->tiles, tiles.data(), tiles.size() * sizeof(MmaTile),
                            hipMemcpyHostToDevice, s), "hipMemcpyAsync tiles");
    if (rc == FDAS_OK) {
        hipLaunchKernelGGL(mma_taps_kernel, dim3((unsigned)Z), dim3(256), 0, s, taps, scale,
                           out->row_inv, zmax_h, stride);
        rc = check_launch("mma_taps_kernel");
    }
    if (rc == FDAS_OK) {
        hipLaunchKernelGGL(mma_pack_kernel, dim3((unsigned)tiles.size(), (unsigned)nk_max), dim3(64), 0,
                           s, static_cast<v4i*>(out->frags), out->tiles, taps, stride, scale, zmax_h, bo);
        rc = check_launch("mma_pack_kernel");
    }
    fail(hipStreamSynchronize(s), "hipStreamSynchronize");   // before the host tables go
    (void)hipFree(taps);
    (void)hipFree(scale);
    out->ntiles = (int)tiles.size();
    out->nk_max = nk_max;
    out->bytes = (size_t)boff * 16 + tiles.size() * sizeof(MmaTile) + (size_t)Z * sizeof(float);
    if (rc != FDAS_OK) free_mma_bank(out);
    return rc;
}

int launch_segment_prep_mma(int L, const SegPrepArgs& a, hipStream_t s)
{
    if (a.nseg == 0) return FDAS_OK;
    switch (L) {
        case 256: return seg_mma_p<128>(a, s);
        case 512: return seg_mma_p<256>(a, s);
        case 1024: return seg_mma_p<512>(a, s);
        case 2048: return seg_mma_p<1024>(a, s);
        case 4096: return seg_mma_p<2048>(a, s);
        case 8192: return seg_mma_p<4096>(a, s);
        case 16384: return seg_mma_p<8192>(a, s);
        default:
            set_error("direct segment length %d unsupported", L / 2);
            return FDAS_ERR_UNSUPPORTED;
    }
}

// Launch shapes of the direct correlation: M tiles per wavefront in flight
// (1: 162 VGPRs, two or three wavefronts per SIMD; 2: half the B fragment reads
// per MFMA but one wavefront per SIMD), steps per workgroup (the template tile's
// fragments are staged once per workgroup) and threads per workgroup (more
// wavefronts share one staged tile).  Production trial (2^23, zmax 200, fftlen
// 8192), fundamental / sub-harmonic ms: {1, 16, 256} 4.1 / 8.2, {1, 16, 512}
// 4.0 / 8.0, {1, 16, 768} 4.2 / 8.2, {1, 8, 256} 4.1 / 8.1, {2, 16, 256}
// 4.3 / 9.1: the shape barely matters (the kernel is bound by its operand
// traffic and epilogue arithmetic beside the MFMAs, not by occupancy).
namespace {
struct MmaShape {
    int MT, G, T;
};
constexpr MmaShape kMmaShapes[] = {{1, 16, 512}, {1, 16, 256}, {1, 8, 256}, {1, 16, 768},
                                   {2, 16, 256}};
constexpr int kMmaVariants = (int)(sizeof(kMmaShapes) / sizeof(kMmaShapes[0]));
std::atomic<int> g_mma_variant{0};

template <int MT, int NT>
int corr_mma_t(const MmaCorrArgs& a, int G, hipStream_t s)
{
    const MmaBank& b = *a.bank;
    const size_t lds = (size_t)b.nk_max * 256 * 16;
    static std::once_flag attr_once;
    static hipError_t attr_err = hipSuccess;
    std::call_once(attr_once, [] {
        attr_err = hipFuncSetAttribute(reinterpret_cast<const void*>(&corr_mma_kernel<MT, NT>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 << 10);
    });
    FDAS_HIP(attr_err);
    const long groups = (a.nsteps + G - 1) / G;
    const long tp = (b.ntiles + kMmaXcds - 1) / kMmaXcds * kMmaXcds;
    const long per = std::max(1L, (2147483647L / NT) / tp);
    for (long g0 = 0; g0 < groups; g0 += per) {
        const long n = std::min(per, groups - g0);
        hipLaunchKernelGGL((corr_mma_kernel<MT, NT>), dim3((unsigned)(tp * n)), dim3(NT), lds, s,
                           static_cast<const v4i*>(b.frags), b.tiles, b.ntiles, b.row_inv,
                           static_cast<const v4i*>(a.seg), a.seg_stride, a.seg_off, a.nch, a.seg_inv,
                           a.H, a.h, a.nsteps, a.W, a.Z, a.nm, G, static_cast<unsigned short*>(a.out),
                           (int)g0);
        FDAS_TRY(check_launch("corr_mma_kernel"));
    }
    return FDAS_OK;
}
}  // namespace

int mma_variant_count() { return kMmaVariants; }
int mma_variant() { return g_mma_variant.load(); }
int set_mma_variant(int v)
{
    if (v < 0 || v >= kMmaVariants) {
        set_error("no direct-correlation variant %d", v);
        return FDAS_ERR_INVALID;
    }
    g_mma_variant.store(v);
    return FDAS_OK;
}
int mma_variant_units(int v) { return v >= 0 && v < kMmaVariants ? kMmaShapes[v].MT : 0; }
int mma_variant_steps(int v) { return v >= 0 && v < kMmaVariants ? kMmaShapes[v].G : 0; }
int mma_variant_threads(int v) { return v >= 0 && v < kMmaVariants ? kMmaShapes[v].T : 0; }

int launch_corr_mma(const MmaCorrArgs& a, hipStream_t s)
{
    if (a.nsteps <= 0) return FDAS_OK;
    const int G = mma_variant_steps(a.variant);
    switch (mma_variant_units(a.variant) * 10000 + mma_variant_threads(a.variant)) {
        case 10256: return corr_mma_t<1, 256>(a, G, s);
        case 10512: return corr_mma_t<1, 512>(a, G, s);
        case 10768: return corr_mma_t<1, 768>(a, G, s);
        case 20256: return corr_mma_t<2, 256>(a, G, s);
        default:
            set_error("no direct-correlation variant %d", a.variant);
            return FDAS_ERR_INVALID;
    }
}

}  // namespace fdas
