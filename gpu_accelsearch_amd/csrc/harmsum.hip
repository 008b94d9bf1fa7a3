// harmsum.hip -- harmonic summing over tiles of the fundamental f-fdot plane.
//
// Restates perform_harmonic_summing / addPowers (source/bpd_search.cu:310-360,
// source/bpd_gpu_kernels.cu:119-170): for every fundamental cell (row j, column i)
// P[j][i] += P_h[zind_h(j)][rind_h(i)] for h = 1..H-1 in that order, and the stage
// thresholds after harmonics 1, 3, 7, 15 that the reference leaves as a stub
// (source/bpd_search.cu:350-356).
//
// One workgroup owns a tile of HS_RJ rows x HS_CB columns of one step (4 x 896 or
// 4 x 768 by default).  The cells of sub-harmonic h that the tile reads form a
// rectangle of about HS_CB f + 1 columns of the sub-harmonic rows its fundamental
// rows map to.  The rectangles of as many harmonics as fit are staged in LDS
// column by column: a column's 16-byte slot holds, in fundamental-row order, the
// value each of the tile's rows reads there (the row map is uniform over the
// workgroup, so it is applied once, in the staging loads), and one 16-byte LDS
// read per column then gives a thread the addends of all its rows with no per-cell
// address arithmetic.  Each fundamental cell thus costs about sum_h f_h^2 distinct
// sub-harmonic reads (2.19 at 8 harmonics) instead of H-1 scattered ones.
#include <atomic>
#include <cstdlib>

#include "fdas_internal.hpp"
#include "harmidx.hpp"
#include "precision.hpp"

namespace fdas {

// Tile shapes: HS_T threads, HS_RJ (2..4) rows x HS_CB columns, HS_LDS staged
// floats.  A rectangle has at most HS_CB + 2 columns of 4 floats (one per tile
// row, the unused ones zero); HS_MAXK column slots per thread cover it.
#define HS_TILE_CONSTANTS                                                      \
    constexpr int HS_CPT = HS_CB / HS_T;                                       \
    constexpr int HS_MAXK = (HS_CB + 2 + HS_T - 1) / HS_T;                     \
    static_assert(HS_RJ >= 2 && HS_RJ <= 4, "rows of a tile");                 \
    static_assert(4 * (HS_CB + 2) <= HS_LDS, "one rectangle must fit");        \
    static_assert(HS_CB % HS_T == 0 && HS_RJ * HS_CPT <= 32, "a hit bit per cell");

__device__ __forceinline__ float comp4(const float4& v, int i)
{
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}

// grid: (ceil(U/HS_CB), ceil(Z0/HS_RJ), steps); fund[s][Z0][U] holds the
// fundamental powers (plane element PE: FP32 or P16) and, when store is set, the
// summed FP32 powers go to out[s][Z0][U] (every cell is summed and thresholded
// either way).  With FP32 planes out may be fund itself: each cell is read and
// written by the same thread.
template <int HS_T, int HS_RJ, int HS_CB, int HS_LDS, typename PE>
__global__ __launch_bounds__(HS_T) void harmsum_kernel(const PE* fund, float* out,
                                                       const PE* __restrict__ sub,
                                                       HarmSumArgs hs, int step0, int Z0,
                                                       const int* __restrict__ fullrlo,
                                                       const int* __restrict__ rlo_h,
                                                       fdas_candidate* __restrict__ cands,
                                                       int* __restrict__ ncand, int store)
{
    HS_TILE_CONSTANTS
    __shared__ __align__(16) float tile[HS_LDS];
    const int t = threadIdx.x;
    const int c0 = blockIdx.x * HS_CB, j0 = blockIdx.y * HS_RJ, s = blockIdx.z;
    const int gstep = step0 + s;
    const int U = hs.U;
    const int nrow = min(HS_RJ, Z0 - j0);
    const int ncol = min(HS_CB, U - c0);
    const unsigned q0 = 2u * (unsigned)fullrlo[gstep];
    const int* rl = rlo_h + (size_t)gstep * MAX_HARM;
    const PE* base_row = fund + ((size_t)s * Z0 + j0) * U;

    // fundamental powers, loaded from clamped (always valid) cells without a
    // branch so that every load is in flight at once; they are decoded and
    // tested once the first round's loads are in flight too, so both arrive
    // behind one memory wait.  Cells outside the tile start at -inf, so no stage
    // sum of theirs passes a cut, and they are never stored.
    float acc[HS_RJ][HS_CPT];
    PE raw[HS_RJ][HS_CPT];
    unsigned hit = 0;
    bool first = true;
#pragma unroll
    for (int r = 0; r < HS_RJ; ++r)
#pragma unroll
        for (int k = 0; k < HS_CPT; ++k) {
            const int col = min(c0 + t + k * HS_T, c0 + ncol - 1);
            raw[r][k] = base_row[(size_t)min(r, nrow - 1) * U + col];
        }
    // Stage test: one max over the thread's cells; only a thread whose max passes
    // the cut (rare) marks its cells one by one.
    auto mark_cells = [&](float pc) {
        float m = acc[0][0];
#pragma unroll
        for (int r = 0; r < HS_RJ; ++r)
#pragma unroll
            for (int k = 0; k < HS_CPT; ++k) m = fmaxf(m, acc[r][k]);
        if (m > pc) {
#pragma unroll
            for (int r = 0; r < HS_RJ; ++r)
#pragma unroll
                for (int k = 0; k < HS_CPT; ++k)
                    if (acc[r][k] > pc) hit |= 1u << (r * HS_CPT + k);
        }
    };
    auto mark_fundamental = [&]() {
#pragma unroll
        for (int r = 0; r < HS_RJ; ++r)
#pragma unroll
            for (int k = 0; k < HS_CPT; ++k) {
                const bool ok = r < nrow && t + k * HS_T < ncol;
                acc[r][k] = ok ? PStore<PE>::get(raw[r][k]) : -INFINITY;
            }
        mark_cells(hs.powcut[0]);
    };
    if (hs.H <= 1) mark_fundamental();

    // Geometry of every harmonic's rectangle, computed once per tile by one
    // thread per harmonic and read back by all waves as uniform values:
    // [0] first staged column, [1] staged columns, [2] c, [3] log2 m, [4] 2 rlo_h,
    // [5] LDS offset of the rectangle in its round, [6 + r] sub-harmonic row read
    // by fundamental row r relative to row 0's; gbase[h]: element offset of the
    // rectangle's first value (row 0's sub-harmonic row, column [0]).
    constexpr int GEO = 6 + HS_RJ;
    __shared__ int geo[MAX_HARM][GEO];
    __shared__ long long gbase[MAX_HARM];
    __shared__ int rounds[MAX_HARM + 1];
    __shared__ int nrounds;
    if (t > 0 && t < hs.H) {
        const int c = hs.c[t], lm = hs.lm[t], r2 = 2 * rl[t];
        const int zi0 = hs_zind(hs.zlo0 + DZ * j0, c, lm, hs.zlo[t]);
        const int ri0 = hs_rind(q0 + c0, c, lm, r2);
        geo[t][0] = ri0;
        geo[t][1] = hs_rind(q0 + c0 + ncol - 1, c, lm, r2) + 1 - ri0;
        geo[t][2] = c;
        geo[t][3] = lm;
        geo[t][4] = r2;
#pragma unroll
        for (int r = 0; r < HS_RJ; ++r)
            geo[t][6 + r] = hs_zind(hs.zlo0 + DZ * (j0 + min(r, nrow - 1)), c, lm, hs.zlo[t]) - zi0;
        gbase[t] = hs.sub_off[t] + (long long)s * hs.sub_stride[t] + (long long)zi0 * hs.W[t] + ri0;
    }
    __syncthreads();
    if (t == 0) {
        // harmonics [rounds[i], rounds[i+1]) whose rectangles fit in LDS together
        int n = 0, used = 0;
        rounds[0] = 1;
        for (int h = 1; h < hs.H; ++h) {
            const int sz = 4 * geo[h][1];
            if (used + sz > HS_LDS) {
                rounds[++n] = h;
                used = 0;
            }
            geo[h][5] = used;
            used += sz;
        }
        rounds[++n] = hs.H;
        nrounds = hs.H > 1 ? n : 0;
    }
    __syncthreads();
    auto u = [](int v) { return __builtin_amdgcn_readfirstlane(v); };

    int stage = 0;
    const int nrnd = u(nrounds);
    for (int rd = 0; rd < nrnd; ++rd) {
        const int h = u(rounds[rd]), h2 = u(rounds[rd + 1]);
        if (rd > 0) __syncthreads();   // earlier readers of the tile are done
        for (int g = h; g < h2; ++g) {
            const int nca = u(geo[g][1]), base = u(geo[g][5]);
            // column slots t + k HS_T: the value each tile row reads in that column
            // (loaded once per distinct sub-harmonic row, clamped to the last
            // column so every address is inside the rectangle), all loads before
            // any store, then one 16-byte LDS store per column
            // uniform row pointers (so the loads take a scalar base and a 32-bit
            // per-lane offset)
            const long long gb = gbase[g];
            const long long gbu = ((long long)u((int)(gb >> 32)) << 32) | (unsigned)u((int)gb);
            const int W = hs.W[g];
            const PE* rowp[HS_RJ];
#pragma unroll
            for (int r = 0; r < HS_RJ; ++r) rowp[r] = sub + gbu + (long long)(u(geo[g][6 + r]) * W);
            // every row loaded (a row that reads the same sub-harmonic row as the
            // row before it reads the same cache lines), all loads before any store
            PE e[HS_MAXK][HS_RJ];
#pragma unroll
            for (int k = 0; k < HS_MAXK; ++k) {
                if (k * HS_T < nca) {
                    const unsigned off = (unsigned)min(t + k * HS_T, nca - 1) * (unsigned)sizeof(PE);
#pragma unroll
                    for (int r = 0; r < HS_RJ; ++r)
                        e[k][r] = *reinterpret_cast<const PE*>(reinterpret_cast<const char*>(rowp[r]) + off);
                }
            }
            if (first) {
                mark_fundamental();
                first = false;
            }
            float4* dst = reinterpret_cast<float4*>(tile + base);
#pragma unroll
            for (int k = 0; k < HS_MAXK; ++k) {
                const int cc = t + k * HS_T;
                if (k * HS_T < nca && cc < nca) {
                    float f[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int r = 0; r < HS_RJ; ++r) f[r] = PStore<PE>::get(e[k][r]);
                    dst[cc] = make_float4(f[0], f[1], f[2], f[3]);
                }
            }
        }
        __syncthreads();
        for (int g = h; g < h2; ++g) {
            const int ri0 = u(geo[g][0]), base = u(geo[g][5]);
            const int c = u(geo[g][2]), lm = u(geo[g][3]), r2 = u(geo[g][4]);
            // staged column of each of the thread's columns: hs_rind with the
            // product q c stepped by HS_T c from column to column, clamped to the
            // tile's last column (cells beyond it are -inf and never used)
            const int ri_last = u(hs_rind(q0 + c0 + ncol - 1, c, lm, r2) - ri0);
            unsigned qc = (q0 + (unsigned)(c0 + t)) * (unsigned)c + (1u << (lm - 1));
            const unsigned qstep = (unsigned)(HS_T * c);
            const float4* col = reinterpret_cast<const float4*>(tile + base);
#pragma unroll
            for (int k = 0; k < HS_CPT; ++k) {
                const float4 cv = col[min((int)(qc >> lm) - (r2 + ri0), ri_last)];
                qc += qstep;
#pragma unroll
                for (int r = 0; r < HS_RJ; ++r) acc[r][k] += comp4(cv, r);
            }
            if (((g + 1) & g) == 0) mark_cells(hs.powcut[++stage]);
        }
    }

    // rare path: stage sums of the marked cells, re-accumulated in the same order
    // (before the store, which may overwrite the fundamental plane in place)
    if (hit) {
        for (int b = 0; b < HS_RJ * HS_CPT; ++b) {
            if (!((hit >> b) & 1u)) continue;
            const int r = b / HS_CPT, k = b % HS_CPT;
            const int j = j0 + r, col = c0 + t + k * HS_T;
            const int zz = hs.zlo0 + DZ * j;
            // the fundamental power, read back (the cell is stored only below)
            float a2 = PStore<PE>::get(base_row[(size_t)r * U + col]);
            int st = 0;
            if (a2 > hs.powcut[0]) hs_emit(cands, ncand, hs.cand_cap, gstep, j, col, 0, a2);
            for (int h = 1; h < hs.H; ++h) {
                const int c = hs.c[h], lm = hs.lm[h];
                const int rind = hs_rind(q0 + col, c, lm, 2 * rl[h]);
                const int zind = hs_zind(zz, c, lm, hs.zlo[h]);
                a2 += PStore<PE>::get(sub[hs.sub_off[h] + s * hs.sub_stride[h] + (long long)zind * hs.W[h] + rind]);
                if (((h + 1) & h) == 0) {
                    ++st;
                    if (a2 > hs.powcut[st]) hs_emit(cands, ncand, hs.cand_cap, gstep, j, col, st, a2);
                }
            }
        }
    }

    if (store) {
        float* out_row = out + ((size_t)s * Z0 + j0) * U;
#pragma unroll
        for (int r = 0; r < HS_RJ; ++r)
#pragma unroll
            for (int k = 0; k < HS_CPT; ++k) {
                const int col = c0 + t + k * HS_T;
                if (r < nrow && col < c0 + ncol) out_row[(size_t)r * U + col] = acc[r][k];
            }
    }
}

template <int HS_T, int HS_RJ, int HS_CB, int HS_LDS, typename PE>
static int harmsum_pe(const void* d_fund, float* d_out, const void* d_sub, const HarmSumArgs& hs,
                      int step0, int nsteps, int Z0, const int* d_fullrlo, const int* d_rlo_h,
                      fdas_candidate* d_cands, int* d_ncand, int store, hipStream_t s)
{
    for (int h = 1; h < hs.H; ++h) {
        if ((hs.W[h] & 3) || (hs.sub_off[h] & 3) || (hs.sub_stride[h] & 3)) {
            set_error("harmonic sum: sub-plane %d rows are not 4-float aligned", h);
            return FDAS_ERR_INVALID;
        }
    }
    if (((uintptr_t)d_sub) & 15) {
        set_error("harmonic sum: sub-plane buffer is not 16-byte aligned");
        return FDAS_ERR_INVALID;
    }
    dim3 grid((unsigned)((hs.U + HS_CB - 1) / HS_CB), (unsigned)((Z0 + HS_RJ - 1) / HS_RJ),
              (unsigned)nsteps);
    hipLaunchKernelGGL((harmsum_kernel<HS_T, HS_RJ, HS_CB, HS_LDS, PE>), grid, dim3(HS_T), 0, s,
                       static_cast<const PE*>(d_fund), d_out, static_cast<const PE*>(d_sub), hs, step0,
                       Z0, d_fullrlo, d_rlo_h, d_cands, d_ncand, store);
    return check_launch("harmsum_kernel");
}

template <int HS_T, int HS_RJ, int HS_CB, int HS_LDS>
static int harmsum_t(const void* d_fund, float* d_out, const void* d_sub, int prec,
                     const HarmSumArgs& hs, int step0, int nsteps, int Z0, const int* d_fullrlo,
                     const int* d_rlo_h, fdas_candidate* d_cands, int* d_ncand, int store,
                     hipStream_t s)
{
    if (prec == PREC_F32)
        return harmsum_pe<HS_T, HS_RJ, HS_CB, HS_LDS, float>(d_fund, d_out, d_sub, hs, step0, nsteps,
                                                             Z0, d_fullrlo, d_rlo_h, d_cands, d_ncand,
                                                             store, s);
    return harmsum_pe<HS_T, HS_RJ, HS_CB, HS_LDS, unsigned short>(d_fund, d_out, d_sub, hs, step0,
                                                                  nsteps, Z0, d_fullrlo, d_rlo_h,
                                                                  d_cands, d_ncand, store, s);
}

// Tile shapes (threads, rows, columns, staged floats).  The default, variant 0,
// is 128 threads x 4 rows with 896 or 768 columns, whichever leaves fewer idle
// columns in the last tile of a U-wide row (U = 7734: 896, 3638: 768).  Small
// LDS tiles keep 4 waves per SIMD resident, and more rows per tile reuse more
// staged sub-harmonic rows; on the production trial it measured 4.22 ms against
// 4.96 for 2 x 1024 tiles (7.8 against 9.4 ms at 16 harmonics, and 4.63 against
// 5.39 ms at fftlen 4096).  The autotuner times every shape.
namespace {
struct HsShape {
    int T, RJ, CB, LDS;   // CB 0: 896 or 768 chosen per launch from U
};
constexpr HsShape kHsShapes[] = {{128, 4, 0, 4608}, {128, 4, 768, 4096}, {128, 4, 896, 4608},
                                 {128, 3, 1024, 4160}, {128, 2, 1024, 4160}, {128, 3, 896, 3840},
                                 {256, 2, 2048, 8256}, {128, 4, 640, 3072}};
constexpr int kHsCount = (int)(sizeof(kHsShapes) / sizeof(kHsShapes[0]));
std::atomic<int> g_hs_variant{0};

// Idle columns in the last tile of a row of U columns with tiles of cb.
int tail_waste(int U, int cb) { return (U + cb - 1) / cb * cb - U; }
}  // namespace

int harmsum_variant_count() { return kHsCount; }
int harmsum_variant() { return g_hs_variant.load(); }
int set_harmsum_variant(int v)
{
    if (v < 0 || v >= kHsCount) {
        set_error("no harmonic-sum variant %d", v);
        return FDAS_ERR_INVALID;
    }
    g_hs_variant.store(v);
    return FDAS_OK;
}
void harmsum_variant_shape(int v, int* threads, int* rows, int* cols)
{
    const bool ok = v >= 0 && v < kHsCount;
    *threads = ok ? kHsShapes[v].T : 0;
    *rows = ok ? kHsShapes[v].RJ : 0;
    *cols = ok ? kHsShapes[v].CB : 0;
}

int launch_harmsum(const void* d_fund, float* d_out, const void* d_sub, int prec, const HarmSumArgs& hs,
                   int step0, int nsteps, int Z0, const int* d_fullrlo, const int* d_rlo_h,
                   fdas_candidate* d_cands, int* d_ncand, hipStream_t s, int variant)
{
    if (nsteps <= 0) return FDAS_OK;
    const int st = d_out ? 1 : 0;
#define HS_ARGS d_fund, d_out, d_sub, prec, hs, step0, nsteps, Z0, d_fullrlo, d_rlo_h, d_cands, d_ncand, st, s
    if (variant < 0 || variant >= kHsCount) variant = g_hs_variant.load();
    switch (variant) {
        case 1: return harmsum_t<128, 4, 768, 4096>(HS_ARGS);
        case 2: return harmsum_t<128, 4, 896, 4608>(HS_ARGS);
        case 3: return harmsum_t<128, 3, 1024, 4160>(HS_ARGS);
        case 4: return harmsum_t<128, 2, 1024, 4160>(HS_ARGS);
        case 5: return harmsum_t<128, 3, 896, 3840>(HS_ARGS);
        case 6: return harmsum_t<256, 2, 2048, 8256>(HS_ARGS);
        case 7: return harmsum_t<128, 4, 640, 3072>(HS_ARGS);
        default:
            if (tail_waste(hs.U, 896) <= tail_waste(hs.U, 768)) return harmsum_t<128, 4, 896, 4608>(HS_ARGS);
            return harmsum_t<128, 4, 768, 4096>(HS_ARGS);
    }
#undef HS_ARGS
}

}  // namespace fdas
