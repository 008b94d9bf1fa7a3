// Host simulation of the direct correlation's data layout (mma_layout.hpp): the
// segment chunks and template fragments are built exactly as seg_mma_kernel and
// mma_pack_kernel store them, the 32x32x32 int8 MFMA (template fragments as its A
// operand, segment chunks as its B operand) is emulated with the lane layout
// checked on the device (lane x + 32 g holds A row x / B column x and K bytes
// [16 g, 16 g + 16); register r of lane l holds C[8 (r/4) + 4 g + r%4][x]), and
// every kept column is compared exactly with the direct integer correlation
//   c[2p + phi] = sum_j a[bo + p + j] conj(T_phi[j]).
// Built by tests/test_mma_layout.py with g++ and called through ctypes.
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "../gpu_accelsearch_amd/csrc/mma_layout.hpp"

using namespace fdas;

namespace {
struct Rng {
    uint64_t s;
    int next(int lo, int hi)   // uniform in [lo, hi]
    {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return lo + (int)((s >> 33) % (uint64_t)(hi - lo + 1));
    }
};
struct C64 {
    long long re, im;
};
}  // namespace

// Returns the number of mismatching outputs (0: pass), or -1 for an index
// outside its array.  hw0, hw1: half-widths of the tile's two templates (<= bo).
extern "C" int mma_sim(int bo, int W, int hw0, int hw1, unsigned seed)
{
    Rng rng{seed * 2654435761ULL + 12345};
    int nm = 0, nch = 0;
    mma_geometry(W, bo, &nm, &nch);
    const int nv = 8 * nch;
    const int numdata = (W + 1) / 2 + 2 * bo + 1;
    if (numdata > nv) return -1;
    std::vector<int> ar(nv, 0), ai(nv, 0);
    for (int u = 0; u < numdata; ++u) {
        ar[u] = rng.next(-kQMax, kQMax);
        ai[u] = rng.next(-kQMax, kQMax);
    }
    const int hws[2] = {hw0, hw1};
    std::vector<int> tr[2], ti[2];
    for (int t = 0; t < 2; ++t) {
        tr[t].resize(4 * hws[t]);
        ti[t].resize(4 * hws[t]);
        for (int i = 0; i < 4 * hws[t]; ++i) {
            tr[t][i] = rng.next(-kQMax, kQMax);
            ti[t][i] = rng.next(-kQMax, kQMax);
        }
    }
    // segment byte arrays (value n: re at 2n, im at 2n + 1)
    std::vector<int8_t> hi(2 * nv), lo(2 * nv);
    for (int u = 0; u < nv; ++u) {
        hi[2 * u] = (int8_t)mma_split_hi(ar[u]);
        hi[2 * u + 1] = (int8_t)mma_split_hi(ai[u]);
        lo[2 * u] = (int8_t)mma_split_lo(ar[u]);
        lo[2 * u + 1] = (int8_t)mma_split_lo(ai[u]);
    }
    // tile K range and B fragments [k][f][lane][16]
    int jlo = 0, nk = 0;
    mma_tile_range(bo, hw0 > hw1 ? hw0 : hw1, &jlo, &nk);
    std::vector<int8_t> frag((size_t)nk * 4 * 64 * 16, 0);
    for (int k = 0; k < nk; ++k)
        for (int l = 0; l < 64; ++l) {
            const int n = l & 31, g = l >> 5, t = mma_col_template(n);
            for (int sg = 0; sg < 8; ++sg) {
                const int q = mma_b_tap(bo, hws[t], jlo, n, g, k, sg);
                if (q >= 4 * hws[t]) return -1;
                const int xr = q >= 0 ? tr[t][q] : 0, xi = q >= 0 ? ti[t][q] : 0;
                const int vals[4][2] = {{mma_split_hi(xr), mma_split_hi(xi)},
                                        {mma_split_lo(xr), mma_split_lo(xi)},
                                        {mma_split_hi(-xi), mma_split_hi(xr)},
                                        {mma_split_lo(-xi), mma_split_lo(xr)}};
                for (int f = 0; f < 4; ++f)
                    for (int c = 0; c < 2; ++c)
                        frag[(((size_t)k * 4 + f) * 64 + l) * 16 + 2 * sg + c] = (int8_t)vals[f][c];
            }
        }
    int bad = 0;
    for (int m = 0; m < nm; ++m) {
        // accumulators [6][lane][16]: re HH, re HL+LH, re LL, im HH, im HL+LH, im LL
        std::vector<long long> acc(6 * 64 * 16, 0);
        for (int k = 0; k < nk; ++k) {
            const int8_t* A[2][64];
            for (int l = 0; l < 64; ++l) {
                const int c = mma_a_chunk(m, l & 31, l >> 5, jlo, k);
                if (c < 0 || c >= nch) return -1;
                A[0][l] = &hi[16 * (size_t)c];
                A[1][l] = &lo[16 * (size_t)c];
            }
            const int8_t* B[4][64];
            for (int f = 0; f < 4; ++f)
                for (int l = 0; l < 64; ++l) B[f][l] = &frag[(((size_t)k * 4 + f) * 64 + l) * 16];
            // (accumulator, A half, B fragment): the kernel's 8 MFMAs
            const int ops[8][3] = {{0, 0, 0}, {1, 0, 1}, {1, 1, 0}, {2, 1, 1},
                                   {3, 0, 2}, {4, 0, 3}, {4, 1, 2}, {5, 1, 3}};
            for (const auto& op : ops)
                for (int l = 0; l < 64; ++l)
                    for (int r = 0; r < 16; ++r) {
                        // A operand: template fragments (row n), B: segment chunks (column x)
                        const int n = 8 * (r >> 2) + 4 * (l >> 5) + (r & 3), x = l & 31;
                        long long s = 0;
                        for (int g = 0; g < 2; ++g)
                            for (int q = 0; q < 16; ++q)
                                s += (long long)B[op[2]][n + 32 * g][q] * A[op[1]][x + 32 * g][q];
                        acc[((size_t)op[0] * 64 + l) * 16 + r] += s;
                    }
        }
        for (int l = 0; l < 64; ++l)
            for (int r = 0; r < 16; ++r) {
                const int col = mma_out_col(m, r, l);
                if (col >= W) continue;
                if (mma_col_template(mma_out_n(r, l)) != mma_out_template(r)) return -1;
                const int t = mma_out_template(r), hw = hws[t];
                auto A_ = [&](int q) { return acc[((size_t)q * 64 + l) * 16 + r]; };
                const C64 got{65536 * A_(0) + 256 * A_(1) + A_(2), 65536 * A_(3) + 256 * A_(4) + A_(5)};
                const int p = col / 2, phi = col & 1;
                C64 want{0, 0};
                for (int j = -bo; j <= bo; ++j) {
                    int q = -1;
                    if (phi == 0 && j >= -hw && j < hw) q = 2 * hw + 2 * j;
                    if (phi == 1 && j > -hw && j <= hw) q = 2 * hw + 2 * j - 1;
                    if (q < 0) continue;
                    const long long xr = ar[bo + p + j], xi = ai[bo + p + j];
                    want.re += xr * tr[t][q] + xi * ti[t][q];
                    want.im += xi * tr[t][q] - xr * ti[t][q];
                }
                if (got.re != want.re || got.im != want.im) ++bad;
            }
    }
    return bad;
}
