// mma_layout.hpp -- index arithmetic of the direct correlation on the matrix cores
// (precision mode FDAS_PREC_I16_MMA), shared by the kernels (direct.hip) and the
// host simulation (tests/mma_host_sim.cpp).
//
// The kept columns of a correlation row (source/bpd_search.cu:243-307) are
//   c[2p + phi] = sum_{j = -bo}^{bo} a[bo + p + j] conj(T_phi[j]),
//   T_0[j] = taps[2 hw + 2 j]      (j in [-hw, hw)),
//   T_1[j] = taps[2 hw + 2 j - 1]  (j in [-hw + 1, hw]),
// where a are the segment's normalised amplitudes, taps the template's 4 hw
// response values (source/bpd_subharmonics.cu:146-226), hw the template's
// half-width and bo the harmonic's (the largest).  The 2x spread of
// source/bpd_search.cu:56-67 leaves one tap parity per column, and the power
// |c|^2 equals the reference's |IFFT|^2 / L^2.
//
// Matrix form.  With p = 8 P + rho (rho < 8) and the segment index
// u = bo + p + j = 8 (P + J) + sigma (sigma < 8), output block P at K index
// (J, sigma) reads a[8 (P + J) + sigma] (aligned 16-byte chunks of 8 complex int8
// pairs), and template column (t, rho, phi) holds
// conj T_{z_t, phi}[8 J + sigma - bo - rho].  One 32x32x32 int8 MFMA (16 complex
// taps as 32 int8 values: re, im interleaved) takes the 32 template columns of a
// tile (two templates x 8 shifts x 2 parities) as its A operand and the segment
// chunks of 32 output blocks as its B operand, so it covers 256 output pairs of
// two templates, every segment operand of a lane is one aligned 16-byte chunk,
// and a lane's result registers hold 4 runs of 4 consecutive plane columns.
//
// MFMA lane layout (v_mfma_i32_32x32x32_i8, checked on the device): lane
// l = x + 32 g holds A row x and B column x, K bytes [16 g, 16 g + 16) (A and B
// use the same byte order); accumulator register r of lane l holds
// C[8 (r / 4) + 4 g + r % 4][x].
#pragma once

#ifndef FDAS_HD
#ifdef __HIPCC__
#define FDAS_HD __host__ __device__ __forceinline__
#else
#define FDAS_HD inline
#endif
#endif

namespace fdas {

// Quantisation range: integers in [-kQMax, kQMax], so that a value and its
// negative both split into two signed bytes x = 256 h + l, h = (x + 128) >> 8,
// l in [-128, 127], h in [-127, 127].
constexpr int kQMax = 32639;
// Largest K steps of one template tile (its B fragments are staged in LDS):
// 24 K steps = 96 KB, which holds every tile up to template half-width 180
// (zmax 300).
constexpr int kMmaMaxK = 24;

FDAS_HD int mma_split_hi(int x) { return (x + 128) >> 8; }
FDAS_HD int mma_split_lo(int x) { return x - 256 * mma_split_hi(x); }

// Output blocks of 8 pairs and M tiles of 32 blocks for a row of W columns, and
// chunks of 8 segment values staged per step: every chunk an M tile reads,
// 32 m + x + jlo + 2 k + g, stays below nch.
FDAS_HD void mma_geometry(int W, int bo, int* nm, int* nch)
{
    const int pairs = (W + 1) / 2;
    const int blocks = (pairs + 7) / 8;
    *nm = (blocks + 31) / 32;
    *nch = 32 * *nm + (2 * bo + 7) / 8 + 2;
}

// K range of a tile whose templates have half-width at most hw: tap blocks J in
// [jlo, jlo + 2 nk) cover 8 J + sigma in [bo - hw, bo + hw + 7].
FDAS_HD void mma_tile_range(int bo, int hw, int* jlo, int* nk)
{
    const int lo = (bo - hw) / 8;            // bo >= hw
    const int hi = (bo + hw + 7) / 8;
    *jlo = lo;
    *nk = (hi - lo + 2) / 2;
}

// Template column n of a tile (the A row of lane n mod 32): template t, shift
// rho, parity phi.  The column's plane column within its output block is n % 16.
FDAS_HD int mma_col_template(int n) { return n >> 4; }
FDAS_HD int mma_col_rho(int n) { return (n & 15) >> 1; }
FDAS_HD int mma_col_phi(int n) { return n & 1; }

// Tap index (into the template's 4 hw taps) that column n holds at K step k,
// lane half g, value sigma, or -1 where the template has no tap.
FDAS_HD int mma_b_tap(int bo, int hw, int jlo, int n, int g, int k, int sigma)
{
    const int J = jlo + 2 * k + g;
    const int j = 8 * J + sigma - bo - mma_col_rho(n);
    if (mma_col_phi(n) == 0) return (j >= -hw && j < hw) ? 2 * hw + 2 * j : -1;
    return (j > -hw && j <= hw) ? 2 * hw + 2 * j - 1 : -1;
}

// Segment chunk (8 values) that lane (x, g) of M tile m reads at K step k (the B
// column x of the MFMA is output block 32 m + x).
FDAS_HD int mma_a_chunk(int m, int x, int g, int jlo, int k) { return 32 * m + x + jlo + 2 * k + g; }

// Template column held in accumulator register r of lane l, its template, and the
// plane column of that register in M tile m: registers 4 q .. 4 q + 3 hold 4
// consecutive plane columns of template q / 2.
FDAS_HD int mma_out_n(int r, int l) { return 8 * (r >> 2) + 4 * (l >> 5) + (r & 3); }
FDAS_HD int mma_out_template(int r) { return r >> 3; }
FDAS_HD int mma_out_col(int m, int r, int l) { return 16 * (32 * m + (l & 31)) + (mma_out_n(r, l) & 15); }

}  // namespace fdas
