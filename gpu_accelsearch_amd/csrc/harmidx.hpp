// harmidx.hpp -- harmonic-sum index maps shared by the tiled harmonic-sum kernel
// and the fused epilogue of the fundamental correlation.
#pragma once

#include "fdas_internal.hpp"

namespace fdas {

// Sub-harmonic column read by fundamental column col (source/bpd_gpu_kernels.cu:139-141):
// subr*2 = floor((2 rr) f + 1/2) with f = c / 2^lm and q = 2 rr = 2 fullrlo + col,
// less 2 rlo_h.  q*c < 2^31 for series up to 2^27 samples and c < 16.
__device__ __forceinline__ int hs_rind(unsigned q, int c, int lm, int rlo2)
{
    return (int)((q * (unsigned)c + (1u << (lm - 1))) >> lm) - rlo2;
}

// Sub-harmonic row read by fundamental row zz = zlo + 2j: (int)(zz f) truncates
// toward zero, then (subz - zlo_h) / 2 (source/bpd_gpu_kernels.cu:147-153).
__device__ __forceinline__ int hs_zind(int zz, int c, int lm, int zlo_h)
{
    const int x = zz * c;
    const int subz = x >= 0 ? (x >> lm) : -((-x) >> lm);
    return (subz - zlo_h) >> 1;
}

// Append one candidate (cells past the capacity are counted, not stored).
__device__ __forceinline__ void hs_emit(fdas_candidate* cands, int* ncand, int cap, int step,
                                        int z, int col, int stage, float p)
{
    const int idx = atomicAdd(ncand, 1);
    if (idx < cap) {
        fdas_candidate c;
        c.step = step;
        c.z_index = z;
        c.r_index = col;
        c.stage = stage;
        c.power = p;
        c.trial = 0;
        cands[idx] = c;
    }
}

}  // namespace fdas
