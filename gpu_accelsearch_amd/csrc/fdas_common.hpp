// fdas_common.hpp -- shared types and helpers for the HIP f-fdot search.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef FDAS_HD
#define FDAS_HD __host__ __device__ __forceinline__
#endif

namespace fdas {

// Interleaved single-precision complex, layout-identical to the reference's
// 8-byte {re, im} complex type.
struct alignas(8) cplx {
    float x, y;
};

FDAS_HD cplx mk(float x, float y) { cplx c; c.x = x; c.y = y; return c; }
FDAS_HD cplx operator+(cplx a, cplx b) { return mk(a.x + b.x, a.y + b.y); }
FDAS_HD cplx operator-(cplx a, cplx b) { return mk(a.x - b.x, a.y - b.y); }
FDAS_HD cplx cmul(cplx a, cplx b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// a * conj(b)
FDAS_HD cplx cmulc(cplx a, cplx b) { return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
FDAS_HD cplx conjg(cplx a) { return mk(a.x, -a.y); }
FDAS_HD cplx scale(cplx a, float s) { return mk(a.x * s, a.y * s); }

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }

// Search constants of the reference (headers/bpd_common.h:30-45).
constexpr int NUMBETWEEN = 2;
constexpr double DR = 0.5;
constexpr int RDR = 2;
constexpr int DZ = 2;
constexpr double RDZ = 0.5;
constexpr double DBLCORRECT = 1e-14;
constexpr int NUMFINTBINS = 16;
constexpr int MAX_HARM = 16;

}  // namespace fdas
