// templates.hip -- search geometry and the device-resident template bank.
//
// Geometry restates source/bpd_subharmonics.cu:10-67 and headers/bpd_math.h:13.
// Templates (source/bpd_subharmonics.cu:78-279) are generated on the GPU in FP64,
// one thread per tap, rounded to FP32, placed wrap-around in rows of length L
// and transformed in place by the row FFT (rows of length R*2^k, R = 3, 5, 7,
// come out in the decimated order X[R m + q] at q*L/R + m that the length-R*M
// correlation reads).
#include <algorithm>
#include <cmath>

#include "fdas_internal.hpp"
#include "template_tap.hpp"

namespace fdas {

// ------------------------------------------------------------------ geometry
int h_nearest_int(double x) { return (int)(x < 0 ? std::ceil(x - 0.5) : std::floor(x + 0.5)); }

int h_calc_required_z(double f, double zfull) { return h_nearest_int(RDZ * zfull * f) * DZ; }

double h_calc_required_r(double f, double rfull) { return (int)(RDR * rfull * f + 0.5) * DR; }

int h_z_response_halfwidth(double z)
{
    z = std::fabs(z);
    int m = (int)(z * (0.00089 * z + 0.3131) + NUMFINTBINS);
    m = (m < NUMFINTBINS) ? NUMFINTBINS : m;
    if (z > 100 && m > 0.6 * z) m = (int)(0.6 * z);
    return m;
}

long long h_next_pow2(long long x)
{
    long long i = 1;
    while (i < x) i <<= 1;
    return i;
}

int h_fast_log2(int n)
{
    int x = 0;
    while (n > 1) {
        n >>= 1;
        x++;
    }
    return x;
}

int h_fft_length(int uselen, int numharm, int harmnum, int zmax)
{
    const double f = (double)harmnum / (double)numharm;
    const int bins_needed = (uselen * harmnum) / numharm + 2;
    const int end_effects = 2 * NUMBETWEEN * h_z_response_halfwidth(h_calc_required_z(f, zmax));
    return (int)h_next_pow2(bins_needed + end_effects);
}

// FFT length of harmonic cur/mx: the reference's power of two L, or with
// FDAS_LENGTHS_SHORTEST the shortest of 5L/8, 3L/4 and 7L/8 (lengths R*2^k, R = 5,
// 3, 7, that the kernels support: 7L/8 only as 7168) for a sub-harmonic whose
// every step still fits.  With D = ceil(drhi) - floor(drlo) a step's segment holds
// numdata = D + 2 bo + 1 bins and its plane row nr = 2 D + 2 columns from offset
// 2 bo, so the normalised bins (numdata <= L/2) and the template span of the last
// column (2 bo + nr + 2 bo <= L) both fit iff 2 D + 4 bo + 2 <= L.  A step
// starts at an integer bin and f is dyadic, so D depends on the start only
// modulo mx: the residues give the worst step.
bool odd_length_supported(int L)
{
    const int r = odd_factor(L);
    const int m = L / r;
    if (r == 1 || (m & (m - 1))) return false;
    // radix 7 only at 7 x 1024: 3584 (shared LDS, 128 threads) measured slower than 4096
    return r == 3 ? m >= 256 && m <= 4096 : r == 5 ? m >= 256 && m <= 2048 : m == 1024;
}

int h_harmonic_length(int uselen, int mx, int cur, int zmax, int mode)
{
    const int L = h_fft_length(uselen, mx, cur, zmax);
    if (mode != FDAS_LENGTHS_SHORTEST || cur == mx || L < 1024) return L;
    const double f = (double)cur / (double)mx;
    const int bo = h_z_response_halfwidth((double)h_calc_required_z(f, zmax));
    const double half = (double)(uselen / 2);
    int dmax = 0;
    for (int rho = 0; rho < 2 * mx; ++rho) {
        const double drlo = h_calc_required_r(f, (double)rho);
        const double drhi = h_calc_required_r(f, (double)rho + half - DR);
        dmax = std::max(dmax, (int)(std::ceil(drhi) - std::floor(drlo)));
    }
    for (const int c : {5 * (L / 8), 3 * (L / 4), 7 * (L / 8)})
        if (odd_length_supported(c) && 2 * dmax + 4 * bo + 2 <= c) return c;
    return L;
}

int h_max_uselen(int fftlen, int zmax)
{
    int u = fftlen & ~1;
    while (u > 0 && h_fft_length(u, 1, 1, zmax) > fftlen) u -= 2;
    return u;
}

void h_harmonic_order(int numharm, int* cur, int* max)
{
    const int stages = h_fast_log2(numharm) + 1;
    int k = 0;
    for (int s = 0; s < stages; s++) {
        const int mx = (s < 1) ? 1 : (1 << s);
        for (int c = 1; c <= mx; c += 2) {
            cur[k] = c;
            max[k] = mx;
            k++;
        }
    }
}

// Row k of the bank (z = -zmax_h + 2k): taps [2hw, 4hw) go to [0, 2hw) and
// taps [0, 2hw) go to [L-2hw, L) (source/bpd_subharmonics.cu:259-265).  Grid:
// (tap blocks, rows); the bank is zeroed beforehand.
__global__ __launch_bounds__(256) void template_rows_kernel(cplx* __restrict__ bank, int zmax_h,
                                                            int L)
{
    const int k = blockIdx.y;
    const int z = -zmax_h + DZ * k;
    const int hw = d_z_response_halfwidth((double)z);
    const int numkern = 2 * NUMBETWEEN * hw;
    const int half = numkern / 2;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= numkern) return;
    const cplx v = template_tap(z, numkern, i);
    const int pos = (i >= half) ? (i - half) : (L - half + i);
    bank[(size_t)k * L + pos] = v;
}

int build_template_bank(cplx* d_bank, int zmax_h, int nz, int L, hipStream_t s)
{
    const int hwmax = h_z_response_halfwidth((double)zmax_h);
    if (4 * hwmax > L) {
        set_error("template of z=%d (%d taps) does not fit FFT length %d", zmax_h, 4 * hwmax, L);
        return FDAS_ERR_INVALID;
    }
    FDAS_HIP(hipMemsetAsync(d_bank, 0, (size_t)nz * L * sizeof(cplx), s));
    dim3 grid((unsigned)((4 * hwmax + 255) / 256), (unsigned)nz), block(256);
    hipLaunchKernelGGL(template_rows_kernel, grid, block, 0, s, d_bank, zmax_h, L);
    FDAS_TRY(check_launch("template_rows_kernel"));
    const int r = odd_factor(L);
    if (r > 1) return fft_odd_rows(d_bank, d_bank, r, L / r, nz, s);   // decimated order
    if (L <= 16384) return fft_rows(d_bank, d_bank, L, nz, L, L, -1, 1.0f, 0, s);
    cplx* work = nullptr;
    FDAS_HIP(hipMalloc(&work, 2 * (size_t)L * sizeof(cplx)));
    int rc = FDAS_OK;
    for (int k = 0; k < nz && rc == FDAS_OK; ++k)
        rc = fft_c2c(d_bank + (size_t)k * L, d_bank + (size_t)k * L, L, -1, work, s);
    (void)hipStreamSynchronize(s);
    (void)hipFree(work);
    return rc;
}

}  // namespace fdas
