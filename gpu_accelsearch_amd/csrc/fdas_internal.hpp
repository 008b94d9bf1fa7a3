// fdas_internal.hpp -- declarations shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/fdas.h"
#include "fdas_common.hpp"

namespace fdas {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
const char* last_error();
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define FDAS_HIP(call)                                                         \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) return ::fdas::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define FDAS_TRY(call)                  \
    do {                                \
        int rc_ = (call);               \
        if (rc_ != FDAS_OK) return rc_; \
    } while (0)

int check_launch(const char* what);

// ---------------------------------------------------------------- FFT
// Threads per workgroup used for a length-N row transform (N = 64..16384).
int fft_threads(int n);
bool fft_row_supported(long n);
// Device twiddle table of the row transform of length n (cached per device).
const cplx* fft_twiddles(int n, int* err);
// Same for a transform run by t threads (the pass radices depend on n/t).
const cplx* fft_twiddles_t(int n, int t, int* err);
// Table of the wavefront-split transform of length n run by t threads (WaveFFT in
// fft_block.hpp; cached per device).
const cplx* fft_wave_twiddles(int n, int t, int* err);
// Rows of length L into the wavefront-decimated order out[row][(n mod W) L/W + n/W]
// = in[row][n], W = 2^lw (not in place).
int decimate_rows(const cplx* in, cplx* out, int L, int lw, long nrows, hipStream_t s);
// Radix-r step table of the length-r*m transforms (r = 3, 5, 7):
// w^(q k) at (r-1) k + q - 1, w = exp(2 pi i / (r m)), q in [1, r), k in [0, m)
// (cached per device).
const cplx* fft_odd_twiddles(int r, int m, int* err);
// Rows of length r*m (r = 3: m = 256..4096; r = 5: 256..2048; r = 7: 1024)
// forward-transformed into the decimated order:
// out[row*rm + q*m + j] = DFT_-1(in[row*rm + .])[r j + q].  In place allowed.
int fft_odd_rows(const cplx* in, cplx* out, int r, int m, long nrows, hipStream_t s);
// The odd factor of an FFT length: r for r * 2^k (r = 3, 5, 7), 1 for a power of two.
// (The kernels implement r = 3, 5 and 7 x 1024; the radix-r step itself is generic.)
inline int odd_factor(int L)
{
    for (int r : {3, 5, 7})
        if (L % r == 0) return r;
    return 1;
}

// exp(2 pi i idx / M), 0 <= idx < M, as hi[idx >> shift] * lo[idx & (2^shift - 1)]:
// both tables hold values computed in double on the host and rounded to float once
// (about sqrt(M) entries each; cached per device and M for the life of the process),
// so kernels need no double-precision trigonometry.
struct PhaseTable {
    const cplx* lo = nullptr;
    const cplx* hi = nullptr;
    int shift = 0;
};
int phase_table(long long M, PhaseTable* out);
__device__ __forceinline__ cplx phase_at(const PhaseTable& t, long long idx)
{
    return cmul(t.hi[idx >> t.shift], t.lo[idx & ((1LL << t.shift) - 1)]);
}

// Row transforms: out[row*out_stride + k] = scale * DFT_dir(in[row*in_stride + .])[k]
// (times exp(dir 2 pi i (row*k mod tw_m)/tw_m) when tw_m > 0).  In place allowed.
int fft_rows(const cplx* in, cplx* out, int n, long nrows, long in_stride, long out_stride,
             int dir, float scale, long tw_m, hipStream_t s);
// Transpose in[rows][cols] -> out[cols][rows].
int transpose(const cplx* in, cplx* out, long rows, long cols, hipStream_t s);
// Complex transform of any power-of-two length 64..2^28 (four-step beyond 16384).
// work: 2*m complex scratch (unused for m <= 16384).  out may equal in.
int fft_c2c(const cplx* in, cplx* out, long m, int dir, cplx* work, hipStream_t s);
// Real-to-complex transform of n (even) floats -> n/2+1 bins.  The n/2-point
// complex transform of the (even, odd) pairs is a four-step FFT when n/2 is a
// power of two, else a chirp-z (Bluestein) transform through length p >= n-1.
// A plan is geometry only (no device state); work holds work_elems complex.
struct R2CPlan {
    long n = 0, m = 0, p = 0;
    bool pow2 = true;
    size_t work_elems = 0;     // complex scratch the transform needs
};
int r2c_plan_create(long n, R2CPlan* rp);
int fft_r2c_plan(const R2CPlan& rp, const float* in, cplx* out, cplx* work, hipStream_t s);

// ---------------------------------------------------------------- device memory
// Source of a plan's device buffers: hipMalloc/hipFree, or a sub-allocator over
// one reserved region (serving, where hipFree would wait for the whole device).
struct DevAlloc {
    virtual ~DevAlloc() = default;
    virtual int alloc(void** p, size_t bytes) = 0;
    virtual void release(void* p) = 0;
};
DevAlloc* hip_alloc();

// ---------------------------------------------------------------- whitening
struct WhitenSchedule {
    unsigned nbins = 0;
    int nblocks = 0;            // median blocks
    int* d_start = nullptr;     // block start bin
    int* d_len = nullptr;       // block length
    float* d_median = nullptr;  // block medians (filled on the device)
    int nsegs = 0;              // normalisation segments = nblocks + 1
    int* d_seg_lo = nullptr;    // first bin of segment
    int* d_seg_hi = nullptr;    // one past the last bin
};
// Tables are copied on stream s, which is synchronised before returning.
int whiten_schedule_create(unsigned nbins, WhitenSchedule* ws, DevAlloc* al = nullptr,
                           hipStream_t s = 0);
void whiten_schedule_destroy(WhitenSchedule* ws, DevAlloc* al = nullptr);
int whiten(cplx* d_spec, const WhitenSchedule& ws, hipStream_t s);

// ---------------------------------------------------------------- templates
// Host-side geometry of the reference (source/bpd_subharmonics.cu).
int h_nearest_int(double x);
int h_calc_required_z(double f, double zfull);
double h_calc_required_r(double f, double rfull);
int h_z_response_halfwidth(double z);
long long h_next_pow2(long long x);
int h_fast_log2(int n);
int h_fft_length(int uselen, int numharm, int harmnum, int zmax);
int h_max_uselen(int fftlen, int zmax);
// FFT length of harmonic cur/mx under FDAS_LENGTHS_* mode (see include/fdas.h).
int h_harmonic_length(int uselen, int mx, int cur, int zmax, int mode);
// A length R*2^k (R = 3, 5; 7 at 7168) that the correlation and row kernels implement.
bool odd_length_supported(int L);
void h_harmonic_order(int numharm, int* cur, int* max);

// Fill bank[k*L .. (k+1)*L) with the FFT'd template rows z_k = -zmax_h + 2k,
// k < nz (FP64 generation, FP32 storage, FP32 FFT on the device).
int build_template_bank(cplx* d_bank, int zmax_h, int nz, int L, hipStream_t s);

// ---------------------------------------------------------------- search
struct SegDesc {          // one (step, harmonic) segment of the overlap-save loop
    int lobin;            // first spectrum bin fetched
    int numdata;          // bins normalised (median taken over these)
    long long out;        // element offset of its L/2 transformed amplitudes
    int eslot;            // item * H + harmonic: slot of its storage exponent
    int nbins;            // bins of its trial's spectrum (reads beyond give zero)
    long long spec_off;   // offset of its trial's spectrum in the spectrum buffer
    int nch;              // direct correlation: 16-byte chunks per byte array
                          // (out is then in chunks: hi array, lo array after it)
};

struct HarmSumArgs {
    int H;                        // harmonics summed (1 = fundamental only)
    int U;                        // fundamental row width (uselen)
    int zlo0;                     // fundamental zlo (= -zmax)
    int c[MAX_HARM];              // harmonic fraction numerator
    int lm[MAX_HARM];             // log2 of the denominator
    int zlo[MAX_HARM];
    int Z[MAX_HARM];
    int W[MAX_HARM];              // sub-plane row width (a multiple of 4 for h >= 1)
    long long sub_off[MAX_HARM];  // element offset of sub-plane h in the sub buffer
    long long sub_stride[MAX_HARM];  // elements per step (Z*W)
    float powcut[5];
    int cand_cap;
};

// Normalise + transform every segment in desc[0..nseg) (all of FFT length L),
// reading bins spec[d.spec_off + lobin ..] below d.nbins; segment d lands at seg + (d.out - out_base) in the storage format of prec, its
// exponent (reduced precision) at seg_exp[d.eslot - eslot_base].
struct SegPrepArgs {
    const cplx* spec;
    const SegDesc* desc;
    int nseg;
    void* seg;
    long long out_base;
    int* seg_exp;
    int eslot_base;
    int prec;
    int decim;   // log2 W: store the segment in the wavefront-decimated order
                 // a'[(j mod W) (L/2)/W + j/W] = a[j] (0: natural order)
};
int launch_segment_prep(int L, const SegPrepArgs& a, hipStream_t s);

// Fused correlation + inverse FFT + power for one harmonic over a batch of steps:
// out[(s*Z + z)*W + i] = |IFFT(A_s * conj(K_z))[off + i]|^2 / L^2, i < W, with
// A_s at seg + s*seg_stride + seg_off.  Reduced precision: bank/seg packed, with
// exponents bank_exp[z] and seg_exp[s*H + h].
struct CorrArgs {
    const void* bank;
    const int* bank_exp;
    const void* bank_split;      // split-form rows [E | O] (split variants), same format
    const int* bank_split_exp;
    int Z;
    const void* seg;
    long long seg_stride;
    long long seg_off;
    const int* seg_exp;
    int H, h;
    int nsteps;
    int off;
    int W;
    void* out;
    int prec;
};
// variant < 0: the shape currently selected for length L (autotuner table).
int launch_corr(int L, const CorrArgs& a, hipStream_t s, int variant = -1);
int corr_variant_count(int L);
int corr_variant(int L);
int set_corr_variant(int L, int v);
void corr_variant_shape(int L, int v, int* threads, int* steps, int* split = nullptr,
                        int* form = nullptr);
// log2 of the wavefront decimation a correlation variant reads its segment and
// template rows in (0: natural order).
int corr_variant_decim(int L, int v);
// Split-form rows of a full FP32 bank (Z rows of length L).
int launch_bank_split(const cplx* d_bank, int Z, int L, cplx* d_out, hipStream_t s);
// Reduced-precision copy of a bank of Z rows of length L.
int launch_bank_pack(const cplx* d_bank, int Z, int L, int prec, void* d_out, int* d_row_exp,
                     hipStream_t s);

// Direct correlation on the matrix cores (precision FDAS_PREC_I16_MMA, direct.hip;
// index arithmetic in mma_layout.hpp).  A tile is two template rows; its B
// fragments (nk K steps x 4 x 64 lanes x 16 bytes) start at boff (16-byte units).
struct MmaTile {
    int row[2];        // plane rows of its templates (row[1] = -1: none)
    int jlo, nk;       // first tap block, K steps
    long long boff;
};
struct MmaBank {
    void* frags = nullptr;
    MmaTile* tiles = nullptr;
    float* row_inv = nullptr;   // per row: 1 / quantisation scale
    int ntiles = 0, nk_max = 0;
    size_t bytes = 0;
};
// Bank of the Z rows z = -zmax_h + 2k of a harmonic whose largest half-width is bo.
int build_mma_bank(int zmax_h, int Z, int bo, hipStream_t s, MmaBank* out);
void free_mma_bank(MmaBank* b);
// Segment preparation for the direct correlation (segments of FFT-length class L,
// L/2 values); a.seg_exp receives each segment's inverse scale as a float.
int launch_segment_prep_mma(int L, const SegPrepArgs& a, hipStream_t s);
struct MmaCorrArgs {
    const MmaBank* bank;
    const void* seg;           // chunk buffer (16-byte units)
    long long seg_stride;      // chunks per step
    long long seg_off;         // first chunk of this harmonic's hi array
    int nch;                   // chunks per byte array
    const float* seg_inv;      // [step * H + h]
    int H, h, nsteps, W, Z, nm;
    void* out;                 // P16 plane [step][Z][W]
    int variant;               // launch shape (mma_variant at plan creation)
};
int launch_corr_mma(const MmaCorrArgs& a, hipStream_t s);
int mma_variant_count();
int mma_variant();
int set_mma_variant(int v);
int mma_variant_units(int v);   // M tiles per wavefront of a variant
int mma_variant_steps(int v);   // steps per workgroup of a variant
int mma_variant_threads(int v); // threads per workgroup of a variant

// Harmonic sum over d_fund[step][Z0][U] (fundamental powers) and the sub-harmonic
// planes d_sub, both in the plane element of storage mode prec (FP32, or P16 in the
// reduced-precision modes); stage thresholds and candidate emission.  The summed
// FP32 powers go to d_out[step][Z0][U] when d_out is not null (with FP32 planes it
// may equal d_fund).
int harmsum_variant_count();
int harmsum_variant();
int set_harmsum_variant(int v);
void harmsum_variant_shape(int v, int* threads, int* rows, int* cols);
// variant < 0: the shape currently selected process-wide.
int launch_harmsum(const void* d_fund, float* d_out, const void* d_sub, int prec, const HarmSumArgs& hs,
                   int step0, int nsteps, int Z0, const int* d_fullrlo, const int* d_rlo_h,
                   fdas_candidate* d_cands, int* d_ncand, hipStream_t s, int variant = -1);

// Plan creation on a given stream (0: the plan makes its own) with buffers from
// al (nullptr: hipMalloc).  Implemented in plan.hip.
int plan_create_ex(const fdas_params* params, int ntrials, const long* lengths, bool own_series,
                   hipStream_t stream, DevAlloc* al, fdas_plan** out);

// A plan's harmonic-sum tile shape (it starts with the process-wide one).
int plan_set_harmsum_variant(fdas_plan* pl, int v);

// Fundamental FFT length for params (p.fftlen if set; tuned or default else).
int resolve_fftlen(const fdas_params& p);

// Candidates (sift.hip).
double candidate_sigma(double power, int numharm, double numindep);
double power_for_sigma(double sigma, int numharm, double numindep);
double numindep_count(double rlo, double rhi, int numz, int stage);
int sift_candidates(const fdas_sift_trial* trials, int ntrials, int zmax, int uselen, int numz,
                    const fdas_candidate* in, int n_in, const fdas_sift_params* prm,
                    fdas_accelcand* out, int cap, int* n_out);
void cand_properties(fdas_accelcand* c, double T);
int optimize_candidates(const cplx* d_spec, const std::vector<long long>& spec_off,
                        const std::vector<int>& nbins, fdas_accelcand* c, int n, hipStream_t s);
int rz_powers(const cplx* d_spec, int nbins, const double* r, const double* z, int count,
              float* out, hipStream_t s);

// Stage-API kernels (unfused, reference data layout).
int launch_convolve(const cplx* d_fft, const cplx* d_bank, int Z, int L, cplx* d_out,
                    hipStream_t s);
int launch_powers(const cplx* d_ifft, int Z, int numrs, int L, float scale, float* d_out,
                  hipStream_t s);
int launch_add_powers(int Z0, int numrs0, int Lh, int zlo0, int zlo_h, int rlo_h, double f,
                      const cplx* d_ifft_h, float scale, float* d_powers, int fullrlo,
                      hipStream_t s);
int launch_normalize_spread(const cplx* d_amps, int numdata, int nhalf, cplx* d_out,
                            float* d_median_out, hipStream_t s);

}  // namespace fdas
