/*
 * fdas_oracle.h -- CPU restatement of the f-fdot acceleration-search path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP path is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * may load it; the product library never links or calls it.
 *
 * Arithmetic: FFTs and correlations run in FP64.  Where the reference defines a
 * value through its float type (spectrum storage, whitening, the fast inverse
 * square root, per-segment medians) the oracle keeps that float semantics, because
 * those types are part of what the reference computes.
 *
 * Complex arrays are interleaved (re, im) pairs.
 */
#ifndef FDAS_ORACLE_H_
#define FDAS_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* ---- geometry (reference: source/bpd_subharmonics.cu, headers/bpd_math.h) ---- */
int    or_nearest_int(double x);
int    or_calc_required_z(double harm_fract, double zfull);
double or_calc_required_r(double harm_fract, double rfull);
int    or_index_from_r(double r, double rlo);
int    or_index_from_z(double z, double zlo);
int    or_z_response_halfwidth(double z);
long long or_next_pow2(long long x);
int    or_fast_log2(int n);
int    or_total_harmonics(int stages);
int    or_fft_length(int uselen, int numharm, int harmnum, int zmax);
int    or_max_uselen(int fftlen, int zmax);
/* harmonic order [1/1, 1/2, 1/4, 3/4, 1/8, ...]; fills cur[], max[] (numharm entries) */
void   or_harmonic_order(int numharm, int *cur, int *max);

/* ---- special functions and templates ---- */
void   or_fresnel(double x, double *s, double *c);
void   or_r_response(double roffset, int numbetween, int numkern, double *out);
void   or_z_response(double roffset, int numbetween, double z, int numkern, double *out);
/* one wrap-around template row of length L for drift z (taps rounded to float) */
void   or_template_row(int z, int L, double *row);

/* ---- FFT (any length; radix-2 for powers of two, Bluestein otherwise) ---- */
void   or_fft(double *x, long n, int sign);            /* in place, unnormalised */
void   or_rfft(const double *x, long n, double *out);  /* out: n/2+1 complex     */

/* ---- order statistics, fast inverse square root, whitening ---- */
float  or_lower_median(const float *a, int n);
float  or_quake_rsqrt(float x);
void   or_deredden(float *fft, unsigned nbins);

/* ---- per-step geometry (reference: source/bpd_search.cu:72-127) ---- */
typedef struct {
    int cur, max;        /* harmonic fraction cur/max                  */
    int rlo;             /* floor(calc_required_r(f, startr))          */
    double rhi;          /* calc_required_r(f, lastr)                  */
    int zlo;             /* calc_required_z(f, -zmax)                  */
    int numrs, numzs, fftlen;
    int binoffset;       /* halfwidth of the widest template           */
    int lobin, numdata;  /* fetch window of the segment                */
} or_harm_geom;

void   or_step_geometry(int numharm, int zmax, int uselen, double startr, or_harm_geom *g);
int    or_num_steps(long nbins_highest, int rlo, int uselen);

/* Normalised local amplitudes of one segment: fetch L/2 bins from lobin
 * (zero outside [0,nbins)), scale the first numdata by rsqrt(median/ln2).
 * out: L/2 complex floats (as float pairs). Returns the median. */
float  or_segment_amplitudes(const float *spec, long nbins, int lobin, int numdata,
                             int L, float *out);

/* Powers of one f-fdot row via FFT correlation (FP64): out[i], i < numrs. */
void   or_row_powers_fft(const float *amps, const double *Krow_fft, int L, int off,
                         int numrs, double *out);
/* Same row by direct correlation (brute force), for pinning the FFT path. */
void   or_row_powers_direct(const float *amps, int z, int L, int off, int numrs,
                            double *out);

/* ---- whole-search driver ---- */
typedef struct {
    long   n;            /* time-series length                         */
    float  dt;           /* sample time (s); rlo = floor(n*dt) in float */
    int    zmax;
    int    numharm;
    int    fftlen;       /* fundamental FFT length                     */
    int    uselen;       /* 0: largest even U that fits fftlen         */
} or_params;

typedef struct or_ctx or_ctx;

/* Spectrum (FP64 R2C, stored as float like the reference), whitening and the
 * template bank (FP64 FFT of float taps). */
or_ctx *or_create(const float *series, const or_params *p);
void    or_destroy(or_ctx *c);
int     or_ctx_info(const or_ctx *c, int *uselen, int *rlo, int *nsteps, int *numz0,
                    long *nbins);
const float *or_ctx_spectrum(const or_ctx *c);  /* whitened, nbins complex */
/* Summed f-fdot plane of one step: out[z*U + i] (double), plus optionally the
 * per-stage partial-sum planes stage_out[s][z*U+i] (NULL to skip). */
void    or_step_plane(const or_ctx *c, int step, double *out, double *stage_out);

/* Search cells above the stage cuts over steps [step0, step0 + nsteps), OpenMP over
 * steps; entries within margin * max(P, 1) of a cut carry ambiguous = 1.  Order:
 * step, then z, r, stage.  Returns the entry count (first cap written), -1 on
 * allocation failure. */
typedef struct {
    int step, z_index, r_index, stage;
    double power;       /* stage sum (FP64)                                 */
    int ambiguous;      /* |power - cut| <= margin * max(power, 1)          */
} or_cand;
long    or_search_candidates(const or_ctx *c, int step0, int nsteps, const float *powcut,
                             double margin, or_cand *out, long cap);

/* ---- candidate optimiser response (build-defined; the reference's optimiser is
 * a stub, source/bpd_candidates.cu:22-26).  R(q, z) = int_0^1 exp(2 pi i (q u +
 * z u^2 / 2)) du by composite Gauss-Legendre quadrature (not the Fresnel form the
 * templates use, source/bpd_subharmonics.cu:146-207), and the matched power
 * |sum_j a_j conj(R(r - z/2 - j, z))|^2 over integers j in [r - hw(z), r + hw(z)). */
void    or_rz_response(double q, double z, double *re, double *im);
double  or_rz_power(const float *spec, long nbins, double r, double z);

/* ---- FP32 CPU form (fdas_cpu32.c): the CPU baseline.  The reference's pipeline
 * in single precision with OpenMP over steps; power-of-two n/2 only. ---- */
typedef struct c32_ctx c32_ctx;
void      c32_fft(float *x, int n, int dir);   /* radix-4 Stockham, unnormalised */
c32_ctx  *c32_create(const float *series, const or_params *p);
void      c32_destroy(c32_ctx *c);
int       c32_info(const c32_ctx *c, int *uselen, int *nsteps, int *numz0, double *t_r2c,
                   double *t_whiten, double *t_templates);
void      c32_step_plane(const c32_ctx *c, int step, float *out);
double    c32_run_steps(const c32_ctx *c, const int *steps, int count, int nthreads,
                        const float *powcut, long *hits, double *t_norm);

#ifdef __cplusplus
}
#endif
#endif
