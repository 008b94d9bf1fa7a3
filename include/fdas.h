/*
 * fdas.h -- C ABI of the MI355X Fourier-domain acceleration search (libfdas.so).
 *
 * Two layers:
 *   1. Reference-named entry points with the reference's signatures, so code
 *      written against keithasr/gpu_accelsearch links unchanged.  Each names the
 *      reference declaration it replaces.  Errors never exit(): they are kept in a
 *      thread-local message (fdas_last_error) and the call returns its "empty"
 *      value; set FDAS_ABORT_ON_ERROR=1 in the environment to abort instead.
 *   2. A plan ("fdas_*") API that returns status codes and runs the fused,
 *      batched search (whole DM trial per call, device-resident data).
 *
 * All compute runs in hand-written HIP kernels for gfx950; nothing falls back
 * to the CPU.  Pointers named d_* / device_* are device memory.
 */
#ifndef FDAS_H_
#define FDAS_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Interleaved single-precision complex; same 8-byte {re, im} layout as the
 * reference's complex type (headers/bpd_common.h:43 includes it from cufft.h). */
typedef struct fdas_complex {
    float x;
    float y;
} fdas_complex;

/* headers/bpd_subharmonics.cuh:13-25 (same field order). kernels_data is the
 * device-resident FFT'd template bank [totalKernelsTemplates][kernelFFTLength]. */
typedef struct SubharmonicInfo {
    unsigned short maxHarmonic;
    unsigned short currentHarmonic;
    int zmax;
    unsigned short totalKernelsTemplates;
    int kernelFFTLength;
    int *z;
    unsigned short *kernelHalfWidth;
    fdas_complex *kernels_data;
} SubharmonicInfo;

/* headers/bpd_search.cuh:15-32 (same field order). */
typedef struct ffdotInfo {
    unsigned short currentHarmonic;
    unsigned short maxHarmonic;
    int numrs;
    int numzs;
    int rlo;
    int zlo;
    double rhi;
    int fftlen;
} ffdotInfo;

/* ------------------------------------------------------------------------ */
/* Search constants. The reference fixes ACCEL_USELEN at compile time        */
/* (headers/bpd_common.h:24, 32000); here it is a process-wide setting.      */
/* ------------------------------------------------------------------------ */
#define FDAS_DEFAULT_USELEN 32000
int fdas_set_uselen(int uselen);   /* even, < 65536; returns 0 or -1 */
int fdas_get_uselen(void);

/* ------------------------------------------------------------------------ */
/* 1. Reference-named entry points                                           */
/* ------------------------------------------------------------------------ */

/* headers/bpd_subharmonics.cuh:27 -- per-harmonic template banks, generated in
 * FP64 on the GPU and FFT'd on the GPU; kernels_data stays on the device. */
SubharmonicInfo *createSubharmonicInfo(int numHarmonicStages, int zhi);
/* Releases what createSubharmonicInfo allocated (the reference frees the
 * fields inline, AccelSearch.cu:274-279). */
void freeSubharmonicInfo(SubharmonicInfo *shi, int totalHarmonics);

/* headers/bpd_subharmonics.cuh:29-33 */
int calc_required_z(double harmonicFraction, double zfull);
double calc_required_r(double harmonicFraction, double rfull);
int index_from_r(double r, double rlo);
int index_from_z(double z, double zlo);
int getTotalHarmonics(int harmonicStages);
/* source/bpd_subharmonics.cu:31 and :60 (file-local there, exported here) */
int z_response_halfwidth(double z);
int calculateFFTLength(unsigned short numharm, unsigned short harmnum, int max_zfull);

/* headers/bpd_fft.cuh:14 -- host series in, malloc'd host spectrum out
 * (n/2+1 bins; caller free()s). Returns the bin count, 0 on error. */
unsigned realToComplexFFT(float *idata, long inputElements, fdas_complex **odata, int batch);
/* headers/bpd_fft.cuh:15-16 -- batched forward / inverse C2C on device data;
 * *device_odata is allocated here (caller releases it with fdas_free). */
unsigned complexToComplexFFT(fdas_complex *device_idata, unsigned int inputElements,
                             fdas_complex **device_odata, int batch);
unsigned inverseComplexToComplexFFT(fdas_complex *device_idata, unsigned int inputElements,
                                    fdas_complex **device_odata, int batch);

/* headers/bpd_noise.h:16 -- whitening of host bins [0, numOfBins), in place. */
void deredden(fdas_complex *fft, unsigned numOfBins);
/* No counterpart in the reference: deredden on DEVICE bins [0, numOfBins), in
 * place, for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  It calls no allocator: its block medians live
 * in a static device array of 4 MiB per GPU (held until the process ends), which
 * limits numOfBins to [16, 2^26].  Calls that overlap on one GPU race on that
 * array.  Queued on hip_stream (a hipStream_t; NULL: the default stream).
 * Returns FDAS_OK, or FDAS_ERR_INVALID for a null pointer or a size outside the
 * limit.  Results are bit-identical to deredden's. */
int fdas_deredden_noalloc(fdas_complex *d_fft, unsigned numOfBins, void *hip_stream);

/* headers/bpd_search.cuh:34-35 -- host data movement helpers. */
fdas_complex *get_fourier_amplitudes(int loBin, int numbins, fdas_complex *fft,
                                     long dataNumbins, long dataLoBin);
void spreadData(fdas_complex *inputData, int inputElements, int interval,
                unsigned int startIdx, fdas_complex *outputData);

/* headers/bpd_search.cuh:37-46 */
void generate_harmonics_ffdot(unsigned short totalHarmonics, unsigned short harmonicStages,
                              SubharmonicInfo *subharmonicsInfo, double fullrlo, double fullrhi,
                              double zlo, ffdotInfo *output_ffdot,
                              unsigned short *output_binOffsetArray);
void generate_harmonic_ffdot_data(unsigned short totalHarmonics, long totalDataLength,
                                  ffdotInfo *ffdot, unsigned short *binoffsetArray,
                                  long *dataStartIdx, fdas_complex *fft, long dataNumbins,
                                  long dataLoBin, fdas_complex *output_data);
void perform_fft_harmonics(unsigned short totalHarmonics, fdas_complex *device_data,
                           long *dataStartIdx, fdas_complex *output_device_fftdata);
void perform_convolution(unsigned short totalHarmonics, fdas_complex *device_fft_data,
                         ffdotInfo *harmonicInfo, SubharmonicInfo *shi, long *dataStartIdx,
                         fdas_complex *output_deviceConvolvedSignal,
                         long *convolutionHarmonicIdx);
void perform_inversefft(unsigned short totalHarmonics, fdas_complex *device_convolvedSignal,
                        long *convolutionHarmonicIdx, fdas_complex *output_device_ifftdata);
void generate_fundamental_powers(unsigned short totalHarmonics, fdas_complex *device_ifft_data,
                                 long *ifftHarmonicIdx, ffdotInfo *harmonicInfo,
                                 SubharmonicInfo *shi, float *output_devicePowers);
void perform_harmonic_summing(unsigned short totalHarmonics, fdas_complex *device_ifft_data,
                              long *ifftHarmonicIdx, ffdotInfo *harmonicInfo,
                              SubharmonicInfo *shi, float *output_devicePowers, int fullrlo,
                              unsigned short *host_rinds, unsigned short *device_rinds,
                              int totalRindsBins);
void initialize_cuda_plans(unsigned short totalHarmonics, SubharmonicInfo *subharmonicInfo);
void free_cuda_plans(unsigned short totalHarmonics);

/* ------------------------------------------------------------------------ */
/* 2. Plan API                                                               */
/* ------------------------------------------------------------------------ */
enum {
    FDAS_OK = 0,
    FDAS_ERR_INVALID = 1,     /* bad parameter                               */
    FDAS_ERR_HIP = 2,         /* HIP runtime error                           */
    FDAS_ERR_NOMEM = 3,       /* device allocation failed                    */
    FDAS_ERR_UNSUPPORTED = 4, /* shape outside what the kernels implement    */
    FDAS_ERR_OVERFLOW = 5     /* more candidates than the plan's capacity    */
};

/* Storage of spectra segments and templates (FFT arithmetic, power planes and
 * power accumulation are FP32 in every mode): FP32, FP16, BF16, or I16 (int16
 * pairs: 16-bit block floating point).  Reduced-precision values carry power-of-
 * two exponents per template row and segment.
 * I16_MMA correlates without FFTs: the normalised segment values and the
 * templates' time-domain taps are quantised to integers of at most 32639 in
 * magnitude (one scale per segment and per template row), and each kept column
 * is an exact integer sum computed on the matrix cores (int8 halves, 32-bit
 * accumulation), scaled back to an FP32 power.  It uses the reference's
 * power-of-two FFT lengths for the segment geometry and supports zmax <= 300. */
enum { FDAS_PREC_FP32 = 0, FDAS_PREC_FP16 = 1, FDAS_PREC_BF16 = 2, FDAS_PREC_I16 = 3,
       FDAS_PREC_I16_MMA = 4 };

#define FDAS_MAX_HARM 16
#define FDAS_MAX_STAGES 5

typedef struct fdas_params {
    long n;             /* time-series length (even)                          */
    float dt;           /* sample time, s; rlo = floor(n*dt) as in the driver */
    int zmax;           /* 0..1200, even                                      */
    int numharm;        /* 1, 2, 4, 8 or 16                                   */
    int fftlen;         /* fundamental FFT length, 1024..16384; 0: the length
                           fdas_autotune chose for this (zmax, numharm,
                           precision), else the smallest valid length whose
                           template overhang is at most 1/8 of it           */
    int uselen;         /* 0: largest even U that fits fftlen                 */
    int precision;      /* FDAS_PREC_*                                        */
    int chunk_steps;    /* steps per batch of launches, 0: automatic          */
    int max_candidates; /* candidate buffer capacity, 0: 1<<20                */
    float powcut[FDAS_MAX_STAGES]; /* per stage (1,2,4,8,16 harmonics);
                                      <= 0 disables that stage's threshold  */
    int fft_lengths;    /* FDAS_LENGTHS_SHORTEST (0) or FDAS_LENGTHS_REFERENCE */
} fdas_params;

/* Sub-harmonic FFT lengths.  The reference rounds every harmonic's length up to
 * a power of two (source/bpd_subharmonics.cu:10-25).  The correlation outputs
 * the search keeps do not depend on that length as long as the segment's bins
 * and the template span of the last kept column fit, so SHORTEST uses the
 * shortest of 5/8, 3/4 and 7/8 of it (lengths 5*2^k and 3*2^k in 768..12288, and
 * 7168) for each sub-harmonic where that holds on every step; REFERENCE keeps
 * the powers of two.  fdas_plan_info.fftlen reports the lengths used. */
#define FDAS_LENGTHS_SHORTEST 0
#define FDAS_LENGTHS_REFERENCE 1

typedef struct fdas_plan_info {
    int uselen;
    int rlo;
    int nsteps;
    int numharm;
    int numz;                          /* fundamental templates (rows)   */
    int fftlen[FDAS_MAX_HARM];
    int numzs[FDAS_MAX_HARM];
    int harm_cur[FDAS_MAX_HARM];
    int harm_max[FDAS_MAX_HARM];
    int chunk_steps;
    long nbins;                        /* n/2: bins whitened and searched */
    long long cells;                   /* nsteps * uselen * numz          */
    long long sub_cells[FDAS_MAX_HARM]; /* sum over steps of numzs*numrs   */
    size_t device_bytes;               /* plan-owned buffers              */
    int ntrials;                       /* DM trials in the plan (1 = single) */
    size_t template_bytes;             /* shared template banks           */
} fdas_plan_info;
/* In a batch plan rlo and nbins are those of trial 0, nsteps and cells are
 * summed over all trials. */

typedef struct fdas_trial_info {
    long n;          /* series length                                      */
    long nbins;      /* n/2                                                */
    int rlo;         /* floor(n*dt)                                        */
    int nsteps;      /* overlap-save steps of this trial                   */
    int first_step;  /* index of its first step among the plan's steps     */
} fdas_trial_info;

typedef struct fdas_candidate {
    int step;      /* step index within its trial                        */
    int z_index;   /* fundamental row, z = -zmax + 2*z_index             */
    int r_index;   /* column, r = rlo + step*uselen/2 + 0.5*r_index      */
    int stage;     /* 0..4: summed over 1,2,4,8,16 harmonics             */
    float power;   /* summed normalised power                            */
    int trial;     /* DM trial (0 in a single-series plan)               */
} fdas_candidate;

typedef struct fdas_plan fdas_plan;

int fdas_plan_create(const fdas_params *params, fdas_plan **plan);
void fdas_plan_destroy(fdas_plan *plan);
int fdas_plan_get_info(const fdas_plan *plan, fdas_plan_info *info);
/* Host-only geometry of a plan with these parameters, without device state: the
 * overlap-save length U and the FFT length of each of the numharm harmonics
 * (harmonic order of fdas_plan_info; FDAS_LENGTHS_* mode applied). */
int fdas_harmonic_fftlens(const fdas_params *params, int *uselen, int lengths[FDAS_MAX_HARM]);
/* Stream the plan's work is queued on (a hipStream_t; NULL = the plan's own). */
int fdas_plan_set_stream(fdas_plan *plan, void *hip_stream);

/* R2C FFT + whitening of a device series of params->n floats; the whitened
 * spectrum stays in the plan.  Resets the candidate list. */
int fdas_prepare(fdas_plan *plan, const float *d_series);
/* Steps [step0, step0+nsteps) of the prepared spectrum.  If d_planes is not
 * NULL the summed f-fdot plane of step s is written to
 * d_planes + (s-step0)*numz*uselen (row z, column i).  Candidates append. */
int fdas_run_steps(fdas_plan *plan, int step0, int nsteps, float *d_planes);
/* prepare + all steps (one DM trial).  Every cell is harmonic-summed and
 * thresholded; the summed planes are not kept (use fdas_run_steps for them). */
int fdas_search(fdas_plan *plan, const float *d_series);
/* Copy of the whitened spectrum (nbins complex) to device memory d_out. */
int fdas_get_spectrum(fdas_plan *plan, fdas_complex *d_out);
/* Candidates found since the last reset, in (trial, step, z, r, stage) order;
 * *count is their number and the first min(capacity, *count) are written.  If
 * more cells passed the thresholds than the plan holds (max_candidates), which of
 * them were stored would depend on the order the GPU found them in, so no list
 * is given: *count is the number found, *overflowed is 1, and the call returns
 * FDAS_ERR_OVERFLOW (raise max_candidates or the power cuts). */
int fdas_get_candidates(fdas_plan *plan, fdas_candidate *out, int capacity, int *count,
                        int *overflowed);
int fdas_reset_candidates(fdas_plan *plan);
int fdas_synchronize(fdas_plan *plan);

/* Batches: many DM trials (any even lengths) searched with one configuration.
 * Their steps are numbered consecutively (trial 0 first) and run in the same
 * chunked launches.  The plan owns one device input buffer per trial.  Template
 * banks are shared by all live plans of the same (zmax, numharm, uselen,
 * precision) on a device. */
int fdas_batch_create(const fdas_params *params, int ntrials, const long *lengths,
                      fdas_plan **plan);   /* params->n is ignored */
int fdas_batch_trial_info(const fdas_plan *plan, int trial, fdas_trial_info *info);
/* Device input buffer of a trial (lengths[trial] floats); fill it before a run. */
float *fdas_batch_series(fdas_plan *plan, int trial);
/* R2C + whitening of one trial (d_series NULL: the plan's buffer). */
int fdas_prepare_trial(fdas_plan *plan, int trial, const float *d_series);
/* Every trial from the plan's buffers: prepare, reset candidates, all steps.
 * With graphs enabled the first call records the work into a hipGraph and
 * later calls replay it (the buffers' contents may change between calls). */
int fdas_batch_run(fdas_plan *plan);
int fdas_plan_enable_graph(fdas_plan *plan, int enable);
int fdas_batch_spectrum(fdas_plan *plan, int trial, fdas_complex *d_out);

/* Per-kernel timing with HIP events recorded on the plan's stream around each
 * launch (enable resets the counters; adds two event records per launch). */
enum {
    FDAS_K_R2C = 0,        /* real-to-complex transform of the series          */
    FDAS_K_WHITEN = 1,     /* block medians + normalisation                    */
    FDAS_K_SEGPREP = 2,    /* segment fetch, median, normalise, forward FFT    */
    FDAS_K_CORR_SUB = 3,   /* fused correlation/IFFT/power, sub-harmonics      */
    FDAS_K_CORR_FUND = 4,  /* fused correlation/IFFT/power, fundamental        */
    FDAS_K_HARMSUM = 5,    /* tiled harmonic sum + stage thresholds            */
    FDAS_K_NCAT = 6
};
typedef struct fdas_timing {
    int launches[8];
    double ms[8];
} fdas_timing;
int fdas_plan_enable_timing(fdas_plan *plan, int enable);
int fdas_plan_get_timing(fdas_plan *plan, fdas_timing *out);

/* ------------------------------------------------------------------------ */
/* Candidates: significance, sifting, optimisation, properties               */
/* ------------------------------------------------------------------------ */
/* Replace the reference's empty search_ffdot_powers / sortCandidates /
 * eliminateHarmonics / optimizeCandidates / calculateProperties
 * (headers/bpd_candidates.cuh:15-49, source/bpd_candidates.cu:4-33) and its unset
 * powersCutoff / independentSpectraCount (AccelSearch.cu:37-38,95-97).        */

/* Equivalent Gaussian sigma of a summed power of numharm harmonics (Gamma(numharm,1)
 * tail) over numindep independent trials, and its inverse. */
double fdas_candidate_sigma(double power, int numharm, double numindep);
double fdas_power_for_sigma(double sigma, int numharm, double numindep);
/* Independent trials of the stage summing 2^stage harmonics over [rlo, rhi) bins
 * and numz drifts: (rhi-rlo)(numz+1)(2/6.95)/2^stage ((rhi-rlo)/2^stage if numz 1). */
double fdas_numindep(double rlo, double rhi, int numz, int stage);
/* powcut[s] = power whose sigma is `sigma` at stage s, for one trial of a plan. */
int fdas_stage_powcuts(const fdas_plan *plan, int trial, double sigma, float powcut[FDAS_MAX_STAGES]);

typedef struct fdas_accelcand {
    int trial;
    int numharm;                 /* harmonics summed at detection (2^stage)        */
    int step, z_index, r_index, stage;   /* search cell of the detection           */
    float power;                 /* summed normalised power at that cell            */
    float sigma;                 /* equivalent Gaussian sigma (stage numindep)      */
    double r, z;                 /* first harmonic: Fourier frequency, drift (bins) */
    double opt_r, opt_z;         /* after fdas_optimize (else r, z)                 */
    float opt_power, opt_sigma;  /* summed power / sigma at (opt_r, opt_z)          */
    float harm_power[FDAS_MAX_HARM]; /* per harmonic at the optimum                 */
    double freq, fdot;           /* Hz, Hz/s at (opt_r, opt_z): r/T, z/T^2          */
    double period, pdot;         /* s, s/s                                          */
} fdas_accelcand;

typedef struct fdas_sift_params {
    double sigma_min;   /* drop candidates below this sigma                         */
    double r_tol;       /* first-harmonic bins for duplicates/harmonics (0: 1.1)    */
    double z_tol;       /* first-harmonic drift bins (0: 4)                         */
    int max_ratio;      /* harmonic ratios n/d, n,d <= max_ratio (0: 16, <0: none)  */
} fdas_sift_params;

typedef struct fdas_sift_trial {
    long n;             /* series length                                          */
    long nbins;         /* searched bins: n/2                                     */
    int rlo;
    float dt;
} fdas_sift_trial;

/* Search candidates -> first-harmonic candidates with sigma, sorted by sigma
 * (descending); duplicates (|dr| <= r_tol, |dz| <= z_tol, any stage) and n/d
 * harmonics of stronger candidates removed.  *n_out = number kept (the first
 * min(capacity, *n_out) are written). */
int fdas_sift(const fdas_plan *plan, const fdas_candidate *in, int n_in,
              const fdas_sift_params *params, fdas_accelcand *out, int capacity, int *n_out);
/* The same on explicit trial geometry (no plan, no device). */
int fdas_sift_raw(const fdas_sift_trial *trials, int ntrials, int zmax, int uselen, int numz,
                  const fdas_candidate *in, int n_in, const fdas_sift_params *params,
                  fdas_accelcand *out, int capacity, int *n_out);
/* Maximise the summed power of each candidate's harmonics over continuous (r, z)
 * on the plan's whitened spectra (grid refinement on the GPU, FP64 response),
 * then fill opt_*, harm_power and the properties. */
int fdas_optimize(fdas_plan *plan, fdas_accelcand *cands, int n);
/* |sum_j a_j conj(R(r - z/2 - j, z))|^2, j in [r - hw(z), r + hw(z)), on a trial's
 * whitened spectrum (no normalisation) at count points, with R(q, z) =
 * int_0^1 exp(2 pi i (q u + z u^2/2)) du: the response the optimiser maximises. */
int fdas_rz_power(fdas_plan *plan, int trial, const double *r, const double *z, int count,
                  float *out);

/* ------------------------------------------------------------------------ */
/* Autotuning                                                                */
/* ------------------------------------------------------------------------ */
/* Measures, on this device and synthetic data: (1) every launch shape of the
 * fused correlation (threads per workgroup = radix of the LDS passes, steps per
 * workgroup = template-block tile) at each FFT length the candidates use, and
 * selects the fastest process-wide; (2) the harmonic-sum tile shape over whole
 * trials at the first valid candidate length; (3) whole DM trials of base->n samples at
 * each candidate fundamental FFT length (nfftlen 0: 1024..16384), remembering
 * the fastest for plans with fftlen 0 and the same (zmax, numharm, precision). */
#define FDAS_TUNE_MAX 12
#define FDAS_TUNE_MAXV 16
typedef struct fdas_tune_report {
    int fftlen;                              /* chosen fundamental length        */
    int ncand;
    int cand_fftlen[FDAS_TUNE_MAX];
    double cand_ms_per_trial[FDAS_TUNE_MAX]; /* < 0: not valid for the config    */
    int ncorr;
    int corr_len[FDAS_TUNE_MAX];             /* FFT lengths whose shape was tuned */
    int corr_variant[FDAS_TUNE_MAX];         /* selected variant                  */
    int corr_threads[FDAS_TUNE_MAX];
    int corr_steps[FDAS_TUNE_MAX];
    int corr_split[FDAS_TUNE_MAX];           /* 1: two L/2 transforms per step    */
    double corr_us[FDAS_TUNE_MAX][FDAS_TUNE_MAXV]; /* us per launch per variant (0: n/a) */
    int harmsum_variant;                     /* selected harmonic-sum tile shape   */
    int harmsum_threads, harmsum_rows, harmsum_cols;  /* cols 0: 896 or 768, by row width */
    double harmsum_ms[FDAS_TUNE_MAXV];       /* ms per trial per shape (0: n/a)    */
    /* LDS exchange form of corr_variant: 0 one region per transform; 1 the split
     * form (powers of two) or one region shared by the R transforms (lengths
     * R*2^k); 2K a ring of two slots of K regions, one barrier per exchange; 8
     * the wavefront-split form (powers of two: each wavefront transforms 1/W of
     * the length with exchanges inside the wavefront, then one exchange across
     * the W wavefronts per step) */
    int corr_form[FDAS_TUNE_MAX];
} fdas_tune_report;

int fdas_autotune(const fdas_params *base, const int *fftlens, int nfftlen,
                  fdas_tune_report *report);
/* The selections below (and the autotuner's) apply to plans created afterwards:
 * a plan keeps the launch shapes in effect when it was created, so a tuning run
 * never changes the kernels, and hence the rounding, of a plan or server request
 * already running. */
int fdas_corr_variant_count(int fftlen);
int fdas_harmsum_variant_count(void);
int fdas_get_harmsum_variant(void);
int fdas_set_harmsum_variant(int variant);
int fdas_get_corr_variant(int fftlen);
int fdas_set_corr_variant(int fftlen, int variant);
/* Direct correlation (FDAS_PREC_I16_MMA): variant v keeps
 * fdas_mma_variant_units(v) M tiles (256 output pairs of two templates each) per
 * wavefront in flight, and a workgroup of fdas_mma_variant_threads(v) threads
 * takes fdas_mma_variant_steps(v) steps of one template tile. */
int fdas_mma_variant_count(void);
int fdas_get_mma_variant(void);
int fdas_set_mma_variant(int variant);
int fdas_mma_variant_units(int variant);
int fdas_mma_variant_steps(int variant);
int fdas_mma_variant_threads(int variant);
/* Shape of a correlation variant: threads, steps per workgroup and LDS exchange
 * form (as fdas_tune_report.corr_form; 1 at a power of two is the split form). */
int fdas_corr_variant_shape(int fftlen, int variant, int *threads, int *steps_per_wg,
                            int *lds_form);
/* The fundamental FFT length a plan with these parameters would use. */
int fdas_resolve_fftlen(const fdas_params *params);

/* ------------------------------------------------------------------------ */
/* 3. Serving: a continuous-batching request loop                            */
/* ------------------------------------------------------------------------ */
/* Requests (one DM-trial series each, any zmax / numharm / fftlen / precision)
 * are admitted in arrival order while one of max_active slots is free and the
 * request's plan (its f-fdot planes for chunk_steps steps, spectrum, segments,
 * candidates) fits in the server's reserved device memory; admitted requests
 * run concurrently on their own streams.  Templates stay resident per
 * configuration.  A request's results do not depend on what runs beside it. */
typedef struct fdas_server fdas_server;

typedef struct fdas_server_config {
    size_t memory_bytes;   /* device memory reserved for requests (0: half of free) */
    int max_active;        /* requests on the device at once, 1..16 (0: 4)          */
    int chunk_steps;       /* default steps per launch group (0: 256)               */
    int max_candidates;    /* default candidate capacity per request (0: 65536)     */
} fdas_server_config;

typedef struct fdas_request_result {
    long long id;
    int status;            /* FDAS_OK or the error that ended the request        */
    int ncandidates;       /* candidates found; with status FDAS_ERR_OVERFLOW
                              (more than the request's capacity) none are returned */
    int overflowed;
    int nsteps;
    long long cells;
    size_t device_bytes;   /* footprint admitted against the memory budget        */
    double queue_ms;       /* submit -> admitted                                  */
    double service_ms;     /* admitted -> complete on the device                  */
    double latency_ms;     /* submit -> complete                                  */
    char message[256];     /* error text when status != FDAS_OK                   */
} fdas_request_result;

typedef struct fdas_server_stats {
    long long submitted, completed, failed;
    int active, pending, max_active_seen;
    size_t memory_bytes, memory_in_use, memory_peak;
    double p50_ms, p99_ms, mean_ms;   /* latency of completed requests           */
    double busy_s;                    /* time with at least one request active   */
    double trials_per_s;              /* completed / (last completion - first submit) */
} fdas_server_stats;

int fdas_server_create(const fdas_server_config *cfg, fdas_server **srv);
/* Finishes queued and running requests, then frees everything. */
void fdas_server_destroy(fdas_server *srv);
/* Builds (or finds) the template banks of a configuration and keeps them. */
int fdas_server_prewarm(fdas_server *srv, const fdas_params *params);
/* Queues a copy of h_series (params->n floats); returns at once with its id.
 * params->chunk_steps / max_candidates of 0 take the server's defaults. */
int fdas_server_submit(fdas_server *srv, const fdas_params *params, const float *h_series,
                       long long *id);
/* Blocks until request id is done; fills res and up to capacity candidates
 * (trial 0, step within the series), then forgets the request. */
int fdas_server_wait(fdas_server *srv, long long id, fdas_request_result *res,
                     fdas_candidate *cands, int capacity);
int fdas_server_get_stats(fdas_server *srv, fdas_server_stats *out);
int fdas_server_reset_stats(fdas_server *srv);

/* ------------------------------------------------------------------------ */
/* Input files                                                               */
/* ------------------------------------------------------------------------ */
/* headers/bpd_data.h:13 readData: raw float32 samples (at most 2^26) of a PRESTO
 * .dat file into a buffer with 2000 zero floats before and after the data;
 * *output_data points at the first sample.  Returns the sample count, or -1
 * with fdas_last_error set (the reference exits).  Free with fdas_free_data. */
long readData(char *path, float **output_data);
void fdas_free_data(float *data);

/* PRESTO .inf metadata; the reference's driver hard-codes the sample time
 * (AccelSearch.cu:90). */
typedef struct fdas_inf {
    char basename[256];
    char telescope[64];
    long n;              /* Number of bins in the time series      */
    double dt;           /* Width of each time series bin (sec)    */
    double dm;           /* Dispersion measure (cm-3 pc)           */
    double epoch_mjd;    /* Epoch of observation (MJD)             */
    int barycentred;
} fdas_inf;
int fdas_read_inf(const char *path, fdas_inf *out);

/* ------------------------------------------------------------------------ */
/* Memory plumbing and diagnostics                                           */
/* ------------------------------------------------------------------------ */
int fdas_malloc(void **d_ptr, size_t bytes);
int fdas_free(void *d_ptr);
int fdas_memcpy_htod(void *d_dst, const void *h_src, size_t bytes);
int fdas_memcpy_dtoh(void *h_dst, const void *d_src, size_t bytes);
int fdas_memcpy_dtod(void *d_dst, const void *d_src, size_t bytes);
int fdas_memset(void *d_dst, int value, size_t bytes);
int fdas_device_synchronize(void);
const char *fdas_last_error(void);
void fdas_clear_error(void);
const char *fdas_version(void);

#ifdef __cplusplus
}
#endif

#endif /* FDAS_H_ */
