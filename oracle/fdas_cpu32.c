/*
 * fdas_cpu32.c -- FP32 CPU form of the f-fdot search: the CPU baseline.
 *
 * TEST AND BASELINE INFRASTRUCTURE ONLY.  Built into liboracle.so beside the FP64
 * oracle; only tests/ and bench.py's CPU-baseline leg load it, and the product
 * library never links or calls it.
 *
 * It runs the reference's pipeline (AccelSearch.cu:52-247) in single precision on
 * the host cores, in the reference's own structure: R2C transform of the series,
 * whitening (source/bpd_noise.c:12-92, or_deredden), template bank
 * (source/bpd_subharmonics.cu:233-279: FP64 taps, FP32 rows, FP32 FFT), and per
 * step and harmonic the segment fetch and normalisation
 * (source/bpd_search.cu:130-184, or_segment_amplitudes), the 2x spread, the
 * forward FFT, then per template X * conj(K), the inverse FFT and the powers of
 * the kept columns (source/bpd_search.cu:228-307), and the harmonic sum in the
 * reference's order with its index maps (source/bpd_search.cu:310-360,
 * source/bpd_gpu_kernels.cu:119-170) plus the stage cuts.  Steps run in parallel
 * (OpenMP).  FFTs are a radix-4 Stockham transform of this file, FP32 arithmetic
 * with twiddles computed in double and rounded once.
 */
#include "fdas_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define C32_PI 3.14159265358979323846264338327950288
#define C32_MAXH 16

/* ---------------------------------------------------------------- FFT (FP32) */
typedef struct {
    int n;
    float *w;   /* w[2k], w[2k+1] = cos, sin of -2 pi k / n, k < n */
} fft32_plan;

static int fft32_init(fft32_plan *p, int n)
{
    p->n = n;
    p->w = (float *)malloc(sizeof(float) * 2 * (size_t)n);
    if (!p->w) return -1;
    for (int k = 0; k < n; k++) {
        double a = -2.0 * C32_PI * (double)k / (double)n;
        p->w[2 * k] = (float)cos(a);
        p->w[2 * k + 1] = (float)sin(a);
    }
    return 0;
}

static void fft32_free(fft32_plan *p)
{
    free(p->w);
    p->w = NULL;
}

/* In-place unnormalised transform of x (n complex, interleaved), dir -1 forward,
 * +1 inverse; work holds n complex.  Stockham self-sorting radix-4 passes (stride
 * s = n / length of the current sub-transform) and one radix-2 pass for odd log2 n. */
static void fft32(const fft32_plan *p, float *x, float *work, int dir)
{
    const int n = p->n;
    const float *w = p->w;
    const float sg = (float)dir;   /* sine sign: -1 forward (table), +1 inverse */
    float *a = x, *b = work;
    int len = n, s = 1;
    while (len >= 4) {
        const int m = len / 4;
        for (int q0 = 0; q0 < m; q0++) {
            const float w1r = w[2 * (q0 * s)], w1i = -sg * w[2 * (q0 * s) + 1];
            const float w2r = w[2 * (2 * q0 * s)], w2i = -sg * w[2 * (2 * q0 * s) + 1];
            const float w3r = w[2 * (3 * q0 * s)], w3i = -sg * w[2 * (3 * q0 * s) + 1];
            const float *x0 = a + 2 * (size_t)s * q0, *x1 = a + 2 * (size_t)s * (q0 + m);
            const float *x2 = a + 2 * (size_t)s * (q0 + 2 * m), *x3 = a + 2 * (size_t)s * (q0 + 3 * m);
            float *y0 = b + 2 * (size_t)s * (4 * q0), *y1 = y0 + 2 * (size_t)s;
            float *y2 = y1 + 2 * (size_t)s, *y3 = y2 + 2 * (size_t)s;
            for (int q = 0; q < s; q++) {
                const float ar = x0[2 * q], ai = x0[2 * q + 1], br = x1[2 * q], bi = x1[2 * q + 1];
                const float cr = x2[2 * q], ci = x2[2 * q + 1], dr = x3[2 * q], di = x3[2 * q + 1];
                const float apcr = ar + cr, apci = ai + ci, amcr = ar - cr, amci = ai - ci;
                const float bpdr = br + dr, bpdi = bi + di;
                /* dir * i * (b - d) */
                const float jr = -sg * (bi - di), ji = sg * (br - dr);
                const float t1r = amcr + jr, t1i = amci + ji;
                const float t2r = apcr - bpdr, t2i = apci - bpdi;
                const float t3r = amcr - jr, t3i = amci - ji;
                y0[2 * q] = apcr + bpdr;
                y0[2 * q + 1] = apci + bpdi;
                y1[2 * q] = t1r * w1r - t1i * w1i;
                y1[2 * q + 1] = t1r * w1i + t1i * w1r;
                y2[2 * q] = t2r * w2r - t2i * w2i;
                y2[2 * q + 1] = t2r * w2i + t2i * w2r;
                y3[2 * q] = t3r * w3r - t3i * w3i;
                y3[2 * q + 1] = t3r * w3i + t3i * w3r;
            }
        }
        float *t = a;
        a = b;
        b = t;
        len = m;
        s *= 4;
    }
    if (len == 2) {
        for (int q = 0; q < s; q++) {
            const float ar = a[2 * q], ai = a[2 * q + 1];
            const float br = a[2 * (q + s)], bi = a[2 * (q + s) + 1];
            b[2 * q] = ar + br;
            b[2 * q + 1] = ai + bi;
            b[2 * (q + s)] = ar - br;
            b[2 * (q + s) + 1] = ai - bi;
        }
        float *t = a;
        a = b;
        b = t;
    }
    if (a != x) memcpy(x, a, sizeof(float) * 2 * (size_t)n);
}

void c32_fft(float *x, int n, int dir)
{
    fft32_plan p;
    if (fft32_init(&p, n)) return;
    float *work = (float *)malloc(sizeof(float) * 2 * (size_t)n);
    fft32(&p, x, work, dir);
    free(work);
    fft32_free(&p);
}

/* ---------------------------------------------------------------- context */
struct c32_ctx {
    or_params p;
    int uselen, rlo, nsteps, numharm;
    long nbins;
    float *spec;                  /* whitened spectrum, nbins complex          */
    int Z[C32_MAXH], L[C32_MAXH];
    float *K[C32_MAXH];           /* FFT'd template rows [Z][L] complex         */
    fft32_plan plan[C32_MAXH];
    double t_r2c, t_whiten, t_templates;
};

static int is_pow2l(long n) { return n > 0 && (n & (n - 1)) == 0; }

c32_ctx *c32_create(const float *series, const or_params *p)
{
    const long n = p->n, m = n / 2;
    if (!is_pow2l(m) || m < 4 || p->numharm > C32_MAXH) return NULL;
    c32_ctx *c = (c32_ctx *)calloc(1, sizeof(c32_ctx));
    c->p = *p;
    c->numharm = p->numharm;
    c->uselen = p->uselen ? p->uselen : or_max_uselen(p->fftlen, p->zmax);
    const float nd = (float)n * p->dt;   /* AccelSearch.cu:101, float product */
    c->rlo = (int)floor(1.0 * nd);
    c->nbins = m;
    c->nsteps = or_num_steps(c->nbins, c->rlo, c->uselen);

    /* R2C: the m-point transform of (x[2j], x[2j+1]) and the split into bins */
    double t0 = omp_get_wtime();
    fft32_plan pm;
    float *z = (float *)malloc(sizeof(float) * 2 * (size_t)m);
    float *work = (float *)malloc(sizeof(float) * 2 * (size_t)m);
    c->spec = (float *)malloc(sizeof(float) * 2 * (size_t)m);
    fft32_init(&pm, (int)m);
    memcpy(z, series, sizeof(float) * 2 * (size_t)m);
    fft32(&pm, z, work, -1);
#pragma omp parallel for
    for (long k = 0; k < m; k++) {
        const long k2 = k == 0 ? 0 : m - k;
        const float zr = z[2 * k], zi = z[2 * k + 1], cr = z[2 * k2], ci = -z[2 * k2 + 1];
        const float fer = 0.5f * (zr + cr), fei = 0.5f * (zi + ci);
        const float dr = zr - cr, di = zi - ci;
        const float for_ = 0.5f * di, foi = -0.5f * dr;
        const double a = -C32_PI * (double)k / (double)m;
        const float wr = (float)cos(a), wi = (float)sin(a);
        c->spec[2 * k] = fer + (wr * for_ - wi * foi);
        c->spec[2 * k + 1] = fei + (wr * foi + wi * for_);
    }
    fft32_free(&pm);
    free(z);
    free(work);
    c->t_r2c = omp_get_wtime() - t0;

    t0 = omp_get_wtime();
    or_deredden(c->spec, (unsigned)c->nbins);
    c->t_whiten = omp_get_wtime() - t0;

    t0 = omp_get_wtime();
    int cur[64], mx[64];
    or_harmonic_order(c->numharm, cur, mx);
    for (int h = 0; h < c->numharm; h++) {
        const double f = (double)cur[h] / (double)mx[h];
        const int zmax_h = or_calc_required_z(f, p->zmax);
        c->Z[h] = (zmax_h / 2) * 2 + 1;
        c->L[h] = or_fft_length(c->uselen, mx[h], cur[h], p->zmax);
        const int L = c->L[h];
        fft32_init(&c->plan[h], L);
        c->K[h] = (float *)malloc(sizeof(float) * 2 * (size_t)L * c->Z[h]);
#pragma omp parallel
        {
            double *row = (double *)malloc(sizeof(double) * 2 * (size_t)L);
            float *wk = (float *)malloc(sizeof(float) * 2 * (size_t)L);
#pragma omp for schedule(dynamic)
            for (int k = 0; k < c->Z[h]; k++) {
                float *dst = c->K[h] + 2 * (size_t)L * k;
                or_template_row(-zmax_h + 2 * k, L, row);
                for (int q = 0; q < 2 * L; q++) dst[q] = (float)row[q];
                fft32(&c->plan[h], dst, wk, -1);
            }
            free(row);
            free(wk);
        }
    }
    c->t_templates = omp_get_wtime() - t0;
    return c;
}

void c32_destroy(c32_ctx *c)
{
    if (!c) return;
    for (int h = 0; h < c->numharm; h++) {
        free(c->K[h]);
        fft32_free(&c->plan[h]);
    }
    free(c->spec);
    free(c);
}

int c32_info(const c32_ctx *c, int *uselen, int *nsteps, int *numz0, double *t_r2c,
             double *t_whiten, double *t_templates)
{
    *uselen = c->uselen;
    *nsteps = c->nsteps;
    *numz0 = c->Z[0];
    *t_r2c = c->t_r2c;
    *t_whiten = c->t_whiten;
    *t_templates = c->t_templates;
    return 0;
}

/* One step: plane[Z0][U] receives the summed powers; returns the cells above a
 * stage cut (counted at each stage end) and adds the segment normalisation time. */
static long c32_step(const c32_ctx *c, int step, float *plane, const float *powcut,
                     double *t_norm)
{
    const int U = c->uselen, H = c->numharm, Z0 = c->Z[0];
    const double startr = c->rlo + (double)step * (double)(U / 2);
    const int fullrlo = (int)startr;
    or_harm_geom g[64];
    or_step_geometry(H, c->p.zmax, U, startr, g);
    const int Lmax = c->L[0] > c->L[H - 1] ? c->L[0] : c->L[H - 1];
    float *amps = (float *)malloc(sizeof(float) * 2 * (size_t)Lmax);
    float *X = (float *)malloc(sizeof(float) * 2 * (size_t)Lmax);
    float *y = (float *)malloc(sizeof(float) * 2 * (size_t)Lmax);
    float *wk = (float *)malloc(sizeof(float) * 2 * (size_t)Lmax);
    float *sub = NULL;
    long hits = 0;
    int stage = 0;
    for (int h = 0; h < H; h++) {
        const int L = c->L[h], Zh = c->Z[h], off = 2 * g[h].binoffset;
        const int nr = h == 0 ? U : g[h].numrs;
        const double t0 = omp_get_wtime();
        or_segment_amplitudes(c->spec, c->nbins, g[h].lobin, g[h].numdata, L, amps);
        *t_norm += omp_get_wtime() - t0;
        memset(X, 0, sizeof(float) * 2 * (size_t)L);
        for (int i = 0; i < L / 2; i++) {   /* spread: odd positions zero */
            X[4 * i] = amps[2 * i];
            X[4 * i + 1] = amps[2 * i + 1];
        }
        fft32(&c->plan[h], X, wk, -1);
        float *dst = plane;
        if (h > 0) {
            sub = (float *)realloc(sub, sizeof(float) * (size_t)Zh * nr);
            dst = sub;
        }
        const float scale = (float)(1.0 / ((double)L * (double)L));
        for (int k = 0; k < Zh; k++) {
            const float *K = c->K[h] + 2 * (size_t)L * k;
            for (int q = 0; q < L; q++) {   /* X * conj(K) */
                const float xr = X[2 * q], xi = X[2 * q + 1], kr = K[2 * q], ki = K[2 * q + 1];
                y[2 * q] = xr * kr + xi * ki;
                y[2 * q + 1] = xi * kr - xr * ki;
            }
            fft32(&c->plan[h], y, wk, +1);
            float *o = dst + (size_t)k * nr;
            for (int i = 0; i < nr; i++) {
                const float a = y[2 * (off + i)], b = y[2 * (off + i) + 1];
                o[i] = (a * a + b * b) * scale;
            }
        }
        if (h > 0) {
            const double f = (double)g[h].cur / (double)g[h].max;
            for (int j = 0; j < Z0; j++) {
                const int zz = g[0].zlo + j * 2;
                const int subz = (int)(zz * f);
                const int zind = (int)((subz - g[h].zlo) * 0.5);
                for (int i = 0; i < U; i++) {
                    const double rr = fullrlo + i * 0.5;
                    const double subr = (int)(2.0 * rr * f + 0.5) * 0.5;
                    const unsigned short rind = (unsigned short)((subr - g[h].rlo) * 2.0);
                    plane[(size_t)j * U + i] += sub[(size_t)zind * nr + rind];
                }
            }
        }
        if (h == 0 || ((h + 1) & h) == 0) {
            if (h > 0) stage++;
            const float cut = powcut ? powcut[stage] : 0.0f;
            if (cut > 0.0f)
                for (size_t q = 0; q < (size_t)Z0 * U; q++) hits += plane[q] > cut;
        }
    }
    free(sub);
    free(amps);
    free(X);
    free(y);
    free(wk);
    return hits;
}

void c32_step_plane(const c32_ctx *c, int step, float *out)
{
    double t = 0.0;
    c32_step(c, step, out, NULL, &t);
}

/* Steps steps[0..count) in parallel over nthreads (0: all).  Returns the wall
 * seconds; *hits = cells above their stage cut, *t_norm = thread-seconds spent in
 * segment fetch and normalisation. */
double c32_run_steps(const c32_ctx *c, const int *steps, int count, int nthreads,
                     const float *powcut, long *hits, double *t_norm)
{
    if (nthreads > 0) omp_set_num_threads(nthreads);
    long total = 0;
    double tn = 0.0;
    const double t0 = omp_get_wtime();
#pragma omp parallel reduction(+ : total, tn)
    {
        float *plane = (float *)malloc(sizeof(float) * (size_t)c->Z[0] * c->uselen);
#pragma omp for schedule(dynamic)
        for (int k = 0; k < count; k++) total += c32_step(c, steps[k], plane, powcut, &tn);
        free(plane);
    }
    const double wall = omp_get_wtime() - t0;
    if (hits) *hits = total;
    if (t_norm) *t_norm = tn;
    return wall;
}
