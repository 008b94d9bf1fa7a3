/*
 * fdas_oracle.c -- CPU restatement of the f-fdot acceleration-search path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's CPU-baseline leg; never by the
 * product library.  Each function names the reference file:line it restates.
 *
 * Build with -ffp-contract=off: the reference's float expressions are evaluated
 * with separate roundings (x86-64 SSE, no FMA), and the oracle keeps that.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   geometry        hand-derived tables + the USELEN/FFT-length pairs listed in
 *                   headers/bpd_common.h:14-28
 *   or_fresnel      scipy.special.fresnel, mpmath, and the reference's own
 *                   fresnl() built from source under oracle/_ref
 *   templates       mpmath quadrature of the chirp response, z->0 limit
 *   or_fft/or_rfft  numpy.fft (FP64 pocketfft) and brute-force DFT
 *   median, rsqrt   the reference's fast_median()/inverse_square_root() built
 *                   from source under oracle/_ref
 *   or_deredden     closed forms: constant spectrum, monotone spectrum (known
 *                   block medians), power-of-two scale invariance
 *   correlation     brute-force direct correlation (or_row_powers_direct)
 *   harmonic sum    index-map properties and a planted single-cell check
 */
#include "fdas_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_NUMBETWEEN 2
#define OR_DR 0.5
#define OR_RDR 2
#define OR_DZ 2
#define OR_RDZ 0.5
#define OR_DBLCORRECT 1e-14
#define OR_NUMFINTBINS 16
#define OR_PI 3.14159265358979323846264338327950288
#define OR_SQRT2 1.41421356237309504880168872420969808

/* ------------------------------------------------------------------------- */
/* geometry                                                                   */
/* ------------------------------------------------------------------------- */

/* headers/bpd_math.h:13 -- round half away from zero */
int or_nearest_int(double x) { return (int)(x < 0 ? ceil(x - 0.5) : floor(x + 0.5)); }

/* source/bpd_subharmonics.cu:10-13 */
int or_calc_required_z(double f, double zfull)
{
    return or_nearest_int(OR_RDZ * zfull * f) * OR_DZ;
}

/* source/bpd_subharmonics.cu:16-19 */
double or_calc_required_r(double f, double rfull)
{
    return (int)(OR_RDR * rfull * f + 0.5) * OR_DR;
}

/* source/bpd_subharmonics.cu:48-57 */
int or_index_from_r(double r, double rlo) { return (int)((r - rlo) * OR_RDR + OR_DBLCORRECT); }
int or_index_from_z(double z, double zlo) { return (int)((z - zlo) * OR_RDZ + OR_DBLCORRECT); }

/* source/bpd_subharmonics.cu:31-45 */
int or_z_response_halfwidth(double z)
{
    z = fabs(z);
    int m = (int)(z * (0.00089 * z + 0.3131) + OR_NUMFINTBINS);
    if (m < OR_NUMFINTBINS) m = OR_NUMFINTBINS;
    if (z > 100 && m > 0.6 * z) m = (int)(0.6 * z);
    return m;
}

/* source/bpd_math.c:113-122 */
long long or_next_pow2(long long x)
{
    long long i = 1;
    while (i < x) i <<= 1;
    return i;
}

/* source/bpd_math.c:4-14 */
int or_fast_log2(int n)
{
    int x = 0;
    while (n > 1) { n >>= 1; x++; }
    return x;
}

/* source/bpd_subharmonics.cu:228-231 */
int or_total_harmonics(int stages) { return 1 << (stages - 1); }

/* source/bpd_subharmonics.cu:60-67, with ACCEL_USELEN as a parameter */
int or_fft_length(int uselen, int numharm, int harmnum, int zmax)
{
    double f = (double)harmnum / (double)numharm;
    int bins_needed = (uselen * harmnum) / numharm + 2;
    int end_effects = 2 * OR_NUMBETWEEN * or_z_response_halfwidth(or_calc_required_z(f, zmax));
    return (int)or_next_pow2(bins_needed + end_effects);
}

/* Largest even U whose fundamental FFT length is <= fftlen (the reference fixes
 * U at compile time, headers/bpd_common.h:14-28, e.g. 7470 -> 8K at zmax=300). */
int or_max_uselen(int fftlen, int zmax)
{
    int u = fftlen;
    if (u & 1) u--;
    while (u > 0 && or_fft_length(u, 1, 1, zmax) > fftlen) u -= 2;
    return u;
}

/* source/bpd_subharmonics.cu:319-354 (stage loop order) */
void or_harmonic_order(int numharm, int *cur, int *max)
{
    int stages = or_fast_log2(numharm) + 1, k = 0;
    for (int s = 0; s < stages; s++) {
        int mx = (s < 1) ? 1 : (1 << s);
        for (int c = 1; c <= mx; c += 2) { cur[k] = c; max[k] = mx; k++; }
    }
}

/* ------------------------------------------------------------------------- */
/* Fresnel integrals C(x)=int_0^x cos(pi t^2/2), S(x)=int_0^x sin(pi t^2/2)     */
/* Own FP64 method: Maclaurin series for |x| < 1.5; for |x| >= 1.5 the relation */
/* C+iS = (1+i)/2 * (1 - erfc(w)), w = sqrt(pi)/2 (1-i) x, with erfc from its   */
/* Laplace continued fraction evaluated by the modified Lentz method.           */
/* Follows the role of fresnl(), source/bpd_fresnl.c:160-200 (including its    */
/* x > 36974 -> 0.5 saturation).                                                */
/* ------------------------------------------------------------------------- */
void or_fresnel(double xin, double *s_out, double *c_out)
{
    double x = fabs(xin), c, s;
    if (x < 1.5) {
        /* x * sum_k sgn_k a^k / (k! (2k+1)), a = pi x^2 / 2; even k -> C, odd k -> S */
        double a = 0.5 * OR_PI * x * x, t = 1.0, sc = 0.0, ss = 0.0;
        for (int k = 0; k < 60; k++) {
            if (k > 0) t *= a / k;
            double term = t / (2 * k + 1);
            int q = k & 3;
            double sg = (q < 2) ? 1.0 : -1.0;
            if (k & 1) ss += sg * term; else sc += sg * term;
            if (t < 1e-18) break;
        }
        c = x * sc;
        s = x * ss;
    } else if (x > 36974.0) {
        c = 0.5;
        s = 0.5;
    } else {
        /* h = 1/(b1 + a2/(b2 + a3/(b3 + ...))), b_k = 1 - i pi x^2 + 4(k-1),
         * a_k = -(2k-3)(2k-2); erfc(w) = (1-i) x e^{i pi x^2/2} h             */
        double px2 = OR_PI * x * x;
        double br = 1.0, bi = -px2;
        const double tiny = 1e-300;
        double cr = 1.0 / tiny, ci = 0.0;
        double den = br * br + bi * bi;
        double dr = br / den, di = -bi / den;
        double hr = dr, hi = di;
        for (int k = 2; k < 2000; k++) {
            double n = 2 * k - 3;
            double an = -n * (n + 1.0);
            br += 4.0;
            /* d = 1/(an*d + b) */
            double tr = an * dr + br, ti = an * di + bi;
            den = tr * tr + ti * ti;
            dr = tr / den; di = -ti / den;
            /* c = b + an/c */
            den = cr * cr + ci * ci;
            double qr = an * cr / den, qi = -an * ci / den;
            cr = br + qr; ci = bi + qi;
            double delr = cr * dr - ci * di, deli = cr * di + ci * dr;
            double nhr = hr * delr - hi * deli, nhi = hr * deli + hi * delr;
            hr = nhr; hi = nhi;
            if (fabs(delr - 1.0) + fabs(deli) < 1e-17) break;
        }
        /* h *= (1-i) x */
        double gr = (hr + hi) * x, gi = (hi - hr) * x;
        double ph = 0.5 * px2;
        double er = cos(ph), ei = sin(ph);
        double pr = er * gr - ei * gi, pi_ = er * gi + ei * gr;
        double ur = 1.0 - pr, ui = -pi_;
        /* (0.5 + 0.5i) * u */
        c = 0.5 * (ur - ui);
        s = 0.5 * (ur + ui);
    }
    if (xin < 0.0) { c = -c; s = -s; }
    *c_out = c;
    *s_out = s;
}

/* source/bpd_subharmonics.cu:78-135: Fourier interpolation (z = 0) response */
void or_r_response(double roffset, int numbetween, int numkern, double *out)
{
    double startr = OR_PI * (numkern / (double)(2 * numbetween) + roffset);
    double delta = -OR_PI / numbetween;
    double tmp = sin(0.5 * delta);
    double alpha = -2.0 * tmp * tmp;
    double beta = sin(delta);
    double c = cos(startr), s = sin(startr);
    double r = startr;
    for (int i = 0; i < numkern; i++, r += delta) {
        double sinc = (r == 0.0) ? 1.0 : s / r;
        out[2 * i] = c * sinc;
        out[2 * i + 1] = s * sinc;
        double c0 = c;
        c = alpha * c0 - beta * s + c0;
        s = alpha * s + beta * c0 + s;
    }
    if (roffset < 1e-3) {
        /* Taylor expansion of the response about roffset = 0 */
        double r2 = roffset * roffset;
        out[2 * (numkern / 2)] = 1.0 - (2.0 * OR_PI * OR_PI / 3.0) * r2;
        out[2 * (numkern / 2) + 1] = roffset * (OR_PI - (OR_PI * OR_PI * OR_PI / 3.0) * r2);
    }
}

/* source/bpd_subharmonics.cu:146-226: constant-f-dot (z) response */
void or_z_response(double roffset, int numbetween, double z, int numkern, double *out)
{
    double absz = fabs(z);
    if (absz < 1e-4) {
        or_r_response(roffset, numbetween, numkern, out);
        return;
    }
    double startr = roffset - 0.5 * z;
    double ip;
    double startroffset = (startr < 0) ? 1.0 + modf(startr, &ip) : modf(startr, &ip);
    int signz = (z < 0.0) ? -1 : 1;
    double zd = signz * OR_SQRT2 / sqrt(absz);
    double cons = zd / 2.0;
    double pibyz = OR_PI / z;
    startr += numkern / (double)(2 * numbetween);
    double delta = -1.0 / numbetween;
    double r = startr;
    for (int i = 0; i < numkern; i++, r += delta) {
        double yy = r * zd;
        double zz = yy + z * zd;
        double xx = pibyz * r * r;
        double c = cos(xx), s = sin(xx);
        double fsy, fcy, fsz, fcz;
        or_fresnel(yy, &fsy, &fcy);
        or_fresnel(zz, &fsz, &fcz);
        double tr = signz * (fcz - fcy);
        double ti = fsy - fsz;
        out[2 * i] = (tr * c - ti * s) * cons;
        out[2 * i + 1] = -(tr * s + ti * c) * cons;
    }
    if (startroffset < 1e-3 && absz < 1e-3) {
        /* second-order expansion about (roffset, z) = (0, 0) */
        double pi2 = OR_PI * OR_PI, pi3 = pi2 * OR_PI;
        double zz = z * z, xx = startroffset * startroffset;
        int h = numkern / 2;
        double re = 1.0 - (pi2 / 60.0) * zz;
        double im = -(OR_PI / 6.0) * z;
        re += startroffset * (pi2 / 6.0) * z;
        im += startroffset * (OR_PI - (pi3 / 60.0) * zz);
        re += xx * (-(2.0 * pi2 / 3.0) + (pi2 * pi2 / 105.0) * zz);
        im += xx * ((pi3 / 10.0) * z);
        out[2 * h] = re;
        out[2 * h + 1] = im;
    }
}

/* source/bpd_subharmonics.cu:253-266: taps rounded to float, second half at
 * [0, 2hw), first half at [L-2hw, L). */
void or_template_row(int z, int L, double *row)
{
    int hw = or_z_response_halfwidth((double)z);
    int numkern = 2 * OR_NUMBETWEEN * hw;
    int half = numkern / 2;
    double *taps = (double *)malloc(sizeof(double) * 2 * numkern);
    or_z_response(0.0, OR_NUMBETWEEN, (double)z, numkern, taps);
    memset(row, 0, sizeof(double) * 2 * L);
    for (int q = 0; q < half; q++) {
        row[2 * q] = (float)taps[2 * (half + q)];
        row[2 * q + 1] = (float)taps[2 * (half + q) + 1];
        row[2 * (L - half + q)] = (float)taps[2 * q];
        row[2 * (L - half + q) + 1] = (float)taps[2 * q + 1];
    }
    free(taps);
}

/* ------------------------------------------------------------------------- */
/* FFT (restates the cuFFT transforms by their definition:                    */
/*   X[k] = sum_n x[n] exp(sign * 2 pi i n k / n), unnormalised)               */
/* ------------------------------------------------------------------------- */
static int is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }

static void fft_pow2(double *x, long n, int sign)
{
    /* bit reversal */
    for (long i = 1, j = 0; i < n; i++) {
        long bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = x[2 * i], ti = x[2 * i + 1];
            x[2 * i] = x[2 * j]; x[2 * i + 1] = x[2 * j + 1];
            x[2 * j] = tr; x[2 * j + 1] = ti;
        }
    }
    if (n < 2) return;
    double *tw = (double *)malloc(sizeof(double) * n);   /* n/2 complex */
    for (long k = 0; k < n / 2; k++) {
        double a = 2.0 * OR_PI * (double)k / (double)n;
        tw[2 * k] = cos(a);
        tw[2 * k + 1] = sign * sin(a);
    }
    for (long len = 2; len <= n; len <<= 1) {
        long half = len >> 1, step = n / len;
        for (long i = 0; i < n; i += len) {
            for (long k = 0; k < half; k++) {
                double wr = tw[2 * k * step], wi = tw[2 * k * step + 1];
                double *a = x + 2 * (i + k), *b = x + 2 * (i + k + half);
                double br = b[0] * wr - b[1] * wi, bi = b[0] * wi + b[1] * wr;
                b[0] = a[0] - br; b[1] = a[1] - bi;
                a[0] += br; a[1] += bi;
            }
        }
    }
    free(tw);
}

static void fft_bluestein(double *x, long n, int sign)
{
    long m = (long)or_next_pow2(2 * n - 1);
    double *c = (double *)malloc(sizeof(double) * 2 * n);
    double *a = (double *)calloc(2 * m, sizeof(double));
    double *b = (double *)calloc(2 * m, sizeof(double));
    long long n2 = 2LL * n;
    for (long k = 0; k < n; k++) {
        long long q = ((long long)k * k) % n2;        /* exact chirp phase */
        double ang = OR_PI * (double)q / (double)n;
        c[2 * k] = cos(ang);
        c[2 * k + 1] = sign * sin(ang);
    }
    for (long k = 0; k < n; k++) {            /* a = x * c */
        double xr = x[2 * k], xi = x[2 * k + 1];
        a[2 * k] = xr * c[2 * k] - xi * c[2 * k + 1];
        a[2 * k + 1] = xr * c[2 * k + 1] + xi * c[2 * k];
    }
    for (long k = 0; k < n; k++) {            /* b = conj(c), two-sided */
        b[2 * k] = c[2 * k]; b[2 * k + 1] = -c[2 * k + 1];
        if (k > 0) { b[2 * (m - k)] = c[2 * k]; b[2 * (m - k) + 1] = -c[2 * k + 1]; }
    }
    fft_pow2(a, m, -1);
    fft_pow2(b, m, -1);
    for (long k = 0; k < m; k++) {
        double ar = a[2 * k], ai = a[2 * k + 1], br = b[2 * k], bi = b[2 * k + 1];
        a[2 * k] = ar * br - ai * bi;
        a[2 * k + 1] = ar * bi + ai * br;
    }
    fft_pow2(a, m, +1);
    for (long k = 0; k < n; k++) {
        double vr = a[2 * k] / m, vi = a[2 * k + 1] / m;
        x[2 * k] = vr * c[2 * k] - vi * c[2 * k + 1];
        x[2 * k + 1] = vr * c[2 * k + 1] + vi * c[2 * k];
    }
    free(c); free(a); free(b);
}

void or_fft(double *x, long n, int sign)
{
    if (n <= 1) return;
    if (is_pow2(n)) fft_pow2(x, n, sign);
    else fft_bluestein(x, n, sign);
}

/* source/bpd_fft.cu:4-78 (cufftExecR2C): X[k], k = 0..n/2 */
void or_rfft(const double *xin, long n, double *out)
{
    double *w = (double *)malloc(sizeof(double) * 2 * n);
    for (long i = 0; i < n; i++) { w[2 * i] = xin[i]; w[2 * i + 1] = 0.0; }
    or_fft(w, n, -1);
    memcpy(out, w, sizeof(double) * 2 * (n / 2 + 1));
    free(w);
}

/* ------------------------------------------------------------------------- */
/* order statistic, fast inverse square root, whitening                       */
/* ------------------------------------------------------------------------- */
static int cmp_float(const void *a, const void *b)
{
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

/* source/bpd_math.c:20-91: the quickselect returns the element of 0-based rank
 * (n-1)/2, i.e. the lower median; restated as sort-and-index. */
float or_lower_median(const float *a, int n)
{
    float *t = (float *)malloc(sizeof(float) * (n > 0 ? n : 1));
    memcpy(t, a, sizeof(float) * n);
    qsort(t, n, sizeof(float), cmp_float);
    float v = t[(n - 1) / 2];
    free(t);
    return v;
}

/* source/bpd_math.c:94-105: magic-constant initial guess + one Newton step,
 * evaluated left to right in float with separate roundings. */
float or_quake_rsqrt(float x)
{
    int32_t i;
    float y;
    memcpy(&i, &x, 4);
    i = 0x5f3759df - (i >> 1);
    memcpy(&y, &i, 4);
    float h = 0.5f * x;
    h = h * y;
    h = h * y;
    return y * (1.5f - h);
}

/* source/bpd_noise.c:12-92 (types kept: powers, mean_old, slope and norm are
 * float; the new block mean is double). */
void or_deredden(float *fft, unsigned nbins)
{
    const int initialbuflen = 6, maxbuflen = 200;
    int binnum = 1, numwrote = 1, lastbuflen, buflen, ind;
    long l;
    float dslope = 1.0f, norm, mean_old;
    const double ln2 = log(2.0);

    fft[0] = 1.0f;
    fft[1] = 0.0f;
    float *pow = (float *)malloc(sizeof(float) * (nbins ? nbins : 1));
    for (unsigned k = 0; k < nbins; k++) {
        float xr = fft[2 * k], xi = fft[2 * k + 1];
        float a = xr * xr, b = xi * xi;
        pow[k] = a + b;
    }
    buflen = initialbuflen;
    mean_old = (float)((double)or_lower_median(pow + binnum, buflen) / ln2);
    norm = or_quake_rsqrt(mean_old);
    for (ind = numwrote; ind < binnum + buflen / 2; ind++) {
        fft[2 * ind] *= norm;
        fft[2 * ind + 1] *= norm;
    }
    numwrote += buflen / 2;
    binnum += buflen;
    lastbuflen = buflen;
    buflen = (int)(initialbuflen * log((double)binnum));
    if (buflen > maxbuflen) buflen = maxbuflen;

    while ((unsigned)(binnum + buflen) < nbins) {
        double mean_new = (double)or_lower_median(pow + binnum, buflen) / ln2;
        dslope = (float)((mean_new - (double)mean_old) / (0.5 * (lastbuflen + buflen)));
        for (l = 0, ind = numwrote; ind < binnum + buflen / 2; l++, ind++) {
            float t = dslope * (float)l;
            norm = or_quake_rsqrt(mean_old + t);
            fft[2 * ind] *= norm;
            fft[2 * ind + 1] *= norm;
        }
        numwrote += (int)l;
        binnum += buflen;
        lastbuflen = buflen;
        mean_old = (float)mean_new;
        buflen = (int)(initialbuflen * log((double)binnum));
        if (buflen > maxbuflen) buflen = maxbuflen;
    }
    for (l = 0, ind = numwrote; (unsigned)ind < nbins; l++, ind++) {
        float t = dslope * (float)l;
        norm = or_quake_rsqrt(mean_old + t);
        fft[2 * ind] *= norm;
        fft[2 * ind + 1] *= norm;
    }
    free(pow);
}

/* ------------------------------------------------------------------------- */
/* per-step geometry and segments                                             */
/* ------------------------------------------------------------------------- */

/* source/bpd_search.cu:72-127 (sequential order) and :139-143 */
void or_step_geometry(int numharm, int zmax, int uselen, double startr, or_harm_geom *g)
{
    int cur[64], mx[64];
    or_harmonic_order(numharm, cur, mx);
    double lastr = startr + uselen * OR_DR - OR_DR;   /* AccelSearch.cu:211-212 */
    for (int h = 0; h < numharm; h++) {
        double f = (double)cur[h] / (double)mx[h];
        double drlo = or_calc_required_r(f, startr);
        double drhi = or_calc_required_r(f, lastr);
        int zmax_h = or_calc_required_z(f, zmax);
        g[h].cur = cur[h];
        g[h].max = mx[h];
        g[h].rlo = (int)floor(drlo);
        g[h].rhi = drhi;
        g[h].zlo = or_calc_required_z(f, -zmax);
        if (cur[h] == 1 && mx[h] == 1) {
            g[h].numrs = uselen;
        } else {
            int nr = (int)((ceil(drhi) - floor(drlo)) * OR_RDR + OR_DBLCORRECT) + 1;
            if (nr % OR_RDR) nr = (nr / OR_RDR + 1) * OR_RDR;
            g[h].numrs = nr;
        }
        g[h].numzs = (zmax_h / OR_DZ) * 2 + 1;
        g[h].fftlen = or_fft_length(uselen, mx[h], cur[h], zmax);
        g[h].binoffset = or_z_response_halfwidth((double)-zmax_h);
        g[h].lobin = g[h].rlo - g[h].binoffset;
        int hibin = (int)ceil(g[h].rhi) + g[h].binoffset;
        g[h].numdata = hibin - g[h].lobin + 1;
    }
}

/* AccelSearch.cu:209 -- steps while startr + U/2 < highestBin */
int or_num_steps(long highest, int rlo, int uselen)
{
    int n = 0;
    double startr = rlo;
    double inc = (double)(uselen / 2);
    while (startr + inc < (double)highest) { n++; startr += inc; }
    return n;
}

/* source/bpd_search.cu:16-47, :150-178 */
float or_segment_amplitudes(const float *spec, long nbins, int lobin, int numdata, int L,
                            float *out)
{
    long nice = or_next_pow2(numdata);
    long nfetch = nice > L / 2 ? nice : L / 2;
    if (nice < L / 2) abort();   /* the reference would read past its buffer */
    float *a = (float *)malloc(sizeof(float) * 2 * nfetch);
    for (long i = 0; i < nfetch; i++) {
        long b = (long)lobin + i;
        if (b >= 0 && b < nbins) { a[2 * i] = spec[2 * b]; a[2 * i + 1] = spec[2 * b + 1]; }
        else { a[2 * i] = 0.0f; a[2 * i + 1] = 0.0f; }
    }
    float *p = (float *)malloc(sizeof(float) * numdata);
    for (int i = 0; i < numdata; i++) {
        float xr = a[2 * i], xi = a[2 * i + 1];
        float u = xr * xr, v = xi * xi;
        p[i] = u + v;
    }
    float med = or_lower_median(p, numdata);
    float c = or_quake_rsqrt((float)((double)med / log(2.0)));
    for (int i = 0; i < numdata; i++) { a[2 * i] *= c; a[2 * i + 1] *= c; }
    memcpy(out, a, sizeof(float) * L);   /* L/2 complex */
    free(p);
    free(a);
    return med;
}

/* Y = FFT_L(spread(a)) * conj(K); y = IFFT_L(Y); P[i] = |y[off+i]|^2 / L^2
 * (source/bpd_search.cu:228-307, source/bpd_gpu_kernels.cu:39-116) */
void or_row_powers_fft(const float *amps, const double *K, int L, int off, int numrs,
                       double *out)
{
    double *s = (double *)calloc(2 * (size_t)L, sizeof(double));
    for (int i = 0; i < L / 2; i++) { s[4 * i] = amps[2 * i]; s[4 * i + 1] = amps[2 * i + 1]; }
    or_fft(s, L, -1);
    for (int k = 0; k < L; k++) {
        double xr = s[2 * k], xi = s[2 * k + 1], kr = K[2 * k], ki = K[2 * k + 1];
        s[2 * k] = xr * kr + xi * ki;
        s[2 * k + 1] = xi * kr - xr * ki;
    }
    or_fft(s, L, +1);
    double scale = 1.0 / ((double)L * (double)L);
    for (int i = 0; i < numrs; i++) {
        double yr = s[2 * (off + i)], yi = s[2 * (off + i) + 1];
        out[i] = (yr * yr + yi * yi) * scale;
    }
    free(s);
}

/* Direct form: y[n] = sum_t s[n+t] conj(tap[2hw+t]), t in [-2hw, 2hw). */
void or_row_powers_direct(const float *amps, int z, int L, int off, int numrs, double *out)
{
    int hw = or_z_response_halfwidth((double)z);
    int numkern = 4 * hw, half = 2 * hw;
    double *taps = (double *)malloc(sizeof(double) * 2 * numkern);
    or_z_response(0.0, OR_NUMBETWEEN, (double)z, numkern, taps);
    for (int q = 0; q < 2 * numkern; q++) taps[q] = (float)taps[q];
    for (int i = 0; i < numrs; i++) {
        long n = off + i;
        double yr = 0.0, yi = 0.0;
        for (int t = -half; t < half; t++) {
            long m = ((n + t) % L + L) % L;
            if (m & 1) continue;
            long j = m / 2;
            double ar = amps[2 * j], ai = amps[2 * j + 1];
            double kr = taps[2 * (half + t)], ki = taps[2 * (half + t) + 1];
            yr += ar * kr + ai * ki;
            yi += ai * kr - ar * ki;
        }
        out[i] = yr * yr + yi * yi;
    }
    free(taps);
}

/* ------------------------------------------------------------------------- */
/* whole search                                                               */
/* ------------------------------------------------------------------------- */
struct or_ctx {
    or_params p;
    int uselen, rlo, nsteps, numharm;
    long nbins;             /* N/2: bins handed to whitening and fetches */
    float *spec;            /* whitened spectrum, nbins complex           */
    int zmax_h[64], Z[64], L[64];
    double *K[64];          /* FFT'd template rows [Z][L] complex         */
};

or_ctx *or_create(const float *series, const or_params *p)
{
    or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
    c->p = *p;
    c->numharm = p->numharm;
    c->uselen = p->uselen ? p->uselen : or_max_uselen(p->fftlen, p->zmax);
    long n = p->n;
    /* AccelSearch.cu:101: rlo = floor(1.0 * (N * dt)) with the product in float */
    float nd = (float)n * p->dt;
    c->rlo = (int)floor(1.0 * nd);
    c->nbins = n / 2;   /* fftLength - 1 (AccelSearch.cu:91,93) */
    c->nsteps = or_num_steps(c->nbins, c->rlo, c->uselen);

    double *x = (double *)malloc(sizeof(double) * n);
    for (long i = 0; i < n; i++) x[i] = series[i];
    double *X = (double *)malloc(sizeof(double) * 2 * (n / 2 + 1));
    or_rfft(x, n, X);
    free(x);
    c->spec = (float *)malloc(sizeof(float) * 2 * (c->nbins > 0 ? c->nbins : 1));
    for (long k = 0; k < c->nbins; k++) {
        c->spec[2 * k] = (float)X[2 * k];
        c->spec[2 * k + 1] = (float)X[2 * k + 1];
    }
    free(X);
    or_deredden(c->spec, (unsigned)c->nbins);

    int cur[64], mx[64];
    or_harmonic_order(c->numharm, cur, mx);
    for (int h = 0; h < c->numharm; h++) {
        double f = (double)cur[h] / (double)mx[h];
        c->zmax_h[h] = or_calc_required_z(f, p->zmax);
        c->Z[h] = (c->zmax_h[h] / OR_DZ) * 2 + 1;
        c->L[h] = or_fft_length(c->uselen, mx[h], cur[h], p->zmax);
        int L = c->L[h];
        c->K[h] = (double *)malloc(sizeof(double) * 2 * (size_t)L * c->Z[h]);
#pragma omp parallel for schedule(dynamic)
        for (int k = 0; k < c->Z[h]; k++) {
            double *row = c->K[h] + 2 * (size_t)L * k;
            or_template_row(-c->zmax_h[h] + OR_DZ * k, L, row);
            or_fft(row, L, -1);
        }
    }
    return c;
}

void or_destroy(or_ctx *c)
{
    if (!c) return;
    for (int h = 0; h < c->numharm; h++) free(c->K[h]);
    free(c->spec);
    free(c);
}

int or_ctx_info(const or_ctx *c, int *uselen, int *rlo, int *nsteps, int *numz0, long *nbins)
{
    if (uselen) *uselen = c->uselen;
    if (rlo) *rlo = c->rlo;
    if (nsteps) *nsteps = c->nsteps;
    if (numz0) *numz0 = c->Z[0];
    if (nbins) *nbins = c->nbins;
    return 0;
}

const float *or_ctx_spectrum(const or_ctx *c) { return c->spec; }

/* One step of AccelSearch.cu:209-246: all harmonics, fundamental powers and the
 * harmonic sum of source/bpd_search.cu:310-360 / source/bpd_gpu_kernels.cu:119-170
 * (FP64 accumulation, reference order). stage_out receives the partial sums at the
 * fundamental and after harmonics 1, 3, 7, 15 (numharm 1, 2, 4, 8, 16). */
void or_step_plane(const or_ctx *c, int step, double *out, double *stage_out)
{
    int U = c->uselen, H = c->numharm;
    double startr = c->rlo + (double)step * (double)(U / 2);
    or_harm_geom g[64];
    or_step_geometry(H, c->p.zmax, U, startr, g);
    int Z0 = c->Z[0];
    int fullrlo = (int)startr;
    double *rowbuf = (double *)malloc(sizeof(double) * 65536);
    float *amps = (float *)malloc(sizeof(float) * 65536);

    /* fundamental */
    or_segment_amplitudes(c->spec, c->nbins, g[0].lobin, g[0].numdata, c->L[0], amps);
    int off0 = 2 * g[0].binoffset;
    for (int j = 0; j < Z0; j++)
        or_row_powers_fft(amps, c->K[0] + 2 * (size_t)c->L[0] * j, c->L[0], off0, U,
                          out + (size_t)j * U);
    int stage = 0;
    if (stage_out) memcpy(stage_out, out, sizeof(double) * (size_t)Z0 * U);

    for (int h = 1; h < H; h++) {
        int L = c->L[h], Zh = c->Z[h], nr = g[h].numrs, off = 2 * g[h].binoffset;
        or_segment_amplitudes(c->spec, c->nbins, g[h].lobin, g[h].numdata, L, amps);
        double *sub = (double *)malloc(sizeof(double) * (size_t)Zh * nr);
        for (int k = 0; k < Zh; k++)
            or_row_powers_fft(amps, c->K[h] + 2 * (size_t)L * k, L, off, nr, sub + (size_t)k * nr);
        double f = (double)g[h].cur / (double)g[h].max;
        for (int j = 0; j < Z0; j++) {
            int zz = g[0].zlo + j * OR_DZ;
            int subz = (int)(zz * f);
            int zind = (int)((subz - g[h].zlo) * OR_RDZ);
            for (int i = 0; i < U; i++) {
                double rr = fullrlo + i * OR_DR;
                double subr = (int)(OR_RDR * rr * f + 0.5) * OR_DR;
                unsigned short rind = (unsigned short)((subr - g[h].rlo) * OR_RDR);
                if (zind < 0 || zind >= Zh || rind >= nr) abort();
                out[(size_t)j * U + i] += sub[(size_t)zind * nr + rind];
            }
        }
        free(sub);
        if ((((h + 1) & h) == 0) && stage_out) {
            stage++;
            memcpy(stage_out + (size_t)stage * Z0 * U, out, sizeof(double) * (size_t)Z0 * U);
        }
    }
    free(rowbuf);
    free(amps);
}

/* Candidates of steps [step0, step0 + nsteps): every cell whose partial sum at a
 * stage exceeds powcut[stage] (a cut <= 0 disables its stage), as in the search
 * hook of source/bpd_search.cu:350-356; cells whose stage sum lies within
 * margin * max(P, 1) of a cut are returned flagged ambiguous instead, whichever
 * side they fall on.  Steps run in parallel (OpenMP); the list comes in step
 * order, then z, r, stage.  Returns the number of entries (at most cap are
 * written), or -1 if out of memory. */
long or_search_candidates(const or_ctx *c, int step0, int nsteps, const float *powcut,
                          double margin, or_cand *out, long cap)
{
    int U = c->uselen, Z0 = c->Z[0], stages = or_fast_log2(c->numharm) + 1;
    or_cand **per = (or_cand **)calloc((size_t)(nsteps > 0 ? nsteps : 1), sizeof(or_cand *));
    long *cnt = (long *)calloc((size_t)(nsteps > 0 ? nsteps : 1), sizeof(long));
    int fail = 0;
#pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < nsteps; k++) {
        size_t cells = (size_t)Z0 * U;
        double *plane = (double *)malloc(sizeof(double) * cells);
        double *st = (double *)malloc(sizeof(double) * cells * (size_t)stages);
        long n = 0, capk = 64;
        or_cand *lst = (or_cand *)malloc(sizeof(or_cand) * (size_t)capk);
        if (!plane || !st || !lst) {
#pragma omp atomic write
            fail = 1;
        } else {
            or_step_plane(c, step0 + k, plane, st);
            for (int j = 0; j < Z0; j++)
                for (int i = 0; i < U; i++)
                    for (int s = 0; s < stages; s++) {
                        double thr = powcut[s];
                        if (!(thr > 0.0)) continue;
                        double P = st[(size_t)s * cells + (size_t)j * U + i];
                        int amb = fabs(P - thr) <= margin * (P > 1.0 ? P : 1.0);
                        if (!amb && !(P > thr)) continue;
                        if (n == capk) {
                            capk *= 2;
                            or_cand *g = (or_cand *)realloc(lst, sizeof(or_cand) * (size_t)capk);
                            if (!g) {
#pragma omp atomic write
                                fail = 1;
                                capk /= 2;
                                n--;   /* the entry below overwrites the last one */
                            } else {
                                lst = g;
                            }
                        }
                        or_cand *e = &lst[n++];
                        e->step = step0 + k;
                        e->z_index = j;
                        e->r_index = i;
                        e->stage = s;
                        e->power = P;
                        e->ambiguous = amb;
                    }
        }
        free(plane);
        free(st);
        per[k] = lst;
        cnt[k] = n;
    }
    long total = 0;
    for (int k = 0; k < nsteps; k++) {
        for (long q = 0; q < cnt[k]; q++, total++)
            if (total < cap && out) out[total] = per[k][q];
        free(per[k]);
    }
    free(per);
    free(cnt);
    return fail ? -1 : total;
}

/* ------------------------------------------------------------------ optimiser
 * n-point Gauss-Legendre nodes on [-1, 1] by Newton iteration on P_n. */
static void or_gauss_legendre(int n, double *x, double *w)
{
    for (int i = 0; i < (n + 1) / 2; i++) {
        double t = cos(OR_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = t;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            pp = n * (t * p1 - p0) / (t * t - 1.0);
            double dt = p1 / pp;
            t -= dt;
            if (fabs(dt) < 1e-16) break;
        }
        x[i] = -t;
        x[n - 1 - i] = t;
        w[i] = w[n - 1 - i] = 2.0 / ((1.0 - t * t) * pp * pp);
    }
}

void or_rz_response(double q, double z, double *re, double *im)
{
    static double gx[16], gw[16];
    static int init = 0;
    if (!init) {
        or_gauss_legendre(16, gx, gw);
        init = 1;
    }
    /* phase rate 2 pi (q + z u): about one cycle per panel at most */
    double cyc = fabs(q) + fabs(z) + 1.0;
    int panels = (int)ceil(2.0 * cyc);
    double sr = 0.0, si = 0.0, h = 0.5 / panels;
    for (int p = 0; p < panels; p++) {
        double mid = (2 * p + 1) * h;
        for (int g = 0; g < 16; g++) {
            double u = mid + gx[g] * h;
            double ph = 2.0 * OR_PI * (q * u + 0.5 * z * u * u);
            sr += gw[g] * h * cos(ph);
            si += gw[g] * h * sin(ph);
        }
    }
    *re = sr;
    *im = si;
}

double or_rz_power(const float *spec, long nbins, double r, double z)
{
    int hw = or_z_response_halfwidth(z);
    long j0 = (long)ceil(r - hw), j1 = (long)ceil(r + hw) - 1;   /* [r - hw, r + hw) */
    double ar = 0.0, ai = 0.0;
    for (long j = j0; j <= j1; j++) {
        if (j < 0 || j >= nbins) continue;
        double rr, ri;
        or_rz_response(r - 0.5 * z - j, z, &rr, &ri);
        double xr = spec[2 * j], xi = spec[2 * j + 1];
        ar += xr * rr + xi * ri;
        ai += xi * rr - xr * ri;
    }
    return ar * ar + ai * ai;
}
