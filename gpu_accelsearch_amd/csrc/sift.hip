// sift.hip -- candidate significance, sifting, optimisation and properties.
//
// The reference declares these stages and leaves them empty: search_ffdot_powers,
// sortCandidates, eliminateHarmonics, optimizeCandidates, calculateProperties
// (headers/bpd_candidates.cuh:15-49, source/bpd_candidates.cu:4-33), and the
// per-stage powersCutoff / independentSpectraCount of AccelSearch.cu:37-38,95-97.
// Their semantics here follow the published accelsearch method (Ransom, Eikenberry
// & Middleditch 2002): a summed power of n harmonics is Gamma(n, 1) under noise,
// its significance is the equivalent Gaussian sigma over the number of independent
// trials of the stage, candidates are reported at the first harmonic (summed
// column / n), and optimisation maximises the summed power at continuous (r, z).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "fdas_internal.hpp"
#include "fresnel.hpp"

using namespace fdas;

namespace fdas {

namespace {

constexpr double kLog2Pi = 1.8378770664093453;   // log(2 pi)
constexpr double kLn2 = 0.69314718055994530942;

// log of the upper tail Q(n, x) of Gamma(n, 1) for integer n >= 1:
// Q = e^-x sum_{k<n} x^k / k!, summed in log space.
double log_gamma_tail(int n, double x)
{
    if (x <= 0.0) return 0.0;
    const double lx = std::log(x);
    double m = -INFINITY;
    std::vector<double> t(n);
    for (int k = 0; k < n; ++k) {
        t[k] = k * lx - std::lgamma((double)k + 1.0);
        m = std::max(m, t[k]);
    }
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += std::exp(t[k] - m);
    return -x + m + std::log(s);
}

// log of the Gaussian upper tail 1 - Phi(s).
double log_normal_tail(double s)
{
    if (s < 30.0) return std::log(0.5 * std::erfc(s / std::sqrt(2.0)));
    const double s2 = s * s;
    return -0.5 * s2 - std::log(s) - 0.5 * kLog2Pi +
           std::log1p(-1.0 / s2 + 3.0 / (s2 * s2) - 15.0 / (s2 * s2 * s2));
}

// sigma with log(1 - Phi(sigma)) = logp (Newton on the log tail).
double sigma_from_log_tail(double logp)
{
    if (logp >= 0.0) return -INFINITY;
    double s;
    if (logp > std::log(0.5)) {
        s = -1.0;   // p > 1/2: negative sigma, start below zero
    } else {
        const double a = -2.0 * logp;
        s = std::sqrt(std::max(a - std::log(std::max(a, 1.0)) - kLog2Pi, 0.1));
    }
    for (int it = 0; it < 100; ++it) {
        const double f = log_normal_tail(s) - logp;
        // d/ds log(1 - Phi(s)) = -phi(s) / (1 - Phi(s))
        const double d = -std::exp(-0.5 * s * s - 0.5 * kLog2Pi - log_normal_tail(s));
        const double step = f / d;
        s -= step;
        if (std::fabs(step) < 1e-12 * std::max(1.0, std::fabs(s))) break;
    }
    return s;
}

}  // namespace

double candidate_sigma(double power, int numharm, double numindep)
{
    if (numharm < 1) return 0.0;
    const double lp1 = log_gamma_tail(numharm, power);   // one trial
    const double N = std::max(numindep, 1.0);
    double lp;
    if (lp1 + std::log(N) < -20.0) {
        lp = lp1 + std::log(N);   // 1 - (1-p)^N = N p to relative 1e-9
    } else {
        const double p1 = std::exp(lp1);
        lp = std::log(-std::expm1(N * std::log1p(-std::min(p1, 1.0 - 1e-16))));
    }
    return sigma_from_log_tail(lp);
}

double power_for_sigma(double sigma, int numharm, double numindep)
{
    double lo = 0.0, hi = 1.0;
    while (candidate_sigma(hi, numharm, numindep) < sigma && hi < 1e7) hi *= 2.0;
    for (int it = 0; it < 200 && hi - lo > 1e-10 * hi; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (candidate_sigma(mid, numharm, numindep) < sigma) lo = mid;
        else hi = mid;
    }
    return hi;
}

// Independent trials of a stage summing 2^stage harmonics over Fourier bins
// [rlo, rhi) and numz drifts spaced dz = 2: the f-fdot cells decorrelate over
// about 6.95 z-bins, and summing 2^stage harmonics divides the distinct
// fundamental frequencies by 2^stage (the published accelsearch count).
double numindep_count(double rlo, double rhi, int numz, int stage)
{
    const double span = std::max(rhi - rlo, 1.0);
    const double d = (double)(1 << stage);
    if (numz <= 1) return span / d;
    return span * (numz + 1) * (DZ / 6.95) / d;
}

// ------------------------------------------------------------------ optimiser
namespace {

// R(q, z) = int_0^1 exp(2 pi i (q u + z u^2 / 2)) du in FP64.  |z| >= 0.1: the
// Fresnel form of the template taps; smaller |z|: Gauss-Legendre quadrature on
// 32 panels (the Fresnel difference loses digits as z -> 0).
__device__ void resp_d(double q, double z, double* re, double* im)
{
    const double kPi = 3.14159265358979323846;
    if (fabs(z) >= 0.1) {
        const int sz = z < 0 ? -1 : 1;
        const double zd = sz * 1.41421356237309504880 / sqrt(fabs(z));
        const double cons = zd / 2.0;
        const double yy = q * zd, zz = yy + z * zd;
        const double xx = kPi / z * q * q;
        double sx, cx, fsy, fcy, fsz, fcz;
        sincos(xx, &sx, &cx);
        fresnel_d(yy, &fsy, &fcy);
        fresnel_d(zz, &fsz, &fcz);
        const double tr = sz * (fcz - fcy), ti = fsy - fsz;
        *re = (tr * cx - ti * sx) * cons;
        *im = -(tr * sx + ti * cx) * cons;
        return;
    }
    // 8-point Gauss-Legendre nodes/weights on [-1, 1]
    const double xg[4] = {0.1834346424956498, 0.5255324099163290, 0.7966664774136267,
                          0.9602898564975363};
    const double wg[4] = {0.3626837833783620, 0.3137066458778873, 0.2223810344533745,
                          0.1012285362903763};
    constexpr int PANELS = 32;
    double sr = 0.0, si = 0.0;
    for (int p = 0; p < PANELS; ++p) {
        const double a = (double)p / PANELS, h = 0.5 / PANELS, mid = a + h;
        for (int g = 0; g < 8; ++g) {
            const double u = mid + (g < 4 ? -xg[g] : xg[g - 4]) * h;
            const double w = wg[g & 3] * h;
            double s, c;
            sincos(2.0 * kPi * (q * u + 0.5 * z * u * u), &s, &c);
            sr += w * c;
            si += w * s;
        }
    }
    *re = sr;
    *im = si;
}

__device__ int hw_of(double z)
{
    z = fabs(z);
    int m = (int)(z * (0.00089 * z + 0.3131) + NUMFINTBINS);
    m = (m < NUMFINTBINS) ? NUMFINTBINS : m;
    if (z > 100 && m > 0.6 * z) m = (int)(0.6 * z);
    return m;
}

struct OptTask {
    long long spec_off;   // trial spectrum in the plan's spectrum buffer
    int nbins;
    int numharm;
    double r, z;          // first-harmonic start point
    double dr, dz;        // initial grid spacing (first-harmonic units)
    float norm[MAX_HARM]; // 1 / (local median power / ln 2) per harmonic (the kernel's)
};

struct OptResult {
    double r, z;
    float power;
    float harm[MAX_HARM];
};

constexpr int OPT_G = 9;           // grid points per axis
constexpr int OPT_ITERS = 5;       // refinements (spacing / 3 each)

// Normalised power of harmonic k (1-based) at first-harmonic (r, z):
// |sum_j a_j conj(R(k r - k z / 2 - j, k z))|^2 with the search's tap convention
// (template tap i = R(numkern/4 - z/2 - i/2, z), source/bpd_subharmonics.cu:146-207).
__device__ float harm_power(const cplx* __restrict__ spec, int nbins, double r, double z, int k,
                            float norm)
{
    const double rk = k * r, zk = k * z;
    const int hw = hw_of(zk);
    // taps j in [rk - hw, rk + hw), as the search's template rows
    const int j0 = (int)ceil(rk - hw), j1 = (int)ceil(rk + hw) - 1;
    double ar = 0.0, ai = 0.0;
    for (int j = j0; j <= j1; ++j) {
        if (j < 0 || j >= nbins) continue;
        double rr, ri;
        resp_d(rk - 0.5 * zk - j, zk, &rr, &ri);
        const cplx x = spec[j];
        ar += x.x * rr + x.y * ri;   // x * conj(R)
        ai += x.y * rr - x.x * ri;
    }
    return (float)((ar * ar + ai * ai) * norm);
}

constexpr int OPT_WIN = 256;   // bins of the local power estimate per harmonic

// Local normalisation of harmonic k (1-based) of a candidate at first-harmonic
// r: ln 2 / (lower median of |a|^2 over the OPT_WIN bins around k r, clamped to
// the spectrum), 0 for a zero median.  Exact selection by rank counting in LDS.
__device__ float local_norm(const cplx* __restrict__ spec, int nbins, double r, int k, float* pw,
                            float* med, int t)
{
    const int wn = min(OPT_WIN, nbins);
    long lo = (long)floor(k * r) - wn / 2;
    lo = max(0L, min(lo, (long)nbins - wn));
    __syncthreads();   // earlier readers of pw are done
    if (t < wn) {
        const cplx a = spec[lo + t];
        pw[t] = a.x * a.x + a.y * a.y;
    }
    __syncthreads();
    const int rank = (wn - 1) / 2;
    if (t < wn) {
        const float v = pw[t];
        int lt = 0, le = 0;
        for (int q = 0; q < wn; ++q) {
            lt += pw[q] < v;
            le += pw[q] <= v;
        }
        if (lt <= rank && rank < le) *med = v;   // every such thread holds the same value
    }
    __syncthreads();
    const float m = *med;
    return m > 0.f ? (float)(kLn2 / (double)m) : 0.f;
}

__global__ __launch_bounds__(256) void optimize_kernel(const cplx* __restrict__ spec_all,
                                                       const OptTask* __restrict__ tasks,
                                                       OptResult* __restrict__ out)
{
    __shared__ float sum[OPT_G * OPT_G];
    __shared__ double cen[2];
    __shared__ int best;
    __shared__ float pw[OPT_WIN];
    __shared__ float med;
    __shared__ float norm[MAX_HARM];
    OptTask tk = tasks[blockIdx.x];
    const cplx* spec = spec_all + tk.spec_off;
    const int t = threadIdx.x;
    for (int k = 1; k <= tk.numharm; ++k) {
        const float nk = local_norm(spec, tk.nbins, tk.r, k, pw, &med, t);
        if (t == 0) norm[k - 1] = nk;
    }
    __syncthreads();
    for (int k = 0; k < tk.numharm; ++k) tk.norm[k] = norm[k];
    if (t == 0) {
        cen[0] = tk.r;
        cen[1] = tk.z;
    }
    double dr = tk.dr, dz = tk.dz;
    for (int it = 0; it < OPT_ITERS; ++it) {
        for (int i = t; i < OPT_G * OPT_G; i += blockDim.x) sum[i] = 0.f;
        __syncthreads();
        const int work = OPT_G * OPT_G * tk.numharm;
        for (int w = t; w < work; w += blockDim.x) {
            const int p = w % (OPT_G * OPT_G), k = w / (OPT_G * OPT_G) + 1;
            const double r = cen[0] + (p % OPT_G - OPT_G / 2) * dr;
            const double z = cen[1] + (p / OPT_G - OPT_G / 2) * dz;
            atomicAdd(&sum[p], harm_power(spec, tk.nbins, r, z, k, tk.norm[k - 1]));
        }
        __syncthreads();
        if (t == 0) {
            int b = 0;
            for (int p = 1; p < OPT_G * OPT_G; ++p)
                if (sum[p] > sum[b]) b = p;
            best = b;
            cen[0] += (b % OPT_G - OPT_G / 2) * dr;
            cen[1] += (b / OPT_G - OPT_G / 2) * dz;
        }
        __syncthreads();
        dr /= 3.0;
        dz /= 3.0;
    }
    // final powers per harmonic at the optimum
    if (t < tk.numharm) {
        out[blockIdx.x].harm[t] = harm_power(spec, tk.nbins, cen[0], cen[1], t + 1, tk.norm[t]);
    }
    if (t == 0) {
        out[blockIdx.x].r = cen[0];
        out[blockIdx.x].z = cen[1];
        out[blockIdx.x].power = sum[best];
    }
}

}  // namespace

int optimize_candidates(const cplx* d_spec, const std::vector<long long>& spec_off,
                        const std::vector<int>& nbins, fdas_accelcand* c, int n, hipStream_t s)
{
    if (n <= 0) return FDAS_OK;
    std::vector<OptTask> tasks(n);
    for (int i = 0; i < n; ++i) {
        OptTask& tk = tasks[i];
        const int tr = c[i].trial;
        tk.spec_off = spec_off[tr];
        tk.nbins = nbins[tr];
        tk.numharm = c[i].numharm;
        tk.r = c[i].r;
        tk.z = c[i].z;
        // one search-grid cell of the summed column, in first-harmonic units
        tk.dr = 0.5 * DR / c[i].numharm;
        tk.dz = 0.5 * DZ / c[i].numharm;
        for (int k = 0; k < MAX_HARM; ++k) tk.norm[k] = 0.f;   // set on the device
    }
    OptTask* d_tasks = nullptr;
    OptResult* d_out = nullptr;
    int rc = FDAS_OK;
    if (hipMalloc(&d_tasks, n * sizeof(OptTask)) != hipSuccess ||
        hipMalloc(&d_out, n * sizeof(OptResult)) != hipSuccess) {
        set_error("optimiser: device allocation failed");
        rc = FDAS_ERR_NOMEM;
    }
    std::vector<OptResult> res(n);
    if (rc == FDAS_OK &&
        hipMemcpyAsync(d_tasks, tasks.data(), n * sizeof(OptTask), hipMemcpyHostToDevice, s) != hipSuccess)
        rc = FDAS_ERR_HIP;
    if (rc == FDAS_OK) {
        hipLaunchKernelGGL(optimize_kernel, dim3(n), dim3(256), 0, s, d_spec, d_tasks, d_out);
        rc = check_launch("optimize_kernel");
    }
    if (rc == FDAS_OK &&
        (hipMemcpyAsync(res.data(), d_out, n * sizeof(OptResult), hipMemcpyDeviceToHost, s) != hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess)) {
        set_error("optimiser: copy back failed");
        rc = FDAS_ERR_HIP;
    }
    (void)hipFree(d_tasks);
    (void)hipFree(d_out);
    FDAS_TRY(rc);
    for (int i = 0; i < n; ++i) {
        c[i].opt_r = res[i].r;
        c[i].opt_z = res[i].z;
        c[i].opt_power = res[i].power;
        for (int k = 0; k < MAX_HARM; ++k) c[i].harm_power[k] = k < c[i].numharm ? res[i].harm[k] : 0.f;
    }
    return FDAS_OK;
}

namespace {

__global__ __launch_bounds__(256) void rz_power_kernel(const cplx* __restrict__ spec, int nbins,
                                                       const double* __restrict__ r,
                                                       const double* __restrict__ z, int count,
                                                       float* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = harm_power(spec, nbins, r[i], z[i], 1, 1.0f);
}

int gcd(int a, int b) { return b ? gcd(b, a % b) : a; }

}  // namespace

int rz_powers(const cplx* d_spec, int nbins, const double* r, const double* z, int count,
              float* out, hipStream_t s)
{
    if (count <= 0) return FDAS_OK;
    double *d_r = nullptr, *d_z = nullptr;
    float* d_o = nullptr;
    int rc = FDAS_OK;
    if (hipMalloc(&d_r, count * sizeof(double)) != hipSuccess ||
        hipMalloc(&d_z, count * sizeof(double)) != hipSuccess ||
        hipMalloc(&d_o, count * sizeof(float)) != hipSuccess) {
        set_error("fdas_rz_power: device allocation failed");
        rc = FDAS_ERR_NOMEM;
    }
    if (rc == FDAS_OK &&
        (hipMemcpyAsync(d_r, r, count * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
         hipMemcpyAsync(d_z, z, count * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess))
        rc = FDAS_ERR_HIP;
    if (rc == FDAS_OK) {
        hipLaunchKernelGGL(rz_power_kernel, dim3((count + 255) / 256), dim3(256), 0, s, d_spec, nbins,
                           d_r, d_z, count, d_o);
        rc = check_launch("rz_power_kernel");
    }
    if (rc == FDAS_OK &&
        (hipMemcpyAsync(out, d_o, count * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess))
        rc = FDAS_ERR_HIP;
    (void)hipStreamSynchronize(s);
    (void)hipFree(d_r);
    (void)hipFree(d_z);
    (void)hipFree(d_o);
    return rc;
}

void cand_properties(fdas_accelcand* c, double T)
{
    const double r = c->opt_r, z = c->opt_z;
    c->freq = r / T;
    c->fdot = z / (T * T);
    c->period = c->freq != 0.0 ? 1.0 / c->freq : 0.0;
    c->pdot = c->freq != 0.0 ? -c->fdot / (c->freq * c->freq) : 0.0;
}

int sift_candidates(const fdas_sift_trial* trials, int ntrials, int zmax, int uselen, int numz,
                    const fdas_candidate* in, int n_in, const fdas_sift_params* prm,
                    fdas_accelcand* out, int cap, int* n_out)
{
    fdas_sift_params P{};
    if (prm) P = *prm;
    const double rtol = P.r_tol > 0 ? P.r_tol : 1.1;
    const double ztol = P.z_tol > 0 ? P.z_tol : 4.0;
    const int maxr = P.max_ratio == 0 ? 16 : P.max_ratio;
    std::vector<fdas_accelcand> all;
    all.reserve(n_in);
    for (int i = 0; i < n_in; ++i) {
        const fdas_candidate& c = in[i];
        if (c.trial < 0 || c.trial >= ntrials || c.stage < 0 || c.stage > 4) {
            set_error("fdas_sift: candidate %d has trial %d / stage %d out of range", i, c.trial, c.stage);
            return FDAS_ERR_INVALID;
        }
        const fdas_sift_trial& t = trials[c.trial];
        fdas_accelcand a;
        memset(&a, 0, sizeof(a));
        a.trial = c.trial;
        a.step = c.step;
        a.z_index = c.z_index;
        a.r_index = c.r_index;
        a.stage = c.stage;
        a.numharm = 1 << c.stage;
        a.power = c.power;
        const double rcol = t.rlo + (double)c.step * (uselen / 2) + DR * c.r_index;
        const double zcol = -zmax + DZ * (double)c.z_index;
        a.r = rcol / a.numharm;
        a.z = zcol / a.numharm;
        a.sigma = (float)candidate_sigma(c.power, a.numharm, numindep_count(t.rlo, (double)t.nbins, numz, c.stage));
        if (a.sigma < P.sigma_min) continue;
        a.opt_r = a.r;
        a.opt_z = a.z;
        a.opt_power = a.power;
        a.opt_sigma = a.sigma;
        cand_properties(&a, (double)t.n * t.dt);
        all.push_back(a);
    }
    std::stable_sort(all.begin(), all.end(), [](const fdas_accelcand& a, const fdas_accelcand& b) {
        if (a.sigma != b.sigma) return a.sigma > b.sigma;
        if (a.power != b.power) return a.power > b.power;
        if (a.trial != b.trial) return a.trial < b.trial;
        return a.r < b.r;
    });
    // keep[trial]: kept candidates ordered by first-harmonic r
    std::vector<std::multimap<double, int>> keep(ntrials);
    std::vector<fdas_accelcand> kept;
    auto near = [&](int trial, double r0, double z0, double rt, double zt) {
        auto& m = keep[trial];
        for (auto it = m.lower_bound(r0 - rt); it != m.end() && it->first <= r0 + rt; ++it)
            if (std::fabs(kept[it->second].z - z0) <= zt) return true;
        return false;
    };
    for (const fdas_accelcand& a : all) {
        bool drop = near(a.trial, a.r, a.z, rtol, ztol);   // same signal, neighbouring cell or stage
        // a is the n/d harmonic of a stronger kept candidate k: a.r = (n/d) k.r
        for (int d = 1; !drop && maxr > 0 && d <= maxr; ++d) {
            for (int n = 1; n <= maxr && !drop; ++n) {
                if (n == d || gcd(n, d) != 1) continue;
                const double q = (double)n / d;
                drop = near(a.trial, a.r / q, a.z / q, rtol * std::max(1.0, 1.0 / q),
                            ztol * std::max(1.0, 1.0 / q));
            }
        }
        if (drop) continue;
        keep[a.trial].insert({a.r, (int)kept.size()});
        kept.push_back(a);
    }
    *n_out = (int)kept.size();
    if (out && cap > 0) memcpy(out, kept.data(), sizeof(fdas_accelcand) * std::min<size_t>(cap, kept.size()));
    return FDAS_OK;
}

}  // namespace fdas

// ====================================================================== C ABI
extern "C" double fdas_candidate_sigma(double power, int numharm, double numindep)
{
    return candidate_sigma(power, numharm, numindep);
}

extern "C" double fdas_power_for_sigma(double sigma, int numharm, double numindep)
{
    return power_for_sigma(sigma, numharm, numindep);
}

extern "C" double fdas_numindep(double rlo, double rhi, int numz, int stage)
{
    return numindep_count(rlo, rhi, numz, stage);
}

extern "C" int fdas_sift_raw(const fdas_sift_trial* trials, int ntrials, int zmax, int uselen,
                             int numz, const fdas_candidate* in, int n_in,
                             const fdas_sift_params* params, fdas_accelcand* out, int capacity,
                             int* n_out)
{
    if (!trials || ntrials < 1 || (!in && n_in > 0) || n_in < 0 || !n_out || uselen <= 0) {
        set_error("fdas_sift_raw: bad argument");
        return FDAS_ERR_INVALID;
    }
    return sift_candidates(trials, ntrials, zmax, uselen, numz, in, n_in, params, out, capacity, n_out);
}
