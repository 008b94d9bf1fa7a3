// plan.hip -- the plan API of include/fdas.h: a batch of DM trials (one or many,
// of any even lengths) searched with the same zmax / numharm / fftlen.
//
// Every (trial, step) pair of the batch is an "item"; the overlap-save loop of
// AccelSearch.cu:203-247 runs over items in chunks, each chunk being one launch
// per FFT-length class (segment preparation), one per sub-harmonic and one for
// the fundamental (fused correlation + inverse FFT + power), and one harmonic-sum
// launch.  Template banks are shared by every plan of the same configuration
// (a process-wide registry), so a server keeps one resident copy per zmax.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "fdas_internal.hpp"
#include "mma_layout.hpp"

using namespace fdas;

namespace fdas {

// ------------------------------------------------------------------ template banks
struct BankSet {
    int H = 0;
    int L[MAX_HARM]{}, Z[MAX_HARM]{};
    void* bankq[MAX_HARM]{};      // stored form used by the kernels
    int* bank_exp[MAX_HARM]{};    // per-row exponents (reduced precision)
    void* splitq[MAX_HARM]{};     // split-form rows [E | O] for the split correlation
    int* split_exp[MAX_HARM]{};
    MmaBank mma[MAX_HARM]{};      // direct-correlation banks (FDAS_PREC_I16_MMA)
    size_t bytes = 0;
    ~BankSet()
    {
        for (int h = 0; h < MAX_HARM; ++h) {
            free_mma_bank(&mma[h]);
            (void)hipFree(bankq[h]);
            (void)hipFree(bank_exp[h]);
            (void)hipFree(splitq[h]);
            (void)hipFree(split_exp[h]);
        }
    }
};

// Stored form of one FP32 bank: itself (FP32) or a packed copy; frees the FP32
// bank when packing.
static int store_bank(cplx* bank, int Z, int L, int prec, hipStream_t s, void** out, int** exps,
                      size_t* bytes)
{
    if (prec == FDAS_PREC_FP32) {
        *out = bank;
        *bytes += (size_t)Z * L * sizeof(cplx);
        return FDAS_OK;
    }
    void* q = nullptr;
    int* e = nullptr;
    int rc = FDAS_OK;
    if (hipMalloc(&q, (size_t)Z * L * 4) != hipSuccess || hipMalloc(&e, (size_t)Z * sizeof(int)) != hipSuccess) {
        set_error("template bank allocation failed");
        rc = FDAS_ERR_NOMEM;
    } else {
        rc = launch_bank_pack(bank, Z, L, prec, q, e, s);
    }
    if (hipStreamSynchronize(s) != hipSuccess && rc == FDAS_OK) rc = FDAS_ERR_HIP;
    (void)hipFree(bank);
    *out = q;
    *exps = e;
    *bytes += (size_t)Z * L * 4 + (size_t)Z * sizeof(int);
    return rc;
}

// device, zmax, numharm, uselen, prec, length mode, row layouts (log2 of each
// harmonic's wavefront decimation, 4 bits per harmonic)
using BankKey = std::tuple<int, int, int, int, int, int, long long>;
static std::mutex g_bank_mu;
static std::map<BankKey, std::weak_ptr<BankSet>> g_banks;

// Banks of harmonic h have length Ls[h] (a power of two, or R*2^k stored in the
// decimated order); the lengths follow from (zmax, H, uselen, mode).  Rows of a
// harmonic with decim[h] > 0 are stored in the wavefront-decimated order of the
// wavefront-split correlation (the split-form rows stay in natural order).
static int build_banks(int zmax, int H, const int* cur, const int* mx, const int* Ls, int uselen,
                       int prec, int mode, const int* decim, hipStream_t s,
                       std::shared_ptr<BankSet>* out)
{
    int dev = 0;
    FDAS_HIP(hipGetDevice(&dev));
    long long layout = 0;
    for (int h = 0; h < H; ++h) layout |= (long long)decim[h] << (4 * h);
    const BankKey key{dev, zmax, H, uselen, prec, mode, layout};
    std::lock_guard<std::mutex> lk(g_bank_mu);
    auto it = g_banks.find(key);
    if (it != g_banks.end()) {
        if (auto sp = it->second.lock()) {
            *out = sp;
            return FDAS_OK;
        }
    }
    auto bs = std::make_shared<BankSet>();
    bs->H = H;
    for (int h = 0; h < H; ++h) {
        const double f = (double)cur[h] / (double)mx[h];
        const int zmax_h = h_calc_required_z(f, zmax);
        const int Z = (zmax_h / DZ) * 2 + 1, L = Ls[h];
        bs->Z[h] = Z;
        bs->L[h] = L;
        if (prec == FDAS_PREC_I16_MMA) {   // time-domain tiles, no FFT'd rows
            FDAS_TRY(build_mma_bank(zmax_h, Z, h_z_response_halfwidth((double)zmax_h), s, &bs->mma[h]));
            bs->bytes += bs->mma[h].bytes;
            continue;
        }
        cplx *bank = nullptr, *split = nullptr;
        FDAS_HIP(hipMalloc(&bank, (size_t)Z * L * sizeof(cplx)));
        FDAS_TRY(build_template_bank(bank, zmax_h, Z, L, s));
        if (odd_factor(L) > 1) {   // no split form at lengths R*2^k
            const int r1 = store_bank(bank, Z, L, prec, s, &bs->bankq[h], &bs->bank_exp[h], &bs->bytes);
            FDAS_TRY(r1);
            continue;
        }
        if (hipMalloc(&split, (size_t)Z * L * sizeof(cplx)) != hipSuccess) {
            (void)hipFree(bank);
            set_error("template bank allocation failed");
            return FDAS_ERR_NOMEM;
        }
        const int rs = launch_bank_split(bank, Z, L, split, s);
        if (rs != FDAS_OK || hipStreamSynchronize(s) != hipSuccess) {
            (void)hipFree(bank);
            (void)hipFree(split);
            return rs != FDAS_OK ? rs : FDAS_ERR_HIP;
        }
        if (decim[h] > 0) {
            cplx* dec = nullptr;
            int rd = hipMalloc(&dec, (size_t)Z * L * sizeof(cplx)) == hipSuccess ? FDAS_OK : FDAS_ERR_NOMEM;
            if (rd == FDAS_OK) rd = decimate_rows(bank, dec, L, decim[h], Z, s);
            if (rd == FDAS_OK && hipStreamSynchronize(s) != hipSuccess) rd = FDAS_ERR_HIP;
            (void)hipFree(bank);
            bank = dec;
            if (rd != FDAS_OK) {
                (void)hipFree(dec);
                (void)hipFree(split);
                if (rd == FDAS_ERR_NOMEM) set_error("template bank allocation failed");
                return rd;
            }
        }
        const int r1 = store_bank(bank, Z, L, prec, s, &bs->bankq[h], &bs->bank_exp[h], &bs->bytes);
        const int r2 = store_bank(split, Z, L, prec, s, &bs->splitq[h], &bs->split_exp[h], &bs->bytes);
        FDAS_TRY(r1);
        FDAS_TRY(r2);
    }
    FDAS_HIP(hipStreamSynchronize(s));
    g_banks[key] = bs;
    *out = bs;
    return FDAS_OK;
}

// ------------------------------------------------------------------ small kernels
__global__ void zero_int_kernel(int* p) { *p = 0; }

}  // namespace fdas

// ------------------------------------------------------------------ the plan
struct TrialInfo {
    long n = 0, nbins = 0;
    int rlo = 0, nsteps = 0, item0 = 0;
    long long spec_off = 0;       // offset of its spectrum in d_spec (complex)
    long long series_off = 0;     // offset of its input in d_series (floats)
    int form = 0;                 // index of its transform + whitening schedule
};

// R2C plan and whitening schedule of one series length (shared by its trials).
struct LengthForm {
    long n = 0;
    R2CPlan r2c;
    WhitenSchedule ws;
};

struct fdas_plan {
    fdas_params p{};
    int uselen = 0, H = 0, Z0 = 0, chunk = 0, nitems = 0;
    std::vector<TrialInfo> trials;
    std::vector<LengthForm> forms;
    int cur[MAX_HARM]{}, mx[MAX_HARM]{}, L[MAX_HARM]{}, Z[MAX_HARM]{}, zmax_h[MAX_HARM]{},
        off[MAX_HARM]{}, W[MAX_HARM]{}, zlo[MAX_HARM]{};
    std::shared_ptr<BankSet> banks;
    int prec = FDAS_PREC_FP32;
    bool mma = false;               // direct correlation (FDAS_PREC_I16_MMA)
    int nm[MAX_HARM]{}, nch[MAX_HARM]{};   // its M tiles and staged chunks per harmonic
    int mma_var = 0;                // its launch shape
    size_t cbytes = sizeof(cplx);   // per segment element (mma: a 16-byte chunk)
    long long seg_off[MAX_HARM]{};
    long long seg_stride = 0;
    long long sub_total = 0;
    long long sub_cells[MAX_HARM]{};
    long long cells = 0;
    std::vector<int> classes;
    std::vector<std::vector<int>> class_harm;
    std::vector<SegDesc*> d_desc;   // per class: [item][harmonics in class]
    int* d_fullrlo = nullptr;       // [item]
    int* d_rlo_h = nullptr;         // [item][MAX_HARM]
    float* d_series = nullptr;      // plan-owned inputs (batches), sum of n floats
    cplx* d_spec = nullptr;         // sum of (n/2+1) complex
    cplx* d_r2c_work = nullptr;
    void* d_seg = nullptr;
    int* d_seg_exp = nullptr;
    void* d_sub = nullptr;          // sub-harmonic power planes (plane element of prec)
    void* d_plane = nullptr;        // fundamental power plane (same element)
    size_t pbytes = sizeof(float);  // plane element: FP32, or P16 in reduced precision
    fdas_candidate* d_cands = nullptr;
    int* d_ncand = nullptr;
    int cand_cap = 0;
    HarmSumArgs hs{};
    // launch shapes in effect when the plan was made (process-wide choices made
    // later, by the autotuner or the setters, apply to later plans only)
    int corr_var[MAX_HARM]{};
    int decim[MAX_HARM]{};          // log2 wavefront decimation of segments and rows
    std::vector<int> class_decim;   // the same per FFT-length class
    int hs_var = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    DevAlloc* al = nullptr;         // source of the plan-owned buffers
    size_t bytes = 0;
    // hipGraph of a whole batch run
    bool use_graph = false;
    hipGraphExec_t graph = nullptr;
    // event timing
    bool timing = false;
    struct Rec {
        int cat;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    size_t nrec = 0;
};

namespace {

// Brackets one launch with events when timing is on.
struct TimedLaunch {
    fdas_plan* pl;
    hipStream_t st;   // the stream the bracketed launch goes to
    size_t idx = (size_t)-1;
    TimedLaunch(fdas_plan* p, int cat) : pl(p), st(p->stream)
    {
        if (!pl->timing) return;
        if (pl->nrec == pl->recs.size()) {
            fdas_plan::Rec r{cat, nullptr, nullptr};
            if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return;
            pl->recs.push_back(r);
        }
        idx = pl->nrec++;
        pl->recs[idx].cat = cat;
        (void)hipEventRecord(pl->recs[idx].a, st);
    }
    ~TimedLaunch()
    {
        if (idx != (size_t)-1) (void)hipEventRecord(pl->recs[idx].b, st);
    }
};

int plan_alloc(fdas_plan* pl, void** ptr, size_t bytes)
{
    FDAS_TRY(pl->al->alloc(ptr, bytes));
    pl->bytes += bytes;
    return FDAS_OK;
}

int plan_upload(fdas_plan* pl, void* dst, const void* src, size_t bytes)
{
    FDAS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, pl->stream));
    return FDAS_OK;
}

int validate(const fdas_params& p)
{
    if (p.zmax < 0 || p.zmax > 1200 || (p.zmax & 1)) {
        set_error("zmax %d must be even and in [0, 1200]", p.zmax);
        return FDAS_ERR_INVALID;
    }
    if (p.numharm < 1 || p.numharm > 16 || (p.numharm & (p.numharm - 1))) {
        set_error("numharm %d must be 1, 2, 4, 8 or 16", p.numharm);
        return FDAS_ERR_INVALID;
    }
    if (p.fftlen < 1024 || p.fftlen > 16384 || (p.fftlen & (p.fftlen - 1))) {
        set_error("fftlen %d must be a power of two in [1024, 16384]", p.fftlen);
        return FDAS_ERR_UNSUPPORTED;
    }
    if (p.precision < FDAS_PREC_FP32 || p.precision > FDAS_PREC_I16_MMA) {
        set_error("precision mode %d unknown", p.precision);
        return FDAS_ERR_INVALID;
    }
    if (p.fft_lengths != FDAS_LENGTHS_SHORTEST && p.fft_lengths != FDAS_LENGTHS_REFERENCE) {
        set_error("FFT length mode %d unknown", p.fft_lengths);
        return FDAS_ERR_INVALID;
    }
    return FDAS_OK;
}

// Host geometry of every item (source/bpd_search.cu:72-127, :139-143) and the
// checks that the kernels' index arithmetic stays inside the planes.
int plan_geometry(fdas_plan* pl, const long* lengths, int ntrials)
{
    if (pl->p.fftlen == 0) pl->p.fftlen = resolve_fftlen(pl->p);
    const fdas_params& p = pl->p;
    FDAS_TRY(validate(p));
    pl->prec = p.precision;
    pl->mma = p.precision == FDAS_PREC_I16_MMA;
    pl->cbytes = p.precision == FDAS_PREC_FP32 ? sizeof(cplx) : pl->mma ? 16 : 4;
    pl->pbytes = p.precision == FDAS_PREC_FP32 ? sizeof(float) : sizeof(unsigned short);
    pl->uselen = p.uselen ? p.uselen : h_max_uselen(p.fftlen, p.zmax);
    if (pl->uselen <= 0 || (pl->uselen & 1) || pl->uselen >= 65536 ||
        h_fft_length(pl->uselen, 1, 1, p.zmax) != p.fftlen) {
        set_error("uselen %d does not give fundamental FFT length %d at zmax %d", pl->uselen,
                  p.fftlen, p.zmax);
        return FDAS_ERR_INVALID;
    }
    const int U = pl->uselen;
    const double half = (double)(U / 2);
    pl->H = p.numharm;
    h_harmonic_order(pl->H, pl->cur, pl->mx);
    long long so = 0;
    for (int h = 0; h < pl->H; ++h) {
        const double f = (double)pl->cur[h] / (double)pl->mx[h];
        pl->zmax_h[h] = h_calc_required_z(f, p.zmax);
        pl->Z[h] = (pl->zmax_h[h] / DZ) * 2 + 1;
        // the direct correlation's cost does not depend on the FFT length: it keeps
        // the reference's powers of two for the segment geometry
        pl->L[h] = h_harmonic_length(U, pl->mx[h], pl->cur[h], p.zmax,
                                     pl->mma ? FDAS_LENGTHS_REFERENCE : p.fft_lengths);
        pl->off[h] = NUMBETWEEN * h_z_response_halfwidth((double)-pl->zmax_h[h]);
        if (pl->mma) {
            int jlo = 0, nk = 0;
            mma_tile_range(pl->off[h] / NUMBETWEEN, pl->off[h] / NUMBETWEEN, &jlo, &nk);
            if (nk > kMmaMaxK) {
                set_error("direct correlation: zmax %d of harmonic %d/%d needs %d K steps (at most "
                          "%d: zmax up to 300)", pl->zmax_h[h], pl->cur[h], pl->mx[h], nk, kMmaMaxK);
                return FDAS_ERR_UNSUPPORTED;
            }
        }
        pl->zlo[h] = h_calc_required_z(f, -p.zmax);
        pl->seg_off[h] = so;
        so += pl->L[h] / 2;
        if (pl->L[h] < 256 || pl->L[h] > 16384) {
            set_error("harmonic %d/%d needs FFT length %d (fused kernels: 256..16384)",
                      pl->cur[h], pl->mx[h], pl->L[h]);
            return FDAS_ERR_UNSUPPORTED;
        }
    }
    pl->seg_stride = so;
    pl->Z0 = pl->Z[0];
    for (int h = 1; h < pl->H; ++h) {
        // rows the harmonic sum reads (monotone in the fundamental row)
        const int c = pl->cur[h], lm = h_fast_log2(pl->mx[h]);
        for (int j : {0, pl->Z0 - 1}) {
            const int zz = pl->zlo[0] + DZ * j;
            const int x = zz * c;
            const int subz = x >= 0 ? (x >> lm) : -((-x) >> lm);
            const int zind = (subz - pl->zlo[h]) >> 1;
            if (zind < 0 || zind >= pl->Z[h]) {
                set_error("harmonic-sum row %d outside [0,%d) for harmonic %d", zind, pl->Z[h], h);
                return FDAS_ERR_INVALID;
            }
        }
    }

    pl->trials.resize(ntrials);
    long long spec = 0, series = 0;
    int items = 0;
    for (int t = 0; t < ntrials; ++t) {
        const long n = lengths[t];
        if (n < 256 || (n & 1) || n > (1L << 27)) {
            set_error("trial %d: series length %ld: even and in [256, 2^27] required", t, n);
            return FDAS_ERR_UNSUPPORTED;
        }
        TrialInfo& tr = pl->trials[t];
        tr.n = n;
        tr.nbins = n / 2;
        // AccelSearch.cu:101: rlo = floor(1.0 * (N * dt)) with a float product
        const float nd = (float)n * p.dt;
        tr.rlo = (int)std::floor(1.0 * nd);
        int ns = 0;
        for (double st = tr.rlo; st + half < (double)tr.nbins; st += half) ++ns;
        if (ns <= 0) {
            set_error("trial %d: no search steps (rlo %d + uselen/2 >= %ld bins)", t, tr.rlo,
                      tr.nbins);
            return FDAS_ERR_INVALID;
        }
        tr.nsteps = ns;
        tr.item0 = items;
        tr.spec_off = spec;
        tr.series_off = series;
        items += ns;
        spec += tr.nbins + 1;
        series += n;
    }
    pl->nitems = items;
    pl->cells = (long long)items * U * pl->Z0;
    return FDAS_OK;
}

int plan_tables(fdas_plan* pl, std::vector<int>& fullrlo, std::vector<int>& rlo_h,
                std::vector<std::vector<SegDesc>>& desc)
{
    const int U = pl->uselen, H = pl->H, S = pl->nitems;
    const double half = (double)(U / 2);
    fullrlo.assign(S, 0);
    rlo_h.assign((size_t)S * MAX_HARM, 0);
    for (int h = 0; h < H; ++h) {
        pl->W[h] = (h == 0) ? U : 0;
        pl->sub_cells[h] = 0;
    }
    for (int h = 0; h < H; ++h) {
        auto it = std::find(pl->classes.begin(), pl->classes.end(), pl->L[h]);
        if (it == pl->classes.end()) {
            pl->classes.push_back(pl->L[h]);
            pl->class_harm.push_back({h});
        } else {
            pl->class_harm[it - pl->classes.begin()].push_back(h);
        }
    }
    desc.assign(pl->classes.size(), {});
    for (size_t c = 0; c < pl->classes.size(); ++c) desc[c].resize((size_t)S * pl->class_harm[c].size());
    for (const TrialInfo& tr : pl->trials) {
        for (int k = 0; k < tr.nsteps; ++k) {
            const int item = tr.item0 + k;
            const double startr = tr.rlo + (double)k * half;
            const double lastr = startr + half - DR;
            fullrlo[item] = (int)startr;
            for (int h = 0; h < H; ++h) {
                const double f = (double)pl->cur[h] / (double)pl->mx[h];
                const double drlo = h_calc_required_r(f, startr);
                const double drhi = h_calc_required_r(f, lastr);
                const int rl = (int)std::floor(drlo);
                const int bo = pl->off[h] / NUMBETWEEN;
                int nr = U;
                if (h > 0) {
                    nr = (int)((std::ceil(drhi) - std::floor(drlo)) * RDR + DBLCORRECT) + 1;
                    if (nr % RDR) nr = (nr / RDR + 1) * RDR;
                    // columns the harmonic sum reads at the first and last fundamental column
                    const double s0 = (int)(RDR * (fullrlo[item] + 0 * DR) * f + 0.5) * DR;
                    const double s1 = (int)(RDR * (fullrlo[item] + (U - 1) * DR) * f + 0.5) * DR;
                    const int r0 = (int)((s0 - rl) * RDR), r1 = (int)((s1 - rl) * RDR);
                    if (r0 < 0 || r1 >= nr) {
                        set_error("harmonic-sum index %d..%d outside [0,%d) at step %d", r0, r1, nr, k);
                        return FDAS_ERR_INVALID;
                    }
                }
                pl->W[h] = std::max(pl->W[h], nr);
                pl->sub_cells[h] += (long long)pl->Z[h] * nr;
                rlo_h[(size_t)item * MAX_HARM + h] = rl;
                const int lobin = rl - bo;
                const int hibin = (int)std::ceil(drhi) + bo;
                const int numdata = hibin - lobin + 1;
                // power-of-two lengths as the reference; a length R*2^k must hold
                // every normalised bin and the template span of the last column
                const bool fits = odd_factor(pl->L[h]) > 1
                                      ? numdata <= pl->L[h] / 2 && nr + 2 * pl->off[h] <= pl->L[h]
                                      : h_next_pow2(numdata) == pl->L[h] / 2;
                if (!fits) {
                    set_error("step %d harmonic %d/%d: %d bins to normalise do not fit FFT length %d",
                              k, pl->cur[h], pl->mx[h], numdata, pl->L[h]);
                    return FDAS_ERR_UNSUPPORTED;
                }
                for (size_t c = 0; c < pl->classes.size(); ++c) {
                    const auto& hl = pl->class_harm[c];
                    auto q = std::find(hl.begin(), hl.end(), h);
                    if (q == hl.end()) continue;
                    SegDesc& e = desc[c][(size_t)item * hl.size() + (q - hl.begin())];
                    e.lobin = lobin;
                    e.numdata = numdata;
                    e.out = (long long)item * pl->seg_stride + pl->seg_off[h];
                    e.eslot = item * H + h;
                    e.nbins = (int)tr.nbins;
                    e.spec_off = tr.spec_off;
                    e.nch = 0;
                }
            }
        }
    }
    long long sub = 0;
    for (int h = 0; h < H; ++h) {
        // sub-harmonic rows padded to 4 floats, so the harmonic sum stages them
        // with 16-byte loads (the extra columns are correlation outputs too)
        if (h > 0) pl->W[h] = (pl->W[h] + 3) & ~3;
        if (pl->off[h] + pl->W[h] > pl->L[h]) {
            set_error("harmonic %d: %d columns from offset %d exceed FFT length %d", h, pl->W[h],
                      pl->off[h], pl->L[h]);
            return FDAS_ERR_INVALID;
        }
        if (h > 0) sub += (long long)pl->Z[h] * pl->W[h];
    }
    pl->sub_total = sub;
    if (pl->mma) {
        // staged chunks per harmonic follow from the final row widths: hi and lo
        // byte arrays of nch chunks each, harmonics one after another per item
        long long so = 0;
        for (int h = 0; h < H; ++h) {
            mma_geometry(pl->W[h], pl->off[h] / NUMBETWEEN, &pl->nm[h], &pl->nch[h]);
            pl->seg_off[h] = so;
            so += 2LL * pl->nch[h];
        }
        pl->seg_stride = so;
        for (size_t c = 0; c < pl->classes.size(); ++c) {
            const auto& hl = pl->class_harm[c];
            for (int item = 0; item < S; ++item)
                for (size_t q = 0; q < hl.size(); ++q) {
                    SegDesc& e = desc[c][(size_t)item * hl.size() + q];
                    e.out = (long long)item * pl->seg_stride + pl->seg_off[hl[q]];
                    e.nch = pl->nch[hl[q]];
                }
        }
    }
    return FDAS_OK;
}

int plan_device(fdas_plan* pl, bool own_series, const std::vector<int>& fullrlo,
                const std::vector<int>& rlo_h, const std::vector<std::vector<SegDesc>>& desc);

// Host tables outlive every copy queued from them: the stream is synchronised
// before they go out of scope, on success and on failure.
int plan_build(fdas_plan* pl, const long* lengths, int ntrials, bool own_series)
{
    FDAS_TRY(plan_geometry(pl, lengths, ntrials));
    std::vector<int> fullrlo, rlo_h;
    std::vector<std::vector<SegDesc>> desc;
    FDAS_TRY(plan_tables(pl, fullrlo, rlo_h, desc));
    const int rc = plan_device(pl, own_series, fullrlo, rlo_h, desc);
    if (pl->stream && hipStreamSynchronize(pl->stream) != hipSuccess && rc == FDAS_OK)
        return hip_fail(hipGetLastError(), "hipStreamSynchronize", __FILE__, __LINE__);
    return rc;
}

int plan_device(fdas_plan* pl, bool own_series, const std::vector<int>& fullrlo,
                const std::vector<int>& rlo_h, const std::vector<std::vector<SegDesc>>& desc)
{
    const fdas_params& p = pl->p;
    const int U = pl->uselen, H = pl->H;

    // chunking: per-item buffer bytes against a device-memory budget
    const size_t per_item = (size_t)pl->seg_stride * pl->cbytes + (size_t)pl->sub_total * pl->pbytes +
                            (size_t)pl->Z0 * U * pl->pbytes + (size_t)H * sizeof(int);
    // default: as many items per chunk as 24 GiB of chunk buffers hold (a whole
    // 2^23-sample trial at zmax 200: fewer, larger launches measured fastest, and
    // overlapping the harmonic sum of one chunk with the correlations of the
    // next on a second stream measured slower, since both kernels fill every CU)
    int chunk = p.chunk_steps;
    if (chunk <= 0) chunk = (int)std::max<size_t>(1, ((size_t)24 << 30) / per_item);
    pl->chunk = std::min(chunk, pl->nitems);

    // launch shapes in effect now, and the data layouts they read
    for (int h = 0; h < H; ++h) {
        pl->corr_var[h] = pl->mma ? 0 : corr_variant(pl->L[h]);
        pl->decim[h] = pl->mma ? 0 : corr_variant_decim(pl->L[h], pl->corr_var[h]);
    }
    pl->hs_var = harmsum_variant();
    pl->mma_var = mma_variant();
    pl->class_decim.assign(pl->classes.size(), 0);
    for (size_t c = 0; c < pl->classes.size(); ++c) pl->class_decim[c] = pl->decim[pl->class_harm[c][0]];

    // device state (everything above is host-side validation and geometry)
    if (!pl->stream) {
        FDAS_HIP(hipStreamCreate(&pl->stream));
        pl->own_stream = true;
    }
    FDAS_TRY(build_banks(p.zmax, H, pl->cur, pl->mx, pl->L, U, pl->prec, p.fft_lengths, pl->decim,
                         pl->stream, &pl->banks));
    for (size_t c = 0; c < desc.size(); ++c) {
        SegDesc* dd = nullptr;
        FDAS_TRY(plan_alloc(pl, (void**)&dd, desc[c].size() * sizeof(SegDesc)));
        pl->d_desc.push_back(dd);
        FDAS_TRY(plan_upload(pl, dd, desc[c].data(), desc[c].size() * sizeof(SegDesc)));
    }
    FDAS_TRY(plan_alloc(pl, (void**)&pl->d_fullrlo, fullrlo.size() * sizeof(int)));
    FDAS_TRY(plan_alloc(pl, (void**)&pl->d_rlo_h, rlo_h.size() * sizeof(int)));
    FDAS_TRY(plan_upload(pl, pl->d_fullrlo, fullrlo.data(), fullrlo.size() * sizeof(int)));
    FDAS_TRY(plan_upload(pl, pl->d_rlo_h, rlo_h.data(), rlo_h.size() * sizeof(int)));
    long long spec = 0, series = 0;
    size_t work = 0;
    for (TrialInfo& tr : pl->trials) {
        spec += tr.nbins + 1;
        series += tr.n;
        size_t f = 0;
        while (f < pl->forms.size() && pl->forms[f].n != tr.n) ++f;
        if (f == pl->forms.size()) {
            pl->forms.emplace_back();
            pl->forms.back().n = tr.n;
            FDAS_TRY(r2c_plan_create(tr.n, &pl->forms.back().r2c));
            FDAS_TRY(whiten_schedule_create((unsigned)tr.nbins, &pl->forms.back().ws, pl->al,
                                            pl->stream));
            work = std::max(work, pl->forms.back().r2c.work_elems);
        }
        tr.form = (int)f;
    }
    if (own_series) FDAS_TRY(plan_alloc(pl, (void**)&pl->d_series, (size_t)series * sizeof(float)));
    FDAS_TRY(plan_alloc(pl, (void**)&pl->d_spec, (size_t)spec * sizeof(cplx)));
    FDAS_TRY(plan_alloc(pl, (void**)&pl->d_r2c_work, work * sizeof(cplx)));
    FDAS_TRY(plan_alloc(pl, &pl->d_seg, (size_t)pl->chunk * pl->seg_stride * pl->cbytes));
    FDAS_TRY(plan_alloc(pl, (void**)&pl->d_seg_exp, (size_t)pl->chunk * H * sizeof(int)));
    if (pl->sub_total > 0)
        FDAS_TRY(plan_alloc(pl, &pl->d_sub, (size_t)pl->chunk * pl->sub_total * pl->pbytes));
    FDAS_TRY(plan_alloc(pl, &pl->d_plane, (size_t)pl->chunk * pl->Z0 * U * pl->pbytes));
    pl->cand_cap = p.max_candidates > 0 ? p.max_candidates : (1 << 20);
    FDAS_TRY(plan_alloc(pl, (void**)&pl->d_cands, (size_t)pl->cand_cap * sizeof(fdas_candidate)));
    FDAS_TRY(plan_alloc(pl, (void**)&pl->d_ncand, sizeof(int)));
    FDAS_HIP(hipMemsetAsync(pl->d_ncand, 0, sizeof(int), pl->stream));

    // harmonic-sum arguments
    HarmSumArgs& hs = pl->hs;
    hs.H = H;
    hs.U = U;
    hs.zlo0 = pl->zlo[0];
    long long acc = 0;
    for (int h = 0; h < H; ++h) {
        hs.c[h] = pl->cur[h];
        hs.lm[h] = h_fast_log2(pl->mx[h]);
        hs.zlo[h] = pl->zlo[h];
        hs.Z[h] = pl->Z[h];
        hs.W[h] = pl->W[h];
        if (h > 0) {
            hs.sub_off[h] = acc;
            hs.sub_stride[h] = (long long)pl->Z[h] * pl->W[h];
            acc += (long long)pl->chunk * hs.sub_stride[h];
        }
    }
    for (int st = 0; st < 5; ++st) hs.powcut[st] = p.powcut[st] > 0.f ? p.powcut[st] : INFINITY;
    hs.cand_cap = pl->cand_cap;
    FDAS_HIP(hipStreamSynchronize(pl->stream));
    return FDAS_OK;
}

int prepare_trial(fdas_plan* pl, int t, const float* d_series)
{
    const TrialInfo& tr = pl->trials[t];
    const LengthForm& lf = pl->forms[tr.form];
    {
        TimedLaunch tl(pl, FDAS_K_R2C);
        FDAS_TRY(fft_r2c_plan(lf.r2c, d_series, pl->d_spec + tr.spec_off, pl->d_r2c_work, pl->stream));
    }
    TimedLaunch tl(pl, FDAS_K_WHITEN);
    return whiten(pl->d_spec + tr.spec_off, lf.ws, pl->stream);
}

int reset_candidates(fdas_plan* pl)
{
    // a kernel, not a memset: this runs inside captured graphs too
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, pl->stream, pl->d_ncand);
    return check_launch("zero_int_kernel");
}

// Chunks of items, each: segment preparation per FFT-length class, the
// correlations of every harmonic, then the harmonic sum.
int run_items(fdas_plan* pl, int item0, int nitems, float* d_planes)
{
    const int U = pl->uselen;
    const BankSet& bs = *pl->banks;
    for (int c0 = item0; c0 < item0 + nitems; c0 += pl->chunk) {
        const int cs = std::min(pl->chunk, item0 + nitems - c0);
        for (size_t c = 0; c < pl->classes.size(); ++c) {
            const int nh = (int)pl->class_harm[c].size();
            SegPrepArgs a{};
            a.spec = pl->d_spec;
            a.desc = pl->d_desc[c] + (size_t)c0 * nh;
            a.nseg = cs * nh;
            a.seg = pl->d_seg;
            // descriptors hold offsets as if every item had a buffer slot
            a.out_base = (long long)c0 * pl->seg_stride;
            a.seg_exp = pl->d_seg_exp;
            a.eslot_base = c0 * pl->H;
            a.prec = pl->prec;
            a.decim = pl->class_decim[c];
            TimedLaunch tl(pl, FDAS_K_SEGPREP);
            if (pl->mma) FDAS_TRY(launch_segment_prep_mma(pl->classes[c], a, pl->stream));
            else FDAS_TRY(launch_segment_prep(pl->classes[c], a, pl->stream));
        }
        auto mma_args = [&](int h, void* out, int W) {
            MmaCorrArgs a{};
            a.bank = &bs.mma[h];
            a.seg = pl->d_seg;
            a.seg_stride = pl->seg_stride;
            a.seg_off = pl->seg_off[h];
            a.nch = pl->nch[h];
            a.seg_inv = reinterpret_cast<const float*>(pl->d_seg_exp);
            a.H = pl->H;
            a.h = h;
            a.nsteps = cs;
            a.W = W;
            a.Z = pl->Z[h];
            a.nm = pl->nm[h];
            a.out = out;
            a.variant = pl->mma_var;
            return a;
        };
        auto corr_args = [&](int h, void* out, int W) {
            CorrArgs a{};
            a.bank = bs.bankq[h];
            a.bank_exp = bs.bank_exp[h];
            a.bank_split = bs.splitq[h];
            a.bank_split_exp = bs.split_exp[h];
            a.Z = pl->Z[h];
            a.seg = pl->d_seg;
            a.seg_stride = pl->seg_stride;
            a.seg_off = pl->seg_off[h];
            a.seg_exp = pl->d_seg_exp;
            a.H = pl->H;
            a.h = h;
            a.nsteps = cs;
            a.off = pl->off[h];
            a.W = W;
            a.out = out;
            a.prec = pl->prec;
            return a;
        };
        for (int h = 1; h < pl->H; ++h) {
            TimedLaunch tl(pl, FDAS_K_CORR_SUB);
            char* sub_h = static_cast<char*>(pl->d_sub) + (size_t)pl->hs.sub_off[h] * pl->pbytes;
            if (pl->mma) FDAS_TRY(launch_corr_mma(mma_args(h, sub_h, pl->W[h]), pl->stream));
            else FDAS_TRY(launch_corr(pl->L[h], corr_args(h, sub_h, pl->W[h]), pl->stream, pl->corr_var[h]));
        }
        // summed powers for the caller, FP32; with FP32 planes the fundamental
        // plane is that buffer itself and the sum goes back in place
        float* out = d_planes ? d_planes + (size_t)(c0 - item0) * pl->Z0 * U : nullptr;
        void* plane = out && pl->prec == FDAS_PREC_FP32 ? static_cast<void*>(out) : pl->d_plane;
        {
            TimedLaunch tl(pl, FDAS_K_CORR_FUND);
            if (pl->mma) FDAS_TRY(launch_corr_mma(mma_args(0, plane, U), pl->stream));
            else FDAS_TRY(launch_corr(pl->L[0], corr_args(0, plane, U), pl->stream, pl->corr_var[0]));
        }
        // The harmonic sum as its own tiled kernel: summing inside the fundamental
        // correlation's epilogue (gathering sub-harmonic cells after each inverse
        // FFT) measured 1.3-1.5x slower at FFT lengths 4096 and 8192.  The summed
        // plane is stored only into a caller's buffer: the internal plane's sole
        // consumer is the threshold, which the kernel applies to every cell.
        TimedLaunch tl(pl, FDAS_K_HARMSUM);
        FDAS_TRY(launch_harmsum(plane, out, pl->d_sub, pl->prec, pl->hs, c0, cs, pl->Z0, pl->d_fullrlo,
                                pl->d_rlo_h, pl->d_cands, pl->d_ncand, pl->stream, pl->hs_var));
    }
    return FDAS_OK;
}

// Every trial from the plan's own series buffers, then every item.
int batch_body(fdas_plan* pl)
{
    for (size_t t = 0; t < pl->trials.size(); ++t)
        FDAS_TRY(prepare_trial(pl, (int)t, pl->d_series + pl->trials[t].series_off));
    FDAS_TRY(reset_candidates(pl));
    return run_items(pl, 0, pl->nitems, nullptr);
}

}  // namespace

namespace fdas {
int plan_create_ex(const fdas_params* params, int ntrials, const long* lengths, bool own_series,
                   hipStream_t stream, DevAlloc* al, fdas_plan** out)
{
    if (!params || !out || ntrials < 1 || !lengths) {
        set_error("fdas plan: null argument or no trials");
        return FDAS_ERR_INVALID;
    }
    *out = nullptr;
    fdas_plan* pl = new fdas_plan();
    pl->p = *params;
    pl->al = al ? al : hip_alloc();
    pl->stream = stream;
    const int rc = plan_build(pl, lengths, ntrials, own_series);
    if (rc != FDAS_OK) {
        fdas_plan_destroy(pl);
        return rc;
    }
    *out = pl;
    return FDAS_OK;
}

// Shared handle on a plan's template banks (keeps them resident while held).
std::shared_ptr<void> plan_banks(fdas_plan* pl) { return pl->banks; }

int plan_set_harmsum_variant(fdas_plan* pl, int v)
{
    if (v < 0 || v >= harmsum_variant_count()) {
        set_error("no harmonic-sum variant %d", v);
        return FDAS_ERR_INVALID;
    }
    if (pl->graph) {
        (void)hipGraphExecDestroy(pl->graph);
        pl->graph = nullptr;
    }
    pl->hs_var = v;
    return FDAS_OK;
}

}  // namespace fdas

// ====================================================================== C ABI
extern "C" void fdas_plan_destroy(fdas_plan* pl)
{
    if (!pl) return;
    if (pl->stream) (void)hipStreamSynchronize(pl->stream);
    if (pl->graph) (void)hipGraphExecDestroy(pl->graph);
    DevAlloc* al = pl->al ? pl->al : hip_alloc();
    for (SegDesc* d : pl->d_desc) al->release(d);
    al->release(pl->d_fullrlo);
    al->release(pl->d_rlo_h);
    al->release(pl->d_series);
    al->release(pl->d_spec);
    al->release(pl->d_r2c_work);
    for (LengthForm& lf : pl->forms) whiten_schedule_destroy(&lf.ws, al);
    al->release(pl->d_seg);
    al->release(pl->d_seg_exp);
    al->release(pl->d_sub);
    al->release(pl->d_plane);
    al->release(pl->d_cands);
    al->release(pl->d_ncand);
    for (auto& r : pl->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    pl->banks.reset();
    if (pl->own_stream) (void)hipStreamDestroy(pl->stream);
    delete pl;
}

extern "C" int fdas_plan_create(const fdas_params* params, fdas_plan** out)
{
    if (!params) {
        set_error("fdas_plan_create: null argument");
        return FDAS_ERR_INVALID;
    }
    const long n = params->n;
    return plan_create_ex(params, 1, &n, false, nullptr, nullptr, out);
}

extern "C" int fdas_batch_create(const fdas_params* params, int ntrials, const long* lengths,
                                 fdas_plan** out)
{
    return plan_create_ex(params, ntrials, lengths, true, nullptr, nullptr, out);
}

extern "C" int fdas_harmonic_fftlens(const fdas_params* params, int* uselen, int* lengths)
{
    if (!params || !uselen || !lengths) {
        set_error("fdas_harmonic_fftlens: null argument");
        return FDAS_ERR_INVALID;
    }
    fdas_params p = *params;
    if (p.fftlen == 0) p.fftlen = resolve_fftlen(p);
    FDAS_TRY(validate(p));
    const int U = p.uselen ? p.uselen : h_max_uselen(p.fftlen, p.zmax);
    if (U <= 0 || (U & 1) || U >= 65536 || h_fft_length(U, 1, 1, p.zmax) != p.fftlen) {
        set_error("uselen %d does not give fundamental FFT length %d at zmax %d", U, p.fftlen, p.zmax);
        return FDAS_ERR_INVALID;
    }
    int cur[MAX_HARM], mx[MAX_HARM];
    h_harmonic_order(p.numharm, cur, mx);
    *uselen = U;
    for (int h = 0; h < p.numharm; ++h) lengths[h] = h_harmonic_length(U, mx[h], cur[h], p.zmax, p.fft_lengths);
    return FDAS_OK;
}

extern "C" int fdas_plan_get_info(const fdas_plan* pl, fdas_plan_info* info)
{
    if (!pl || !info) {
        set_error("fdas_plan_get_info: null argument");
        return FDAS_ERR_INVALID;
    }
    memset(info, 0, sizeof(*info));
    info->uselen = pl->uselen;
    info->rlo = pl->trials[0].rlo;
    info->nsteps = pl->nitems;
    info->numharm = pl->H;
    info->numz = pl->Z0;
    for (int h = 0; h < pl->H; ++h) {
        info->fftlen[h] = pl->L[h];
        info->numzs[h] = pl->Z[h];
        info->harm_cur[h] = pl->cur[h];
        info->harm_max[h] = pl->mx[h];
        info->sub_cells[h] = pl->sub_cells[h];
    }
    info->chunk_steps = pl->chunk;
    info->nbins = pl->trials[0].nbins;
    info->cells = pl->cells;
    info->device_bytes = pl->bytes;
    info->ntrials = (int)pl->trials.size();
    info->template_bytes = pl->banks ? pl->banks->bytes : 0;
    return FDAS_OK;
}

extern "C" int fdas_batch_trial_info(const fdas_plan* pl, int trial, fdas_trial_info* info)
{
    if (!pl || !info || trial < 0 || trial >= (int)pl->trials.size()) {
        set_error("fdas_batch_trial_info: bad argument");
        return FDAS_ERR_INVALID;
    }
    const TrialInfo& tr = pl->trials[trial];
    info->n = tr.n;
    info->nbins = tr.nbins;
    info->rlo = tr.rlo;
    info->nsteps = tr.nsteps;
    info->first_step = tr.item0;
    return FDAS_OK;
}

extern "C" float* fdas_batch_series(fdas_plan* pl, int trial)
{
    if (!pl || !pl->d_series || trial < 0 || trial >= (int)pl->trials.size()) {
        set_error("fdas_batch_series: no plan-owned series for trial %d", trial);
        return nullptr;
    }
    return pl->d_series + pl->trials[trial].series_off;
}

extern "C" int fdas_plan_set_stream(fdas_plan* pl, void* stream)
{
    if (!pl) {
        set_error("fdas_plan_set_stream: null plan");
        return FDAS_ERR_INVALID;
    }
    if (pl->own_stream) {
        (void)hipStreamSynchronize(pl->stream);
        (void)hipStreamDestroy(pl->stream);
        pl->own_stream = false;
    }
    if (pl->graph) {
        (void)hipGraphExecDestroy(pl->graph);
        pl->graph = nullptr;
    }
    if (stream) {
        pl->stream = (hipStream_t)stream;
    } else {
        FDAS_HIP(hipStreamCreate(&pl->stream));
        pl->own_stream = true;
    }
    return FDAS_OK;
}

extern "C" int fdas_prepare_trial(fdas_plan* pl, int trial, const float* d_series)
{
    if (!pl || trial < 0 || trial >= (int)pl->trials.size()) {
        set_error("fdas_prepare_trial: bad argument");
        return FDAS_ERR_INVALID;
    }
    if (!d_series) d_series = fdas_batch_series(pl, trial);
    if (!d_series) return FDAS_ERR_INVALID;
    return prepare_trial(pl, trial, d_series);
}

extern "C" int fdas_prepare(fdas_plan* pl, const float* d_series)
{
    if (!pl || !d_series) {
        set_error("fdas_prepare: null argument");
        return FDAS_ERR_INVALID;
    }
    if (pl->trials.size() != 1) {
        set_error("fdas_prepare: plan has %zu trials (use fdas_prepare_trial)", pl->trials.size());
        return FDAS_ERR_INVALID;
    }
    FDAS_TRY(prepare_trial(pl, 0, d_series));
    return reset_candidates(pl);
}

extern "C" int fdas_reset_candidates(fdas_plan* pl)
{
    if (!pl) {
        set_error("fdas_reset_candidates: null plan");
        return FDAS_ERR_INVALID;
    }
    return reset_candidates(pl);
}

extern "C" int fdas_run_steps(fdas_plan* pl, int step0, int nsteps, float* d_planes)
{
    if (!pl || step0 < 0 || nsteps < 0 || step0 + nsteps > pl->nitems) {
        set_error("fdas_run_steps: steps [%d, %d) outside [0, %d)", step0, step0 + nsteps,
                  pl ? pl->nitems : 0);
        return FDAS_ERR_INVALID;
    }
    return run_items(pl, step0, nsteps, d_planes);
}

extern "C" int fdas_search(fdas_plan* pl, const float* d_series)
{
    FDAS_TRY(fdas_prepare(pl, d_series));
    return run_items(pl, 0, pl->nitems, nullptr);
}

extern "C" int fdas_plan_enable_graph(fdas_plan* pl, int enable)
{
    if (!pl) {
        set_error("fdas_plan_enable_graph: null plan");
        return FDAS_ERR_INVALID;
    }
    if (!pl->d_series) {
        set_error("graph replay needs a batch plan (plan-owned series)");
        return FDAS_ERR_INVALID;
    }
    pl->use_graph = enable != 0;
    if (!pl->use_graph && pl->graph) {
        (void)hipGraphExecDestroy(pl->graph);
        pl->graph = nullptr;
    }
    return FDAS_OK;
}

extern "C" int fdas_batch_run(fdas_plan* pl)
{
    if (!pl || !pl->d_series) {
        set_error("fdas_batch_run: not a batch plan");
        return FDAS_ERR_INVALID;
    }
    if (!pl->use_graph || pl->timing) return batch_body(pl);
    if (!pl->graph) {
        // one eager run creates the lazily built twiddle tables outside the capture
        FDAS_TRY(batch_body(pl));
        FDAS_HIP(hipStreamSynchronize(pl->stream));
        hipGraph_t g = nullptr;
        FDAS_HIP(hipStreamBeginCapture(pl->stream, hipStreamCaptureModeThreadLocal));
        const int rc = batch_body(pl);
        hipGraph_t cap = nullptr;
        const hipError_t e = hipStreamEndCapture(pl->stream, &cap);
        FDAS_TRY(rc);
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
        g = cap;
        const hipError_t e2 = hipGraphInstantiate(&pl->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e2 != hipSuccess) return hip_fail(e2, "hipGraphInstantiate", __FILE__, __LINE__);
        return FDAS_OK;
    }
    FDAS_HIP(hipGraphLaunch(pl->graph, pl->stream));
    return FDAS_OK;
}

extern "C" int fdas_get_spectrum(fdas_plan* pl, fdas_complex* d_out)
{
    return fdas_batch_spectrum(pl, 0, d_out);
}

extern "C" int fdas_batch_spectrum(fdas_plan* pl, int trial, fdas_complex* d_out)
{
    if (!pl || !d_out || trial < 0 || trial >= (int)pl->trials.size()) {
        set_error("fdas_batch_spectrum: bad argument");
        return FDAS_ERR_INVALID;
    }
    const TrialInfo& tr = pl->trials[trial];
    FDAS_HIP(hipMemcpyAsync(d_out, pl->d_spec + tr.spec_off, (size_t)tr.nbins * sizeof(cplx),
                            hipMemcpyDeviceToDevice, pl->stream));
    return FDAS_OK;
}

extern "C" int fdas_get_candidates(fdas_plan* pl, fdas_candidate* out, int capacity, int* count,
                                   int* overflowed)
{
    if (!pl || !count) {
        set_error("fdas_get_candidates: null argument");
        return FDAS_ERR_INVALID;
    }
    int n = 0;
    FDAS_HIP(hipMemcpyAsync(&n, pl->d_ncand, sizeof(int), hipMemcpyDeviceToHost, pl->stream));
    FDAS_HIP(hipStreamSynchronize(pl->stream));
    if (overflowed) *overflowed = n > pl->cand_cap;
    if (n > pl->cand_cap) {
        // the stored subset depends on the order the kernels found the cells in
        *count = n;
        set_error("%d candidates exceed the capacity of %d; raise max_candidates or the power cuts",
                  n, pl->cand_cap);
        return FDAS_ERR_OVERFLOW;
    }
    const int kept = n;
    *count = kept;
    if (!out || capacity <= 0) return FDAS_OK;
    std::vector<fdas_candidate> all(kept);
    if (kept > 0) {
        FDAS_HIP(hipMemcpyAsync(all.data(), pl->d_cands, (size_t)kept * sizeof(fdas_candidate),
                                hipMemcpyDeviceToHost, pl->stream));
        FDAS_HIP(hipStreamSynchronize(pl->stream));
    }
    // kernels record the item; report (trial, step within the trial)
    std::vector<int> first(pl->trials.size());
    for (size_t t = 0; t < first.size(); ++t) first[t] = pl->trials[t].item0;
    for (auto& c : all) {
        const int t = (int)(std::upper_bound(first.begin(), first.end(), c.step) - first.begin()) - 1;
        c.trial = t;
        c.step -= first[t];
    }
    std::sort(all.begin(), all.end(), [](const fdas_candidate& a, const fdas_candidate& b) {
        if (a.trial != b.trial) return a.trial < b.trial;
        if (a.step != b.step) return a.step < b.step;
        if (a.z_index != b.z_index) return a.z_index < b.z_index;
        if (a.r_index != b.r_index) return a.r_index < b.r_index;
        return a.stage < b.stage;
    });
    memcpy(out, all.data(), sizeof(fdas_candidate) * std::min(kept, capacity));
    return FDAS_OK;
}

extern "C" int fdas_synchronize(fdas_plan* pl)
{
    if (!pl) {
        set_error("fdas_synchronize: null plan");
        return FDAS_ERR_INVALID;
    }
    FDAS_HIP(hipStreamSynchronize(pl->stream));
    return FDAS_OK;
}

extern "C" int fdas_plan_enable_timing(fdas_plan* pl, int enable)
{
    if (!pl) {
        set_error("fdas_plan_enable_timing: null plan");
        return FDAS_ERR_INVALID;
    }
    FDAS_HIP(hipStreamSynchronize(pl->stream));
    pl->timing = enable != 0;
    pl->nrec = 0;
    return FDAS_OK;
}

extern "C" int fdas_plan_get_timing(fdas_plan* pl, fdas_timing* out)
{
    if (!pl || !out) {
        set_error("fdas_plan_get_timing: null argument");
        return FDAS_ERR_INVALID;
    }
    FDAS_HIP(hipStreamSynchronize(pl->stream));
    memset(out, 0, sizeof(*out));
    for (size_t i = 0; i < pl->nrec; ++i) {
        float ms = 0.f;
        FDAS_HIP(hipEventElapsedTime(&ms, pl->recs[i].a, pl->recs[i].b));
        out->launches[pl->recs[i].cat] += 1;
        out->ms[pl->recs[i].cat] += ms;
    }
    return FDAS_OK;
}

// ------------------------------------------------------------------ candidates
static std::vector<fdas_sift_trial> sift_trials(const fdas_plan* pl)
{
    std::vector<fdas_sift_trial> t(pl->trials.size());
    for (size_t i = 0; i < t.size(); ++i) {
        t[i].n = pl->trials[i].n;
        t[i].nbins = pl->trials[i].nbins;
        t[i].rlo = pl->trials[i].rlo;
        t[i].dt = pl->p.dt;
    }
    return t;
}

extern "C" int fdas_stage_powcuts(const fdas_plan* pl, int trial, double sigma, float* powcut)
{
    if (!pl || !powcut || trial < 0 || trial >= (int)pl->trials.size()) {
        set_error("fdas_stage_powcuts: bad argument");
        return FDAS_ERR_INVALID;
    }
    const TrialInfo& tr = pl->trials[trial];
    const int stages = h_fast_log2(pl->H) + 1;
    for (int s = 0; s < FDAS_MAX_STAGES; ++s)
        powcut[s] = s < stages ? (float)power_for_sigma(sigma, 1 << s,
                                                        numindep_count(tr.rlo, (double)tr.nbins, pl->Z0, s))
                               : 0.f;
    return FDAS_OK;
}

extern "C" int fdas_sift(const fdas_plan* pl, const fdas_candidate* in, int n_in,
                         const fdas_sift_params* params, fdas_accelcand* out, int capacity, int* n_out)
{
    if (!pl || (!in && n_in > 0) || n_in < 0 || !n_out) {
        set_error("fdas_sift: bad argument");
        return FDAS_ERR_INVALID;
    }
    const std::vector<fdas_sift_trial> t = sift_trials(pl);
    return sift_candidates(t.data(), (int)t.size(), pl->p.zmax, pl->uselen, pl->Z0, in, n_in, params,
                           out, capacity, n_out);
}

extern "C" int fdas_optimize(fdas_plan* pl, fdas_accelcand* c, int n)
{
    if (!pl || (!c && n > 0) || n < 0) {
        set_error("fdas_optimize: bad argument");
        return FDAS_ERR_INVALID;
    }
    std::vector<long long> off(pl->trials.size());
    std::vector<int> nb(pl->trials.size());
    for (size_t i = 0; i < off.size(); ++i) {
        off[i] = pl->trials[i].spec_off;
        nb[i] = (int)pl->trials[i].nbins;
    }
    for (int i = 0; i < n; ++i) {
        if (c[i].trial < 0 || c[i].trial >= (int)pl->trials.size() || c[i].numharm < 1 ||
            c[i].numharm > MAX_HARM) {
            set_error("fdas_optimize: candidate %d has trial %d / numharm %d out of range", i,
                      c[i].trial, c[i].numharm);
            return FDAS_ERR_INVALID;
        }
    }
    FDAS_TRY(optimize_candidates(pl->d_spec, off, nb, c, n, pl->stream));
    for (int i = 0; i < n; ++i) {
        const TrialInfo& tr = pl->trials[c[i].trial];
        c[i].opt_sigma = (float)candidate_sigma(
            c[i].opt_power, c[i].numharm,
            numindep_count(tr.rlo, (double)tr.nbins, pl->Z0, h_fast_log2(c[i].numharm)));
        cand_properties(&c[i], (double)tr.n * pl->p.dt);
    }
    return FDAS_OK;
}

extern "C" int fdas_rz_power(fdas_plan* pl, int trial, const double* r, const double* z, int count,
                             float* out)
{
    if (!pl || trial < 0 || trial >= (int)pl->trials.size() || count < 0 ||
        (count > 0 && (!r || !z || !out))) {
        set_error("fdas_rz_power: bad argument");
        return FDAS_ERR_INVALID;
    }
    const TrialInfo& tr = pl->trials[trial];
    return rz_powers(pl->d_spec + tr.spec_off, (int)tr.nbins, r, z, count, out, pl->stream);
}
