// serve.hip -- continuous-batching request loop (fdas_server_* in include/fdas.h).
//
// Each request is one DM-trial series with its own zmax / numharm / fftlen /
// precision.  A dispatcher thread admits queued requests in arrival order while
// a slot (a non-blocking HIP stream) is free and the request's plan fits in the
// server's device-memory arena; admitted requests run concurrently on their
// slots, so work of requests that arrived later overlaps on the device with
// work already running.  Memory comes from one region reserved up front and
// sub-allocated (first fit, coalescing), so admission and completion never call
// hipMalloc/hipFree, which would wait for the whole device.  Template banks stay
// resident per configuration for the server's lifetime.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <thread>
#include <vector>

#include "fdas_internal.hpp"

namespace fdas {

std::shared_ptr<void> plan_banks(fdas_plan* pl);

// ------------------------------------------------------------------ arena
class Arena final : public DevAlloc {
public:
    static constexpr size_t kAlign = 256;

    int init(size_t bytes)
    {
        size_ = (bytes + kAlign - 1) / kAlign * kAlign;
        FDAS_HIP(hipMalloc((void**)&base_, size_));
        free_[0] = size_;
        return FDAS_OK;
    }
    ~Arena() override
    {
        if (base_) (void)hipFree(base_);
    }
    int alloc(void** p, size_t bytes) override
    {
        const size_t need = std::max<size_t>(kAlign, (bytes + kAlign - 1) / kAlign * kAlign);
        std::lock_guard<std::mutex> lk(mu_);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < need) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > need) free_[off + need] = len - need;
            used_[off] = need;
            in_use_ += need;
            peak_ = std::max(peak_, in_use_);
            *p = base_ + off;
            return FDAS_OK;
        }
        *p = nullptr;
        set_error("serving arena: %zu bytes requested, %zu of %zu in use", need, in_use_, size_);
        return FDAS_ERR_NOMEM;
    }
    void release(void* p) override
    {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu_);
        const size_t off = (size_t)((char*)p - base_);
        auto u = used_.find(off);
        if (u == used_.end()) return;
        size_t start = off, len = u->second;
        in_use_ -= len;
        used_.erase(u);
        auto nx = free_.lower_bound(start);
        if (nx != free_.end() && nx->first == start + len) {
            len += nx->second;
            nx = free_.erase(nx);
        }
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == start) {
                start = pv->first;
                len += pv->second;
                free_.erase(pv);
            }
        }
        free_[start] = len;
    }
    size_t size() const { return size_; }
    size_t in_use()
    {
        std::lock_guard<std::mutex> lk(mu_);
        return in_use_;
    }
    size_t peak()
    {
        std::lock_guard<std::mutex> lk(mu_);
        return peak_;
    }

private:
    char* base_ = nullptr;
    size_t size_ = 0, in_use_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_, used_;
    std::mutex mu_;
};

using Clock = std::chrono::steady_clock;

static double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
    long long id = 0;
    fdas_params p{};
    std::vector<float> series;
    Clock::time_point t_submit, t_admit, t_end;
    fdas_plan* plan = nullptr;
    int status = FDAS_OK;
    char message[256] = "";
    std::vector<fdas_candidate> cands;
    int overflowed = 0, nsteps = 0;
    int nfound = 0;                  // candidates found (more than kept on overflow)
    long long cells = 0;
    size_t bytes = 0;
};

}  // namespace fdas

using namespace fdas;

struct fdas_server {
    fdas_server_config cfg{};
    int device = 0;
    std::unique_ptr<Arena> arena;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;
    std::vector<Request*> slot;
    std::deque<Request*> pending;
    std::map<long long, Request*> done;
    std::map<std::vector<int>, std::shared_ptr<void>> banks;   // resident templates
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool stop = false, exited = false;
    std::set<long long> outstanding;   // submitted, not yet collected
    long long next_id = 1;
    // statistics
    long long submitted = 0, completed = 0, failed = 0;
    int max_active_seen = 0;
    std::vector<double> latencies;
    bool have_first = false;
    Clock::time_point first_submit, last_end;
    double busy_s = 0.0;
    Clock::time_point busy_since;
    int nactive = 0;
};

namespace {

std::vector<int> bank_key(const fdas_params& p)
{
    return {p.zmax, p.numharm, p.fftlen, p.uselen, p.precision, p.fft_lengths};
}

// Creates the request's plan on the slot's stream and queues upload + search.
// FDAS_ERR_NOMEM: the arena cannot hold it now.
int admit(fdas_server* S, Request* r, int k)
{
    fdas_params p = r->p;
    p.chunk_steps = p.chunk_steps > 0 ? p.chunk_steps : S->cfg.chunk_steps;
    p.max_candidates = p.max_candidates > 0 ? p.max_candidates : S->cfg.max_candidates;
    const long n = p.n;
    hipStream_t s = S->streams[k];
    FDAS_TRY(plan_create_ex(&p, 1, &n, true, s, S->arena.get(), &r->plan));
    fdas_plan_info info;
    fdas_plan_get_info(r->plan, &info);
    r->nsteps = info.nsteps;
    r->cells = info.cells;
    r->bytes = info.device_bytes;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto& b = S->banks[bank_key(r->p)];
        if (!b) b = plan_banks(r->plan);
    }
    float* d = fdas_batch_series(r->plan, 0);
    FDAS_HIP(hipMemcpyAsync(d, r->series.data(), (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
    FDAS_TRY(fdas_batch_run(r->plan));
    FDAS_HIP(hipEventRecord(S->events[k], s));
    return FDAS_OK;
}

void finish(Request* r)
{
    int count = 0;
    if (r->status == FDAS_OK) {
        int rc = fdas_get_candidates(r->plan, nullptr, 0, &count, &r->overflowed);
        r->nfound = count;
        if (rc == FDAS_OK) {
            r->cands.resize(count);
            rc = fdas_get_candidates(r->plan, r->cands.data(), count, &count, &r->overflowed);
        }
        if (rc != FDAS_OK) {
            r->status = rc;
            snprintf(r->message, sizeof(r->message), "%s", last_error());
        }
    }
    fdas_plan_destroy(r->plan);
    r->plan = nullptr;
    std::vector<float>().swap(r->series);
}

void record_done(fdas_server* S, Request* r)
{
    S->done[r->id] = r;
    if (r->status == FDAS_OK) {
        ++S->completed;
        S->latencies.push_back(ms_between(r->t_submit, r->t_end));
        S->last_end = r->t_end;
    } else {
        ++S->failed;
    }
    S->cv_done.notify_all();
}

void worker_loop(fdas_server* S)
{
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    const int K = (int)S->slot.size();
    for (;;) {
        bool progress = false;
        // completions
        for (int k = 0; k < K; ++k) {
            Request* r = S->slot[k];
            if (!r) continue;
            const hipError_t q = hipEventQuery(S->events[k]);
            if (q == hipErrorNotReady) continue;
            r->t_end = Clock::now();
            if (q != hipSuccess) {
                r->status = FDAS_ERR_HIP;
                snprintf(r->message, sizeof(r->message), "request failed on the device: %s",
                         hipGetErrorString(q));
            }
            S->slot[k] = nullptr;
            if (--S->nactive == 0) S->busy_s += std::chrono::duration<double>(r->t_end - S->busy_since).count();
            lk.unlock();
            finish(r);
            lk.lock();
            record_done(S, r);
            progress = true;
        }
        // admissions, in arrival order
        while (!S->pending.empty()) {
            int k = 0;
            while (k < K && S->slot[k]) ++k;
            if (k == K) break;
            Request* r = S->pending.front();
            const bool others = S->nactive > 0;
            lk.unlock();
            int rc = admit(S, r, k);
            const Clock::time_point now = Clock::now();
            if (rc != FDAS_OK && r->plan) {
                fdas_plan_destroy(r->plan);
                r->plan = nullptr;
            }
            lk.lock();
            if (rc == FDAS_ERR_NOMEM && others) break;   // wait for memory to free up
            S->pending.pop_front();
            progress = true;
            if (rc != FDAS_OK) {
                r->status = rc;
                snprintf(r->message, sizeof(r->message), "%s%s", last_error(),
                         rc == FDAS_ERR_NOMEM ? " (request larger than the serving memory)" : "");
                r->t_end = now;
                std::vector<float>().swap(r->series);
                record_done(S, r);
                continue;
            }
            r->t_admit = now;
            S->slot[k] = r;
            if (S->nactive++ == 0) S->busy_since = now;
            S->max_active_seen = std::max(S->max_active_seen, S->nactive);
        }
        if (S->stop && S->pending.empty() && S->nactive == 0) {
            S->exited = true;
            S->cv_done.notify_all();
            break;
        }
        if (!progress) {
            if (S->nactive > 0) {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(50));
                lk.lock();
            } else {
                S->cv_work.wait(lk, [&] { return S->stop || !S->pending.empty(); });
            }
        }
    }
}

double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const size_t i = (size_t)std::ceil(q * (double)v.size());
    return v[std::min(v.size() - 1, i > 0 ? i - 1 : 0)];
}

}  // namespace

// ====================================================================== C ABI
extern "C" void fdas_server_destroy(fdas_server* S)
{
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    for (auto& kv : S->done) delete kv.second;
    S->banks.clear();
    for (hipEvent_t e : S->events) (void)hipEventDestroy(e);
    for (hipStream_t s : S->streams) (void)hipStreamDestroy(s);
    S->arena.reset();
    delete S;
}

extern "C" int fdas_server_create(const fdas_server_config* cfg, fdas_server** out)
{
    if (!out) {
        set_error("fdas_server_create: null argument");
        return FDAS_ERR_INVALID;
    }
    *out = nullptr;
    auto* S = new fdas_server();
    if (cfg) S->cfg = *cfg;
    if (S->cfg.max_active <= 0) S->cfg.max_active = 4;
    if (S->cfg.max_active > 16) S->cfg.max_active = 16;
    if (S->cfg.chunk_steps <= 0) S->cfg.chunk_steps = 256;
    if (S->cfg.max_candidates <= 0) S->cfg.max_candidates = 65536;
    int rc = FDAS_OK;
    hipError_t e = hipGetDevice(&S->device);
    if (e != hipSuccess) rc = hip_fail(e, "hipGetDevice", __FILE__, __LINE__);
    if (rc == FDAS_OK && S->cfg.memory_bytes == 0) {
        size_t fr = 0, tot = 0;
        e = hipMemGetInfo(&fr, &tot);
        if (e != hipSuccess) rc = hip_fail(e, "hipMemGetInfo", __FILE__, __LINE__);
        S->cfg.memory_bytes = fr / 2;
    }
    if (rc == FDAS_OK) {
        S->arena = std::make_unique<Arena>();
        rc = S->arena->init(S->cfg.memory_bytes);
    }
    for (int k = 0; rc == FDAS_OK && k < S->cfg.max_active; ++k) {
        hipStream_t s = nullptr;
        hipEvent_t e = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            set_error("fdas_server_create: stream/event creation failed");
            rc = FDAS_ERR_HIP;
        }
        if (s) S->streams.push_back(s);
        if (e) S->events.push_back(e);
        S->slot.push_back(nullptr);
    }
    if (rc != FDAS_OK) {
        fdas_server_destroy(S);
        return rc;
    }
    S->worker = std::thread(worker_loop, S);
    *out = S;
    return FDAS_OK;
}

extern "C" int fdas_server_prewarm(fdas_server* S, const fdas_params* params)
{
    if (!S || !params) {
        set_error("fdas_server_prewarm: null argument");
        return FDAS_ERR_INVALID;
    }
    // a plan of the smallest series with steps builds (or finds) the banks
    fdas_params p = *params;
    p.chunk_steps = 1;
    p.max_candidates = 1;
    long n = std::max<long>(p.n, 256);
    fdas_plan* pl = nullptr;
    int rc = FDAS_ERR_INVALID;
    for (int tries = 0; tries < 24 && rc != FDAS_OK; ++tries, n *= 2) {
        p.n = n;
        rc = plan_create_ex(&p, 1, &n, false, nullptr, nullptr, &pl);
        if (rc != FDAS_OK && rc != FDAS_ERR_INVALID) return rc;   // bad config, not just too short
    }
    if (rc != FDAS_OK) return rc;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->banks[bank_key(*params)] = plan_banks(pl);
    }
    fdas_plan_destroy(pl);
    return FDAS_OK;
}

extern "C" int fdas_server_submit(fdas_server* S, const fdas_params* params, const float* h_series,
                                  long long* id)
{
    if (!S || !params || !h_series || !id || params->n <= 0) {
        set_error("fdas_server_submit: bad argument");
        return FDAS_ERR_INVALID;
    }
    auto* r = new Request();
    r->p = *params;
    r->series.assign(h_series, h_series + params->n);
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) {
            delete r;
            set_error("fdas_server_submit: server is shutting down");
            return FDAS_ERR_INVALID;
        }
        r->id = S->next_id++;
        *id = r->id;
        S->outstanding.insert(r->id);
        ++S->submitted;
        if (!S->have_first) {
            S->have_first = true;
            S->first_submit = r->t_submit;
        }
        S->pending.push_back(r);
    }
    S->cv_work.notify_all();
    return FDAS_OK;
}

extern "C" int fdas_server_wait(fdas_server* S, long long id, fdas_request_result* res,
                                fdas_candidate* cands, int capacity)
{
    if (!S || !res) {
        set_error("fdas_server_wait: null argument");
        return FDAS_ERR_INVALID;
    }
    Request* r = nullptr;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        if (!S->outstanding.count(id)) {
            set_error("fdas_server_wait: request %lld unknown or already collected", id);
            return FDAS_ERR_INVALID;
        }
        S->cv_done.wait(lk, [&] { return S->done.count(id) > 0 || S->exited; });
        auto it = S->done.find(id);
        if (it == S->done.end()) {
            set_error("fdas_server_wait: server stopped before request %lld finished", id);
            return FDAS_ERR_INVALID;
        }
        r = it->second;
        S->done.erase(it);
        S->outstanding.erase(id);
    }
    memset(res, 0, sizeof(*res));
    res->id = r->id;
    res->status = r->status;
    res->ncandidates = r->nfound;
    res->overflowed = r->overflowed;
    res->nsteps = r->nsteps;
    res->cells = r->cells;
    res->device_bytes = r->bytes;
    const bool admitted = r->t_admit.time_since_epoch().count() != 0;
    res->queue_ms = ms_between(r->t_submit, admitted ? r->t_admit : r->t_end);
    res->service_ms = admitted ? ms_between(r->t_admit, r->t_end) : 0.0;
    res->latency_ms = ms_between(r->t_submit, r->t_end);
    snprintf(res->message, sizeof(res->message), "%s", r->message);
    if (cands && capacity > 0)
        memcpy(cands, r->cands.data(), sizeof(fdas_candidate) * std::min<size_t>(capacity, r->cands.size()));
    const int st = r->status;
    delete r;
    if (st != FDAS_OK) set_error("%s", res->message);
    return FDAS_OK;
}

extern "C" int fdas_server_get_stats(fdas_server* S, fdas_server_stats* out)
{
    if (!S || !out) {
        set_error("fdas_server_get_stats: null argument");
        return FDAS_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    memset(out, 0, sizeof(*out));
    out->submitted = S->submitted;
    out->completed = S->completed;
    out->failed = S->failed;
    out->active = S->nactive;
    out->pending = (int)S->pending.size();
    out->max_active_seen = S->max_active_seen;
    out->memory_bytes = S->arena->size();
    out->memory_in_use = S->arena->in_use();
    out->memory_peak = S->arena->peak();
    out->p50_ms = percentile(S->latencies, 0.50);
    out->p99_ms = percentile(S->latencies, 0.99);
    double sum = 0.0;
    for (double v : S->latencies) sum += v;
    out->mean_ms = S->latencies.empty() ? 0.0 : sum / (double)S->latencies.size();
    double busy = S->busy_s;
    if (S->nactive > 0) busy += std::chrono::duration<double>(Clock::now() - S->busy_since).count();
    out->busy_s = busy;
    if (S->completed > 0 && S->have_first) {
        const double span = std::chrono::duration<double>(S->last_end - S->first_submit).count();
        out->trials_per_s = span > 0 ? (double)S->completed / span : 0.0;
    }
    return FDAS_OK;
}

extern "C" int fdas_server_reset_stats(fdas_server* S)
{
    if (!S) {
        set_error("fdas_server_reset_stats: null server");
        return FDAS_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    S->submitted = S->completed = S->failed = 0;
    S->max_active_seen = S->nactive;
    S->latencies.clear();
    S->have_first = false;
    S->busy_s = 0.0;
    S->busy_since = Clock::now();
    return FDAS_OK;
}
