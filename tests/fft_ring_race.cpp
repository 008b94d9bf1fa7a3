// Race check of the LDS exchange ring in gpu_accelsearch_amd/csrc/fft_block.hpp.
// Each workgroup thread runs as an OS thread and block_sync() is a real barrier.
// After every barrier thread 0 sleeps before going on, so the other threads run
// ahead to their next barrier: any LDS region that is stored again before a
// barrier separates it from thread 0's pending loads hands thread 0 overwritten
// values, and its outputs stop matching the DFT.  Successive calls on the same
// LDS (as a workgroup's successive steps) check the barrier between calls.
// Built by tests/test_lds_ring.py with g++ and called through ctypes.
#define FDAS_HOST_BARRIER 1
#include "../gpu_accelsearch_amd/csrc/fft_block.hpp"

#include <chrono>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

namespace fdas {

namespace {
std::mutex g_mu;
std::condition_variable g_cv;
int g_count = 0, g_total = 0;
long g_gen = 0;
int g_sleep_us = 0;
thread_local int g_tid = -1;
}  // namespace

void host_barrier()
{
    {
        std::unique_lock<std::mutex> lk(g_mu);
        const long gen = g_gen;
        if (++g_count == g_total) {
            g_count = 0;
            ++g_gen;
            g_cv.notify_all();
        } else {
            g_cv.wait(lk, [&] { return g_gen != gen; });
        }
    }
    if (g_tid == 0) std::this_thread::sleep_for(std::chrono::microseconds(g_sleep_us));
}

// A barrier among the 64 threads of the calling thread's wavefront (the lanes of
// one wavefront execute together on the GPU).
namespace {
struct WaveBar {
    std::mutex mu;
    std::condition_variable cv;
    int count = 0;
    long gen = 0;
};
WaveBar g_wave[16];
}  // namespace

void host_wave_barrier(int)
{
    WaveBar& b = g_wave[g_tid >> 6];
    std::unique_lock<std::mutex> lk(b.mu);
    const long gen = b.gen;
    if (++b.count == 64) {
        b.count = 0;
        ++b.gen;
        b.cv.notify_all();
    } else {
        b.cv.wait(lk, [&] { return b.gen != gen; });
    }
}

// lane_swap for the calling thread (lane l of its wavefront) through a buffer its
// wavefront shares: every lane publishes its pair, then the lane with bit lb set
// takes the partner's y into x and the lane with it clear the partner's x into y.
namespace {
cplx g_lane_pair[16][64][2];
}  // namespace

void host_lane_swap(int lb, cplx& x, cplx& y)
{
    const int w = g_tid >> 6, l = g_tid & 63, p = l ^ (1 << lb);
    g_lane_pair[w][l][0] = x;
    g_lane_pair[w][l][1] = y;
    host_wave_barrier(g_tid);
    if (l & (1 << lb)) x = g_lane_pair[w][p][1];
    else y = g_lane_pair[w][p][0];
    host_wave_barrier(g_tid);
}

}  // namespace fdas

using namespace fdas;

// The ring's exchanges with the leading barrier left out: a planted mistake, so
// the test can show that the check sees a missing barrier.
template <typename F, int NA, int NR, int P>
static void passes_ring_nolead(cplx (*v)[F::E], cplx* lds, const cplx* tw, int t)
{
    if constexpr (P < F::NPASS) {
        for (int a = 0; a < NA; ++a) {
            cplx* reg = lds + (((P - 1) * NA + a) % NR) * F::LDS_ELEMS;
            F::template store<P - 1>(v[a], reg, t);
            block_sync();
            F::template load<P>(v[a], reg, t);
        }
        for (int a = 0; a < NA; ++a) F::template compute<P>(v[a], t, tw);
        passes_ring_nolead<F, NA, NR, P + 1>(v, lds, tw, t);
    }
}

// calls x NA transforms of length N (sign DIR) by T threads; in/out
// [calls][NA][N] interleaved complex.  FORM 2: a ring of NR slots of K regions
// (run_ring); 0: a region per transform (run, run_multi); 1: one region shared
// by the NA transforms (run_shared).
template <int N, int T, int DIR, int NA, int NR, int K, bool PLANT, int FORM = 2>
static void ring_run(int calls, const float* in, float* out)
{
    using F = BlockFFT<N, T, DIR>;
    constexpr int E = F::E;
    constexpr int NREG = FORM == 2 ? NR * K : FORM == 1 ? 1 : NA;
    std::vector<cplx> tw(F::TW_COUNT > 0 ? F::TW_COUNT : 1), lds((size_t)NREG * F::LDS_ELEMS);
    F::build_twiddles(tw.data());
    g_total = T;
    g_count = 0;
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([&, t] {
            g_tid = t;
            for (int c = 0; c < calls; ++c) {
                cplx v[NA][E];
                for (int a = 0; a < NA; ++a)
                    for (int i = 0; i < E; ++i) {
                        const size_t n = ((size_t)c * NA + a) * N + F::in_index(t, i);
                        v[a][i] = mk(in[2 * n], in[2 * n + 1]);
                    }
                if constexpr (PLANT) {
                    for (int a = 0; a < NA; ++a) F::template compute<0>(v[a], t, tw.data());
                    passes_ring_nolead<F, NA, NR, 1>(v, lds.data(), tw.data(), t);
                } else if constexpr (FORM == 1) {
                    F::template run_shared<NA>(v, lds.data(), tw.data(), t);
                } else if constexpr (FORM == 0 && NA == 1) {
                    F::run(v[0], lds.data(), tw.data(), t);
                } else if constexpr (FORM == 0) {
                    F::template run_multi<NA>(v, lds.data(), tw.data(), t);
                } else {
                    F::template run_ring<NA, NR, K>(v, lds.data(), tw.data(), t);
                }
                for (int a = 0; a < NA; ++a)
                    for (int i = 0; i < E; ++i) {
                        const size_t n = ((size_t)c * NA + a) * N + F::out_index(t, i);
                        out[2 * n] = v[a][i].x;
                        out[2 * n + 1] = v[a][i].y;
                    }
            }
        });
    for (auto& x : th) x.join();
}

// Returns -1 for a shape not listed; lead (if not null) gets whether the ring
// issues the leading barrier for that shape.
extern "C" int ring_race(int n, int t, int dir, int na, int nr, int k, int plant, int calls,
                         int sleep_us, const float* in, float* out, int* lead)
{
    g_sleep_us = sleep_us;
#define CASE(NN, TT, NA, NR, KK)                                                                 \
    if (n == NN && t == TT && na == NA && nr == NR && k == KK) {                                  \
        if (lead) *lead = BlockFFT<NN, TT, 1>::ring_lead<NA, NR, KK>();                           \
        if (plant) {                                                                              \
            if (dir < 0) ring_run<NN, TT, -1, NA, NR, KK, true>(calls, in, out);                  \
            else ring_run<NN, TT, +1, NA, NR, KK, true>(calls, in, out);                          \
        } else {                                                                                  \
            if (dir < 0) ring_run<NN, TT, -1, NA, NR, KK, false>(calls, in, out);                 \
            else ring_run<NN, TT, +1, NA, NR, KK, false>(calls, in, out);                         \
        }                                                                                         \
        return 0;                                                                                 \
    }
    // the library's shapes: one transform at powers of two, R transforms of
    // M = L/R at the lengths R * 2^k (K transforms exchanged per barrier)
    CASE(2048, 256, 1, 2, 1)
    CASE(4096, 512, 1, 2, 1)
    CASE(8192, 512, 1, 2, 1)
    CASE(1024, 128, 3, 2, 1)
    CASE(1024, 256, 3, 2, 1)
    CASE(2048, 256, 3, 2, 1)
    CASE(1024, 128, 5, 2, 1)
    CASE(2048, 256, 5, 2, 1)
    CASE(1024, 256, 7, 2, 1)
    CASE(1024, 128, 7, 2, 1)
    CASE(1024, 256, 7, 2, 2)
    CASE(4096, 512, 3, 2, 1)
#undef CASE
    return -1;
}

// The wavefront-split transform (WaveFFT) over successive calls, alternating
// between two LDS sets as the correlation kernel does (one workgroup barrier per
// call, wavefront barriers inside each wavefront's passes); PLANT reuses one set
// for every call, which must race.  in/out [calls][N].
template <int N, int T, int DIR, bool PLANT>
static void wave_run(int calls, const float* in, float* out)
{
    using WF = WaveFFT<N, T, DIR>;
    constexpr int E = WF::E;
    std::vector<cplx> tw(WF::TW_COUNT), lds(2 * (size_t)WF::SET);
    WF::build_twiddles(tw.data());
    g_total = T;
    g_count = 0;
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([&, t] {
            g_tid = t;
            for (int c = 0; c < calls; ++c) {
                cplx v[E];
                for (int i = 0; i < E; ++i) {
                    const size_t n = (size_t)c * N + WF::in_index(t, i);
                    v[i] = mk(in[2 * n], in[2 * n + 1]);
                }
                WF::run(v, lds.data() + (PLANT ? 0 : (c & 1) * WF::SET), tw.data(), t);
                for (int i = 0; i < E; ++i) {
                    const size_t n = (size_t)c * N + WF::out_index(t, i);
                    out[2 * n] = v[i].x;
                    out[2 * n + 1] = v[i].y;
                }
            }
        });
    for (auto& x : th) x.join();
}

extern "C" int wave_race(int n, int t, int dir, int plant, int calls, int sleep_us, const float* in,
                         float* out)
{
    g_sleep_us = sleep_us;
#define CASE(NN, TT)                                                                  \
    if (n == NN && t == TT) {                                                         \
        if (plant) {                                                                  \
            if (dir < 0) wave_run<NN, TT, -1, true>(calls, in, out);                  \
            else wave_run<NN, TT, +1, true>(calls, in, out);                          \
        } else {                                                                      \
            if (dir < 0) wave_run<NN, TT, -1, false>(calls, in, out);                 \
            else wave_run<NN, TT, +1, false>(calls, in, out);                         \
        }                                                                             \
        return 0;                                                                     \
    }
    CASE(2048, 128)
    CASE(4096, 256)
    CASE(4096, 512)
    CASE(8192, 512)
#undef CASE
    return -1;
}

// The other exchange forms (form 0: a region per transform, 1: one shared
// region), two barriers per exchange, at the shapes the library uses them.
extern "C" int form_race(int n, int t, int dir, int na, int form, int calls, int sleep_us,
                         const float* in, float* out)
{
    g_sleep_us = sleep_us;
#define CASE(NN, TT, NA, FF)                                                                      \
    if (n == NN && t == TT && na == NA && form == FF) {                                           \
        if (dir < 0) ring_run<NN, TT, -1, NA, 2, 1, false, FF>(calls, in, out);                   \
        else ring_run<NN, TT, +1, NA, 2, 1, false, FF>(calls, in, out);                           \
        return 0;                                                                                 \
    }
    CASE(1024, 64, 1, 0)
    CASE(8192, 512, 1, 0)
    CASE(16384, 512, 1, 0)
    CASE(256, 64, 3, 1)
    CASE(256, 64, 5, 1)
    CASE(512, 128, 5, 1)
    CASE(512, 64, 3, 0)
    CASE(512, 128, 7, 0)
#undef CASE
    return -1;
}
