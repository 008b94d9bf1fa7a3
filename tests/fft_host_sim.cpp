// Host simulation of the workgroup FFT in gpu_accelsearch_amd/csrc/fft_block.hpp.
// Each phase (compute / LDS store / LDS load) runs for every simulated thread in
// turn, which is what the barriers between phases guarantee on the GPU, so the
// Stockham index arithmetic and the twiddle tables are checked without a GPU.
// Built by tests/test_fft_host_sim.py with g++ and called through ctypes.
#include "../gpu_accelsearch_amd/csrc/fft_block.hpp"

#include <utility>
#include <vector>

using namespace fdas;

template <int N, int T, int DIR, int P>
static void run_rest(std::vector<cplx>& v, std::vector<cplx>& lds, const std::vector<cplx>& tw)
{
    using F = BlockFFT<N, T, DIR>;
    if constexpr (P < F::NPASS) {
        for (int t = 0; t < T; ++t) F::template store<P - 1>(&v[(size_t)t * F::E], lds.data(), t);
        for (int t = 0; t < T; ++t) F::template load<P>(&v[(size_t)t * F::E], lds.data(), t);
        for (int t = 0; t < T; ++t) F::template compute<P>(&v[(size_t)t * F::E], t, tw.data());
        run_rest<N, T, DIR, P + 1>(v, lds, tw);
    }
}

template <int N, int T, int DIR>
static void sim(const float* in, float* out)
{
    using F = BlockFFT<N, T, DIR>;
    std::vector<cplx> tw(F::TW_COUNT > 0 ? F::TW_COUNT : 1), lds(F::LDS_ELEMS);
    F::build_twiddles(tw.data());
    std::vector<cplx> v((size_t)T * F::E);
    for (int t = 0; t < T; ++t)
        for (int i = 0; i < F::E; ++i) {
            int n = F::in_index(t, i);
            v[(size_t)t * F::E + i] = mk(in[2 * n], in[2 * n + 1]);
        }
    for (int t = 0; t < T; ++t) F::template compute<0>(&v[(size_t)t * F::E], t, tw.data());
    run_rest<N, T, DIR, 1>(v, lds, tw);
    for (int t = 0; t < T; ++t)
        for (int i = 0; i < F::E; ++i) {
            int n = F::out_index(t, i);
            out[2 * n] = v[(size_t)t * F::E + i].x;
            out[2 * n + 1] = v[(size_t)t * F::E + i].y;
        }
}

// Length R*M (R odd): the R length-M transforms run in lock step (one LDS
// region each), around the radix-R step.  Forward reads natural order and
// writes the decimated order X[R m + r] at r*M + m; inverse reads V[R n + r] and
// writes y in natural order.  The result is returned in natural order either way.
template <int R, int M, int T, int DIR, int P>
static void run_rest_odd(std::vector<cplx>& v, std::vector<cplx>& lds, const std::vector<cplx>& tw)
{
    using F = BlockFFT<M, T, DIR>;
    constexpr int E = F::E;
    if constexpr (P < F::NPASS) {
        for (int t = 0; t < T; ++t)
            for (int a = 0; a < R; ++a)
                F::template store<P - 1>(&v[((size_t)t * R + a) * E], lds.data() + a * F::LDS_ELEMS, t);
        for (int t = 0; t < T; ++t)
            for (int a = 0; a < R; ++a)
                F::template load<P>(&v[((size_t)t * R + a) * E], lds.data() + a * F::LDS_ELEMS, t);
        for (int t = 0; t < T; ++t)
            for (int a = 0; a < R; ++a) F::template compute<P>(&v[((size_t)t * R + a) * E], t, tw.data());
        run_rest_odd<R, M, T, DIR, P + 1>(v, lds, tw);
    }
}

template <int R, int M, int T, int DIR>
static void sim_odd(const float* in, float* out)
{
    using F = BlockFFT<M, T, DIR>;
    using RO = RadixOdd<R, M, T>;
    constexpr int E = F::E;
    std::vector<cplx> tw(F::TW_COUNT > 0 ? F::TW_COUNT : 1), lds(R * F::LDS_ELEMS), twr((R - 1) * M);
    F::build_twiddles(tw.data());
    RO::build_twiddles(twr.data());
    std::vector<cplx> v((size_t)T * R * E);
    auto at = [&](int t, int a, int i) -> cplx& { return v[((size_t)t * R + a) * E + i]; };
    for (int t = 0; t < T; ++t)
        for (int a = 0; a < R; ++a)
            for (int i = 0; i < E; ++i) {
                const int n = F::in_index(t, i);
                const int src = DIR < 0 ? n + M * a : R * n + a;
                at(t, a, i) = mk(in[2 * src], in[2 * src + 1]);
            }
    if (DIR < 0)
        for (int t = 0; t < T; ++t) RO::fwd_split(reinterpret_cast<cplx(*)[E]>(&at(t, 0, 0)), t, twr.data());
    for (int t = 0; t < T; ++t)
        for (int a = 0; a < R; ++a) F::template compute<0>(&at(t, a, 0), t, tw.data());
    run_rest_odd<R, M, T, DIR, 1>(v, lds, tw);
    if (DIR > 0)
        for (int t = 0; t < T; ++t) RO::inv_combine(reinterpret_cast<cplx(*)[E]>(&at(t, 0, 0)), t, twr.data());
    for (int t = 0; t < T; ++t)
        for (int a = 0; a < R; ++a)
            for (int i = 0; i < E; ++i) {
                const int k = F::out_index(t, i);
                const int dst = DIR < 0 ? R * k + a : k + M * a;
                out[2 * dst] = at(t, a, i).x;
                out[2 * dst + 1] = at(t, a, i).y;
            }
}

extern "C" int fft_odd_sim(int r, int m, int t, int dir, const float* in, float* out)
{
#define CASE(RR, MM, TT)                                      \
    if (r == RR && m == MM && t == TT) {                      \
        if (dir < 0) sim_odd<RR, MM, TT, -1>(in, out);        \
        else sim_odd<RR, MM, TT, +1>(in, out);                \
        return 0;                                             \
    }
    CASE(3, 128, 64)
    CASE(3, 256, 64)
    CASE(3, 512, 64)
    CASE(3, 1024, 128)
    CASE(3, 1024, 64)
    CASE(3, 2048, 256)
    CASE(3, 2048, 128)
    CASE(5, 128, 64)
    CASE(5, 256, 64)
    CASE(5, 1024, 128)
    CASE(5, 1024, 256)
    CASE(7, 256, 64)
    CASE(7, 512, 64)
    CASE(7, 1024, 256)
    CASE(7, 1024, 128)
#undef CASE
    return -1;
}

// Lane-swap policy that records the register pairs BlockFFT::lane_exchange_last
// exchanges (the sequence is the same in every lane); apply_lane_swaps then makes
// each exchange across the 64 lanes of every wavefront at once, with the device's
// semantics (lane_swap): the lane with the bit set takes the partner's y into x,
// the lane with it clear the partner's x into y.
struct SwapOp {
    int lb, xi, yi;
};
struct RecordSwap {
    std::vector<SwapOp>* ops;
    const cplx* base;
    template <int LB>
    void swap(cplx& x, cplx& y) const
    {
        ops->push_back({LB, (int)(&x - base), (int)(&y - base)});
    }
};

template <int E>
static void apply_lane_swaps(const std::vector<SwapOp>& ops, std::vector<cplx>& v, int T)
{
    for (const SwapOp& op : ops)
        for (int t = 0; t < T; ++t)
            if (t & (1 << op.lb)) {
                const int p = t ^ (1 << op.lb);
                std::swap(v[(size_t)t * E + op.xi], v[(size_t)p * E + op.yi]);
            }
}

// Wavefront-split transform (WaveFFT): inputs in the decimated order
// x[w + W (lane + 64 i)], outputs in natural order; each phase for every thread in
// turn (the wavefront passes, the cross-wavefront store, then the loads).  As on
// the device, the last wavefront exchange is made in registers where
// FW::LANE_LAST holds.
template <int N, int T, int DIR, int P>
static void run_rest_wave(std::vector<cplx>& v, std::vector<cplx>& set, const std::vector<cplx>& tw)
{
    using WF = WaveFFT<N, T, DIR>;
    using FW = typename WF::FW;
    constexpr int E = WF::E;
    if constexpr (P < FW::NPASS) {
        if constexpr (FW::LANE_LAST && P == FW::NPASS - 1) {
            std::vector<SwapOp> ops;
            cplx probe[E];
            FW::lane_exchange_last(probe, RecordSwap{&ops, probe});
            apply_lane_swaps<E>(ops, v, T);
        } else {
            for (int t = 0; t < T; ++t)
                FW::template store<P - 1>(&v[(size_t)t * E], set.data() + (t >> 6) * WF::BAND, t & 63);
            for (int t = 0; t < T; ++t)
                FW::template load<P>(&v[(size_t)t * E], set.data() + (t >> 6) * WF::BAND, t & 63);
        }
        for (int t = 0; t < T; ++t) FW::template compute<P>(&v[(size_t)t * E], t & 63, tw.data());
        run_rest_wave<N, T, DIR, P + 1>(v, set, tw);
    }
}

template <int N, int T, int DIR>
static void sim_wave(const float* in, float* out)
{
    using WF = WaveFFT<N, T, DIR>;
    using FW = typename WF::FW;
    constexpr int E = WF::E;
    std::vector<cplx> tw(WF::TW_COUNT), set(WF::SET);
    WF::build_twiddles(tw.data());
    std::vector<cplx> v((size_t)T * E);
    for (int t = 0; t < T; ++t)
        for (int i = 0; i < E; ++i) {
            const int n = WF::in_index(t, i);
            v[(size_t)t * E + i] = mk(in[2 * n], in[2 * n + 1]);
        }
    for (int t = 0; t < T; ++t) FW::template compute<0>(&v[(size_t)t * E], t & 63, tw.data());
    run_rest_wave<N, T, DIR, 1>(v, set, tw);
    for (int t = 0; t < T; ++t) WF::cross_store(&v[(size_t)t * E], set.data(), tw.data(), t);
    for (int t = 0; t < T; ++t) WF::cross_finish(&v[(size_t)t * E], set.data(), t);
    for (int t = 0; t < T; ++t)
        for (int i = 0; i < E; ++i) {
            const int n = WF::out_index(t, i);
            out[2 * n] = v[(size_t)t * E + i].x;
            out[2 * n + 1] = v[(size_t)t * E + i].y;
        }
}

extern "C" int fft_wave_sim(int n, int t, int dir, const float* in, float* out)
{
#define CASE(NN, TT)                                          \
    if (n == NN && t == TT) {                                 \
        if (dir < 0) sim_wave<NN, TT, -1>(in, out);           \
        else sim_wave<NN, TT, +1>(in, out);                   \
        return 0;                                             \
    }
    CASE(2048, 128)
    CASE(4096, 256)
    CASE(4096, 512)
    CASE(8192, 512)
    CASE(16384, 1024)
#undef CASE
    return -1;
}

// The (N, T) shapes the library instantiates.
extern "C" int fft_sim(int n, int t, int dir, const float* in, float* out)
{
#define CASE(NN, TT)                                          \
    if (n == NN && t == TT) {                                 \
        if (dir < 0) sim<NN, TT, -1>(in, out);                \
        else sim<NN, TT, +1>(in, out);                        \
        return 0;                                             \
    }
    CASE(64, 16)
    CASE(128, 32)
    CASE(128, 64)
    CASE(256, 64)
    CASE(512, 64)
    CASE(1024, 64)
    CASE(1024, 128)
    CASE(2048, 128)
    CASE(2048, 256)
    CASE(4096, 256)
    CASE(4096, 512)
    CASE(8192, 256)
    CASE(8192, 512)
    CASE(8192, 1024)
    CASE(16384, 512)
    CASE(16384, 1024)
#undef CASE
    return -1;
}
