// whiten_runs.hpp -- the whitening block schedule as a table of runs.
//
// source/bpd_noise.c:33-88 walks the spectrum in median blocks: the first starts
// at bin 1 with 6 bins, each later one starts where the previous ended and is
// (int)(6 ln start) bins long (at most 200), for as long as start + length stays
// below the bin count.  The length changes slowly, so consecutive blocks form
// runs of equal length (87 runs up to 2^26 bins): a run table that small travels
// by value in kernel arguments, and a block's start and length follow from its
// run.  Built with g++ too (tests/whiten_runs_sim.cpp), so the schedule is
// checked against the plain loop on the host.
#pragma once

#ifndef __HIPCC__
#ifndef __host__
#define __host__
#define __device__
#define __forceinline__ inline
#endif
#endif
#include <math.h>

namespace fdas {

constexpr int kWhitenMaxRuns = 128;

struct WhitenRuns {
    unsigned nbins;
    int nblocks;
    int nruns;
    int first[kWhitenMaxRuns + 1];   // first block of run r; first[nruns] = nblocks
    int start[kWhitenMaxRuns];       // start bin of that block
    int len[kWhitenMaxRuns];         // length of every block of run r
};

// (int)(6 * log(binnum)) capped at 200, as the reference evaluates it.
inline int whiten_block_len(long long binnum)
{
    const int b = (int)(6 * log((double)binnum));
    return b > 200 ? 200 : b;
}

// The reference's loop with each run found by a doubling search and bisection
// over the block count: "block i of the run has the run's length and passes the
// loop condition" holds for i < c and fails from c on, since the length never
// decreases.  Returns 0, or -1 if the table would exceed kWhitenMaxRuns.
inline int whiten_runs(unsigned nbins, WhitenRuns* w)
{
    w->nbins = nbins;
    w->nruns = 1;
    w->first[0] = 0;
    w->start[0] = 1;
    w->len[0] = 6;
    int k = 1;
    long long s = 7;
    for (;;) {
        const int l = whiten_block_len(s);
        if (s + l >= (long long)nbins) break;
        const long long s0 = s;
        auto in_run = [&](long long i) {
            const long long b = s0 + i * l;
            return whiten_block_len(b) == l && b + l < (long long)nbins;
        };
        long long good = 0, bad = 1;
        while (in_run(bad)) {
            good = bad;
            bad *= 2;
        }
        while (bad - good > 1) {
            const long long mid = (good + bad) / 2;
            if (in_run(mid)) good = mid;
            else bad = mid;
        }
        if (w->nruns == kWhitenMaxRuns) return -1;
        w->first[w->nruns] = k;
        w->start[w->nruns] = (int)s;
        w->len[w->nruns] = l;
        ++w->nruns;
        k += (int)bad;
        s += bad * l;
    }
    w->first[w->nruns] = k;
    w->nblocks = k;
    return 0;
}

// Start and length of block b (0 <= b < nblocks): its run by bisection over first[].
__host__ __device__ __forceinline__ void whiten_run_block(const WhitenRuns& w, int b, int& start,
                                                          int& len)
{
    int lo = 0, hi = w.nruns;   // first[lo] <= b < first[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (w.first[mid] <= b) lo = mid;
        else hi = mid;
    }
    len = w.len[lo];
    start = w.start[lo] + (b - w.first[lo]) * len;
}

}  // namespace fdas
