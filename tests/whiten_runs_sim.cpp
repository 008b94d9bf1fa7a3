// Host build of the whitening run table (gpu_accelsearch_amd/csrc/whiten_runs.hpp):
// expands the runs into every block's (start, length) exactly as the kernels do,
// so tests/test_whiten_runs.py can compare it with the reference's block loop.
// Built with g++ and called through ctypes.
#include "../gpu_accelsearch_amd/csrc/whiten_runs.hpp"

using namespace fdas;

extern "C" int sim_whiten_blocks(unsigned nbins, int* starts, int* lens, int cap, int* nruns)
{
    WhitenRuns w;
    if (whiten_runs(nbins, &w) != 0) return -1;
    *nruns = w.nruns;
    if (w.nblocks > cap) return -2;
    for (int b = 0; b < w.nblocks; ++b) whiten_run_block(w, b, starts[b], lens[b]);
    return w.nblocks;
}
