"""The whitening block schedule as a run table (csrc/whiten_runs.hpp), checked on
the host against the block-by-block loop of source/bpd_noise.c:33-88: the first
block starts at bin 1 with 6 bins, each later one starts where the previous
ended and is (int)(6 ln start) bins long, while start + length < nbins.

Sizes include every bin count that sits at a run boundary of the loop (where a
block just passes or just fails the loop condition) and the largest size the
allocation-free whitening accepts (2^26 bins)."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def sim(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("wr") / "whiten_runs_sim.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", so,
                    os.path.join(ROOT, "tests", "whiten_runs_sim.cpp")], check=True)
    L = C.CDLL(so)
    L.sim_whiten_blocks.restype = C.c_int
    L.sim_whiten_blocks.argtypes = [C.c_uint, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    return L


def loop_blocks(nbins):
    """The reference's loop, one block at a time."""
    starts, lens = [1], [6]
    binnum = 7
    buflen = min(200, int(6 * math.log(binnum)))
    while binnum + buflen < nbins:
        starts.append(binnum)
        lens.append(buflen)
        binnum += buflen
        buflen = min(200, int(6 * math.log(binnum)))
    return np.array(starts), np.array(lens)


def run_blocks(sim, nbins):
    cap = nbins // 6 + 8
    s = np.zeros(cap, np.int32)
    ln = np.zeros(cap, np.int32)
    nr = C.c_int()
    nb = sim.sim_whiten_blocks(nbins, s.ctypes.data, ln.ctypes.data, cap, C.byref(nr))
    assert nb > 0, nb
    return s[:nb], ln[:nb], nr.value


def boundary_sizes(limit):
    """Bin counts around each block end of the loop below limit."""
    starts, lens = loop_blocks(limit)
    ends = starts + lens
    out = set()
    for e in ends[:: max(1, len(ends) // 400)]:
        out |= {int(e) - 1, int(e), int(e) + 1}
    return sorted(x for x in out if x >= 16)


@pytest.mark.parametrize("nbins", [16, 17, 23, 100, 1000, 4097, 65536, 524288, (1 << 22) + 1234,
                                   1 << 22, 3_000_001])
def test_runs_equal_loop(sim, nbins):
    s, ln, _ = run_blocks(sim, nbins)
    ws, wl = loop_blocks(nbins)
    assert np.array_equal(s, ws) and np.array_equal(ln, wl)


def test_runs_equal_loop_at_boundaries(sim):
    for nbins in boundary_sizes(200_000):
        s, ln, _ = run_blocks(sim, nbins)
        ws, wl = loop_blocks(nbins)
        assert np.array_equal(s, ws) and np.array_equal(ln, wl), nbins


def test_runs_at_largest_size(sim):
    nbins = 1 << 26
    s, ln, nruns = run_blocks(sim, nbins)
    ws, wl = loop_blocks(nbins)
    assert len(s) == 663030 and nruns == 87
    assert np.array_equal(s, ws) and np.array_equal(ln, wl)
    # blocks tile [1, end of the last block) with no gap
    assert np.array_equal(s[1:], s[:-1] + ln[:-1])
