"""Index arithmetic of the direct correlation on the matrix cores
(gpu_accelsearch_amd/csrc/mma_layout.hpp), checked on the host: tests/mma_host_sim.cpp
builds the segment chunks and template B fragments as the kernels store them,
emulates the 32x32x32 int8 MFMA with the lane layout measured on the device, and
compares every kept column exactly with the direct integer correlation
c[2p + phi] = sum_j a[bo + p + j] conj(T_phi[j]) (the spread segment's correlation
with the template's taps of one parity, source/bpd_search.cu:56-67,243-307).

Cases cover every harmonic geometry of the production search (zmax 200, fftlen
8192: half-widths 24..114), zmax 300 (180, the largest supported), the smallest
templates (16), ragged row widths (not a multiple of 16 columns) and tiles whose
two templates have different half-widths."""
import ctypes as C
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def sim(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("mma") / "mma_host_sim.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", so,
                    os.path.join(ROOT, "tests", "mma_host_sim.cpp")], check=True)
    L = C.CDLL(so)
    L.mma_sim.restype = C.c_int
    L.mma_sim.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint]
    return L


@pytest.mark.parametrize("bo,W,hw0,hw1", [
    (16, 100, 16, 16),        # smallest templates, one M tile
    (24, 972, 24, 16),        # 1/8 harmonic at zmax 200, mixed tile
    (33, 1936, 33, 33),
    (56, 3870, 56, 55),
    (114, 7734, 114, 114),    # fundamental at zmax 200 (uselen 7734)
    (114, 7734, 17, 16),      # small |z| rows of the fundamental
    (98, 6772, 98, 97),       # 7/8 harmonic
    (180, 7470, 180, 180),    # zmax 300
    (40, 66, 39, 40),         # ragged width, second template larger
])
def test_mma_layout_exact(sim, bo, W, hw0, hw1):
    for seed in (1, 2):
        assert sim.mma_sim(bo, W, hw0, hw1, seed) == 0, (bo, W, hw0, hw1, seed)
