"""Write tests/golden/ref_math.npz from the reference's own C helpers.

oracle/_ref/libbpdref.so is compiled (by oracle/Makefile) from the reference's
source/bpd_math.c and source/bpd_fresnl.c unchanged.  This script runs them on
seeded inputs and stores inputs and outputs only:
  median_*  fast_median()          (source/bpd_math.c:20-91)
  rsqrt_*   inverse_square_root()  (source/bpd_math.c:94-105)
  fresnl_*  fresnl()               (source/bpd_fresnl.c:160-200)
  pow2_*    get2PowerNGreaterThan() (source/bpd_math.c:113-122)
Run: python tests/golden/make_ref_vectors.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402


def main():
    O.build()
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libbpdref.so not built (reference sources absent)")
    rng = np.random.default_rng(20240611)

    # medians: lengths 1..220 (whitening blocks <= 200), 4096 and 8192 (segments)
    lens = list(range(1, 221)) + [1000, 4096, 4097, 8192]
    offs = np.zeros(len(lens) + 1, np.int64)
    offs[1:] = np.cumsum(lens)
    vals = np.empty(offs[-1], np.float32)
    out = np.empty(len(lens), np.float32)
    for i, n in enumerate(lens):
        a = (rng.standard_normal(n) ** 2 + rng.standard_normal(n) ** 2).astype(np.float32)
        if i % 5 == 0:
            a[::4] = a[0]          # ties
        if i % 7 == 0:
            a[: n // 3] = 0.0      # zero padding at a segment edge
        vals[offs[i]:offs[i + 1]] = a
        out[i] = R.fast_median(a.copy(), n)

    # fast inverse square root over many decades, plus subnormals/zero
    x = np.abs(rng.standard_normal(30000)).astype(np.float32)
    x *= (np.float32(10.0) ** rng.integers(-30, 30, x.size)).astype(np.float32)
    x = np.concatenate([x, np.float32([0.0, 1e-40, 1.0, 2.0, 0.6931472, 3.4e38])])
    y = np.array([R.inverse_square_root(float(v)) for v in x], np.float32)

    # Fresnel integrals
    fx = np.concatenate([np.linspace(-45, 45, 3001), rng.uniform(-2, 2, 1000),
                         [1.6, -1.6, 1.60078, 36974.5, -40000.0]])
    fs = np.empty_like(fx)
    fc = np.empty_like(fx)
    s, c = C.c_double(), C.c_double()
    for i, v in enumerate(fx):
        R.fresnl(float(v), C.byref(s), C.byref(c))
        fs[i], fc[i] = s.value, c.value

    p2in = np.array([1, 2, 3, 5, 1023, 1024, 1025, 4096 + 229, 7734 + 458, 65535], np.int64)
    p2out = np.array([R.get2PowerNGreaterThan(int(v)) for v in p2in], np.int64)

    np.savez_compressed(os.path.join(HERE, "ref_math.npz"), median_vals=vals, median_offs=offs,
                        median_out=out, rsqrt_in=x, rsqrt_out=y, fresnl_x=fx, fresnl_s=fs,
                        fresnl_c=fc, pow2_in=p2in, pow2_out=p2out)
    print("wrote", os.path.join(HERE, "ref_math.npz"))


if __name__ == "__main__":
    main()
