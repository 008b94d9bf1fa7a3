"""The serving workload's shapes through the server: a seeded Poisson-arrival mix
of requests with series of 2^20 to 2^25 samples, zmax 50, 200 and 600, 4, 8 and 16
harmonics, FP32 and the reduced-precision mode within the 1e-3 tolerance (I16),
at most 4 requests on the device at once.  Each request's result must be the same
alone and in the mix: bit-identical candidates, and for a request whose
candidates exceed its capacity the same overflow error and count.  Two steps of
the largest request (2^25 samples, zmax 600, 16 harmonics) are checked against
the oracle through fdas_run_steps."""
import threading
import time

import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

RATE = 50.0          # requests per second
CHUNK = 128
POWCUT = (40.0, 48.0, 60.0, 78.0, 100.0)
# (n, zmax, numharm, fftlen, precision, pulsar (r0, z0, first-harmonic power) or None)
C5 = [
    (1 << 25, 600, 16, 16384, "fp32", (200000.3, -150.0, 60.0)),
    (1 << 23, 200, 8, 8192, "i16", (125000.25, 20.0, 40.0)),
    (1 << 20, 50, 4, 2048, "fp32", (30000.6, 10.0, 60.0)),
    (1 << 23, 600, 16, 16384, "i16", (60000.4, 300.0, 50.0)),
    (3_000_000, 200, 16, 8192, "fp32", None),
    (1_500_002, 50, 8, 4096, "i16", (40000.1, -20.0, 50.0)),
    (1 << 21, 600, 4, 16384, "fp32", None),
    (1 << 22, 200, 16, 8192, "i16", (90000.7, 80.0, 45.0)),
    (1 << 25, 200, 8, 8192, "i16", None),
]
# a request whose threshold lets far more cells through than its capacity holds
OVERFLOW = (1 << 20, 50, 4, 2048, "fp32", None)
OVERFLOW_POWCUT = (6.0, 9.0, 14.0, 0.0, 0.0)
OVERFLOW_CAP = 64


@pytest.fixture(scope="module")
def fd():
    import gpu_accelsearch_amd as fd
    fd.lib()
    return fd


@pytest.fixture(scope="module")
def workload():
    from gpu_accelsearch_amd import synth
    pool = synth.noise_pool(1 << 25, seed=31)
    rng = np.random.default_rng(32)
    xs = []
    for n, _, _, _, _, psr in C5 + [OVERFLOW]:
        off = int(rng.integers(0, (1 << 25) - n + 1))
        x = pool[off:off + n].copy()
        if psr:
            x += synth.pulse_train(n, psr[0], psr[1], 0.05, psr[2]).astype(np.float32)
        xs.append(x)
    arrivals = np.cumsum(rng.exponential(1.0 / RATE, len(xs)))
    return xs, arrivals


def _params(fd, req, powcut=POWCUT, cap=0):
    n, zmax, numharm, fftlen, prec, _ = req
    return fd.SearchParams(n=n, zmax=zmax, numharm=numharm, fftlen=fftlen, precision=prec,
                           powcut=powcut, chunk_steps=CHUNK, max_candidates=cap)


def _all_params(fd):
    return [_params(fd, r) for r in C5] + [_params(fd, OVERFLOW, OVERFLOW_POWCUT, OVERFLOW_CAP)]


def _key(info, c):
    return (info["status"], info["ncandidates"], info["overflowed"],
            [(int(a["step"]), int(a["z_index"]), int(a["r_index"]), int(a["stage"]),
              float(a["power"])) for a in c])


def test_serving_workload_alone_and_in_the_mix(fd, workload):
    xs, arrivals = workload
    params = _all_params(fd)
    with fd.Server(memory_bytes=96 << 30, max_active=4, chunk_steps=CHUNK) as srv:
        for p in params:
            srv.prewarm(p)
        alone = []
        for i, (p, x) in enumerate(zip(params, xs)):
            print_case("C5 alone", i, p.n, p.zmax, p.numharm, p.precision)
            info, c = srv.wait(srv.submit(p, x), raise_on_error=False)
            alone.append(_key(info, c))
        srv.reset_stats()
        print_case("C5 mix, Poisson arrivals at", RATE, "per s")
        ids = [None] * len(params)
        t0 = time.perf_counter()

        def feed():
            for i, (p, x) in enumerate(zip(params, xs)):
                delay = arrivals[i] - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                ids[i] = srv.submit(p, x)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        mixed = [_key(*srv.wait(i, raise_on_error=False)) for i in ids]
        st = srv.stats()
    print("C5 stats", {k: st[k] for k in ("completed", "failed", "max_active_seen", "p50_ms",
                                          "p99_ms", "memory_peak")})
    for i, (a, m) in enumerate(zip(alone, mixed)):
        assert a == m, i
    for i, (req, a) in enumerate(zip(C5, alone)):
        assert a[0] == 0 and not a[2], (i, a[:3])
        if req[5] is not None:
            assert a[1] > 0, i
    ov = alone[-1]
    assert ov[0] == 5 and ov[2] == 1 and ov[1] > OVERFLOW_CAP and ov[3] == []
    assert st["max_active_seen"] >= 2, st
    assert st["completed"] == len(C5) and st["failed"] == 1
    assert st["memory_peak"] <= st["memory_bytes"]


def test_serving_largest_request_against_oracle(fd, oracle, workload):
    """Two steps of the 2^25-sample, zmax 600, 16-harmonic request through
    fdas_run_steps: the pulsar's step and the last step, at the FP32 1e-4 bound."""
    from gpu_accelsearch_amd import synth
    xs, _ = workload
    req = C5[0]
    n, zmax, numharm, fftlen, _, psr = req
    x = xs[0]
    with fd.AccelSearch(_params(fd, req)) as s:
        steps = [synth.steps_covering(psr[0], s.rlo, s.uselen), s.nsteps - 1]
        print_case("C5 largest request vs oracle, steps", steps)
        ref = oracle.Search(x, n, 6.4e-5, zmax, numharm, fftlen)
        assert (ref.uselen, ref.nsteps) == (s.uselen, s.nsteps)
        s.prepare(x)
        for st in steps:
            got = s.run_steps(st, 1)[0].astype(np.float64)
            want = ref.step_plane(st)
            e = float(np.max(np.abs(got - want) / np.maximum(want, 1.0)))
            print(f"C5 step {st}: max relative power error {e:.2e}, peak {want.max():.1f}")
            assert e <= 1e-4
        ref.close()
