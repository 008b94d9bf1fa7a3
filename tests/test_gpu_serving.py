"""Serving: requests of different lengths, zmax, harmonic counts and precisions
through the continuous-batching loop.  A request's candidates must be identical
(bit for bit, powers included) alone, in a concurrent mix, and through a plain
single-series plan; admission must respect the memory budget."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

# (n, zmax, numharm, fftlen, precision, r0, z0)
REQUESTS = [
    (1 << 19, 50, 4, 2048, "fp32", 20000.3, 12.0),
    (300_000, 50, 2, 2048, "bf16", 9000.7, -30.0),
    (1 << 20, 100, 8, 4096, "fp32", 40000.2, 60.0),
    (700_002, 200, 8, 8192, "fp16", 30000.6, -120.0),
    (1 << 19, 50, 16, 2048, "fp32", 7000.1, 8.0),
    (500_000, 200, 4, 8192, "bf16", 15000.4, 150.0),
]
POWCUT = (30.0, 38.0, 48.0, 62.0, 80.0)
CHUNK = 64


@pytest.fixture(scope="module")
def fd():
    import gpu_accelsearch_amd as fd
    fd.lib()
    return fd


@pytest.fixture(scope="module")
def series():
    from gpu_accelsearch_amd import synth
    return [synth.plain_series(n, seed=70 + i, r0=r0, z0=z0, power=110.0)
            for i, (n, _, _, _, _, r0, z0) in enumerate(REQUESTS)]


def _params(fd, req):
    n, zmax, numharm, fftlen, prec, _, _ = req
    return fd.SearchParams(n=n, zmax=zmax, numharm=numharm, fftlen=fftlen, precision=prec,
                           powcut=POWCUT)


def _key(c):
    return [(int(a["trial"]), int(a["step"]), int(a["z_index"]), int(a["r_index"]),
             int(a["stage"]), float(a["power"])) for a in c]


def test_alone_mix_and_plan_agree(fd, series):
    print_case("serving alone")
    with fd.Server(memory_bytes=16 << 30, max_active=4, chunk_steps=CHUNK) as srv:
        alone = []
        for req, x in zip(REQUESTS, series):
            info, c = srv.wait(srv.submit(_params(fd, req), x))
            assert info["status"] == 0 and not info["overflowed"]
            alone.append(_key(c))
        assert all(len(a) > 0 for a in alone)
        print_case("serving mix")
        srv.reset_stats()
        ids = [srv.submit(_params(fd, req), x) for req, x in zip(REQUESTS, series)]
        mixed = [_key(srv.wait(i)[1]) for i in ids]
        st = srv.stats()
        assert st["completed"] == len(REQUESTS) and st["failed"] == 0
        assert st["max_active_seen"] >= 2, st
        assert st["memory_peak"] <= st["memory_bytes"]
        assert 0 < st["p50_ms"] <= st["p99_ms"]
        for a, m in zip(alone, mixed):
            assert a == m
    print_case("serving vs plan")
    for req, x, a in zip(REQUESTS, series, alone):
        with fd.AccelSearch(_params(fd, req)) as s:
            s.search(x)
            assert _key(s.candidates()) == a, req


def test_memory_admission(fd, series):
    """A budget that holds one of these requests at a time serialises them."""
    req = REQUESTS[0]
    p = _params(fd, req)
    p.chunk_steps, p.max_candidates = CHUNK, 65536
    with fd.AccelSearch(p) as s:
        foot = s.info.device_bytes
    print_case("serving admission, footprint", foot)
    budget = int(foot * 1.6)
    with fd.Server(memory_bytes=budget, max_active=4, chunk_steps=CHUNK) as srv:
        ids = [srv.submit(_params(fd, req), series[0]) for _ in range(4)]
        outs = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["max_active_seen"] == 1, st
    assert st["memory_peak"] <= budget
    first = _key(outs[0][1])
    assert len(first) > 0
    for info, c in outs:
        assert info["device_bytes"] <= budget
        assert _key(c) == first
    # later requests waited for memory: queue time grows along the queue
    q = [info["queue_ms"] for info, _ in outs]
    assert q[-1] > q[0]


def test_failed_requests_do_not_stop_the_loop(fd, series):
    print_case("serving failures")
    x = series[0]
    with fd.Server(memory_bytes=1 << 30, max_active=2, chunk_steps=CHUNK) as srv:
        bad = srv.submit(fd.SearchParams(n=len(x), zmax=51, numharm=4, fftlen=2048), x)
        huge = np.zeros(1 << 26, np.float32)
        big = srv.submit(fd.SearchParams(n=len(huge), zmax=200, numharm=8, fftlen=8192,
                                         chunk_steps=4096), huge)
        good = srv.submit(_params(fd, REQUESTS[0]), x)
        with pytest.raises(fd.FdasError, match="zmax"):
            srv.wait(bad)
        with pytest.raises(fd.FdasError, match="serving memory"):
            srv.wait(big)
        info, c = srv.wait(good)
        assert info["status"] == 0 and len(c) > 0
        st = srv.stats()
        assert st["failed"] == 2 and st["completed"] == 1
