"""Slot pool (mh_pool_*) and the screening service: a request's GCMC in the
pool equals the same chains in a plain batch, and a request's results are
identical whether it is served alone or together with others."""
import numpy as np
import pytest

from musichip import gcmc as G
from musichip import hip as H
from musichip import serve, workloads

pytestmark = pytest.mark.gpu


def test_pool_chunks_equal_plain_batch(hip):
    print("case: methane pool (chunks of 500) vs plain FP32 batch, 3000 steps", flush=True)
    ms = workloads.methane_system(hip)
    fug, seeds = ms["fugacities"][-3:], [11, 12, 13]
    b = G.GcmcBatch(hip, ms["spec"], fug, seeds, maxN=256, maps=ms["maps"], precision=H.MH_PREC_FP32)
    b.run(3000)
    ref = b.state()
    L = ms["spec"]["full_ell"]
    p = G.Pool(hip, 16, 4, maxN=256, precision=H.MH_PREC_FP32)
    # an unrelated request first, so the methane replicas land in other slots
    t0 = p.admit(ms["spec"], ms["maps"], L, L, [300.0, 300.0], [1, 2], 1000)
    t = p.admit(ms["spec"], ms["maps"], L, L, fug, seeds, 3000)
    for _ in range(6):
        p.launch(500)
    p.wait()
    st = p.status(t)
    assert (st["remaining"] == 0).all() and list(st["slots"]) == [2, 3, 4]
    assert np.array_equal(st["N"], ref["N"])
    assert np.array_equal(st["energy"], ref["energy"])
    assert np.array_equal(st["att"], ref["att"]) and np.array_equal(st["acc"], ref["acc"])
    assert np.array_equal(st["sumN"], ref["sumN"])
    s0 = p.status(t0)
    assert (s0["remaining"] == 0).all() and (s0["nsteps"] == 1000).all()
    p.release(t0)
    p.release(t)
    # slots and system entries are reusable
    t2 = p.admit(ms["spec"], ms["maps"], L, L, fug, seeds, 500)
    p.launch(1000)
    p.wait()
    assert (p.status(t2)["nsteps"] == 500).all()
    p.close()


def test_serving_alone_equals_mix():
    print("case: screening requests, alone vs in a mix of 6", flush=True)
    reqs = [serve.screening_request(r, steps=2000) for r in range(6)]
    srv = serve.Server(nslots=12, chunk=250, maxN=256, hbm_budget_gb=4.0)
    alone, st1 = srv.run([reqs[3]], np.array([0.0]))
    srv.close()
    # a fresh server: the map of request 3 is rebuilt inside the mix
    srv = serve.Server(nslots=12, chunk=250, maxN=256, hbm_budget_gb=4.0)
    mix, st2 = srv.run(reqs, np.array([0.0, 0.0, 0.002, 0.004, 0.004, 0.01]))
    srv.close()
    assert st2["map_cache_hits"] == 0
    a, m = alone[0], mix[3]
    assert a["rid"] == m["rid"] == 3
    for k in ("N", "energy", "mean_N", "att", "acc", "err"):
        assert np.array_equal(a[k], m[k]), k
    assert st2["requests"] == 6 and st2["p99_latency_ms"] >= st2["p50_latency_ms"] > 0
    # with 12 slots at most 4 requests are admitted together
    assert st2["max_active_requests"] == 4
    print("alone %s / mix %s" % (st1, st2), flush=True)


def test_serving_map_cache_reuse_and_eviction():
    """Requests that revisit frameworks hit the device map cache; with a
    cache budget of about two maps, unpinned maps are evicted least recently
    used and rebuilt on a later miss.  Every request's results equal those
    of the same request served alone on a fresh server."""
    print("case: screening requests with repeated frameworks, small map cache", flush=True)
    fws = [0, 1, 0, 2, 1, 3, 0, 2]
    reqs = [serve.screening_request(r, steps=1000, framework=f) for r, f in enumerate(fws)]
    per_map = max(serve.map_nodes(r) for r in reqs) * 68
    srv = serve.Server(nslots=6, chunk=250, maxN=256, hbm_budget_gb=4.0, cache_gb=2.2 * per_map / 1e9)
    arr = np.array([0.0, 0.0, 0.05, 0.05, 0.1, 0.1, 0.15, 0.2])
    mix, st = srv.run(reqs, arr)
    srv.close()
    print(st, flush=True)
    assert st["map_cache_hits"] > 0 and st["map_cache_evictions"] > 0
    assert st["map_cache_hits"] + st["map_cache_misses"] == len(reqs)
    for i in (2, 5, 7):
        srv = serve.Server(nslots=6, chunk=250, maxN=256, hbm_budget_gb=4.0)
        alone, _ = srv.run([reqs[i]], np.array([0.0]))
        srv.close()
        for k in ("N", "energy", "mean_N", "att", "acc", "err"):
            assert np.array_equal(alone[0][k], mix[i][k]), (i, k)
