"""Screening service: continuous batching of map-build + GCMC requests on one
device (no counterpart in the reference, which runs one simulation per
process).

A request = one framework (orthorhombic cell, LJ atoms, optional charges),
an LJ map for single-site methane at 0.2 A (mapmaker_makemap) and a
methane GCMC at three fugacities (gcmc_dosim x 3 state points).

Execution:
  * GCMC replicas of all admitted requests live in one slot pool
    (mh_pool_*); each loop iteration enqueues one persistent launch that
    advances every occupied slot by `chunk` steps;
  * while that launch runs, the maps of waiting requests are built on a
    second context (its own HIP stream), so map builds overlap GCMC work;
  * after the launch, finished requests are read back and released, and
    requests with built maps are admitted into the freed slots;
  * admission is bounded by free slots / system entries and by an HBM
    budget (map + dense lookup copy + replica state per request);
  * maps are cached on the device by framework hash (LRU, bytes-bounded).
Latency = arrival -> results on the host.
"""
import collections
import hashlib
import time

import numpy as np

from . import constants as C
from . import gcmc as G
from . import hip as H
from . import synth
from .inputs import lj_AB

METHANE = (148.0, 3.73)
HICUT = 12.8
GRID = 0.2
HIPOT_KJ = 100.0
FUGACITIES_KPA = (100.0, 1000.0, 5000.0)


def screening_request(rid, steps=10000, framework=None):
    """Seeded synthetic request: framework `framework` (default rid; seed
    4000 + framework) with cell edges U[10, 45] A, U[100, 6000] atoms per
    cell, charged fraction U[0, 0.3] with q ~ U[-1, 1] neutralised, LJ
    types as the methane framework; the cell is replicated so that
    2 rc < L_full on every axis.  The GCMC seeds follow the request id
    (4000 + rid + 10000 j for the j-th fugacity)."""
    fid = rid if framework is None else framework
    rng = np.random.default_rng(4000 + fid)
    L = rng.uniform(10.0, 45.0, 3)
    n = int(rng.integers(100, 6001))
    xyz = rng.uniform(0.0, 1.0, (n, 3)) * L
    eps, sig, typ = synth.fw_types(rng, n)
    q = np.zeros(n)
    nq = int(rng.uniform(0.0, 0.3) * n)
    if nq > 1:
        idx = rng.choice(n, nq, replace=False)
        q[idx] = rng.uniform(-1.0, 1.0, nq)
        q[idx] -= q[idx].mean()
    rep = (np.floor(2.0 * HICUT / L) + 1).astype(int)
    shifts = np.array([[i, j, k] for i in range(rep[0]) for j in range(rep[1]) for k in range(rep[2])], dtype=float)
    full_xyz = (xyz[None, :, :] + shifts[:, None, :] * L).reshape(-1, 3)
    return dict(rid=rid, framework=fid, L=L, full=L * rep, xyz=full_xyz, type=np.tile(typ, len(shifts)),
                q=np.tile(q, len(shifts)), eps=eps, sig=sig, steps=steps,
                seeds=[4000 + rid + 10000 * j for j in range(len(FUGACITIES_KPA))])


def request_frameworks(n, popular=16, share=0.25, seed=4998):
    """Framework of each of n requests: a fraction `share` of them revisits
    one of `popular` frameworks (a screening campaign returning to its
    leading candidates), the others each have their own."""
    rng = np.random.default_rng(seed)
    fw = np.arange(n)
    pick = rng.random(n) < share
    fw[pick] = rng.integers(0, popular, int(pick.sum()))
    return [int(f) for f in fw]


def framework_key(req):
    h = hashlib.sha1()
    for a in (req["L"], req["full"], req["xyz"], req["type"], req["eps"], req["sig"]):
        h.update(np.ascontiguousarray(a).tobytes())
    h.update(("%r %r %r" % (METHANE, GRID, HICUT)).encode())
    return h.hexdigest()


def methane_spec():
    A, B = lj_AB(METHANE[1], METHANE[0])
    return dict(ref=[[0.0, 0.0, 0.0]], stype=[0], sq=[0.0], nst=1,
                lj=([float(A)], [float(B)], [HICUT * HICUT], [C.LOW_DIST_CUTOFF ** 2], [1]), T=300.0,
                moves=[("RINSERT", 1.0), ("RDELETE", 1.0), ("RTRANSLATE", 1.0)], dtrans=0.4)


def map_nodes(req):
    return int(np.prod((req["L"] / GRID).astype(int) + 1))


def footprint(req, maxN):
    """Device bytes a request holds while admitted: map (ptr + records +
    uncompacted lookup copy) + 3 replicas of state."""
    nodes = map_nodes(req)
    return nodes * (4 + 32 + 32) + 3 * maxN * 9 * 8 + 4096


class MapCache:
    """Device maps by framework hash; entries in use are pinned, the others
    are evicted least-recently-used beyond `budget` bytes."""

    def __init__(self, budget_bytes):
        self.budget = budget_bytes
        self.entries = collections.OrderedDict()  # key -> [map, bytes, pins]
        self.hits = self.misses = self.evictions = 0

    def get(self, key):
        e = self.entries.get(key)
        if e is None:
            self.misses += 1
            return None
        self.hits += 1
        self.entries.move_to_end(key)
        e[2] += 1
        return e[0]

    def put(self, key, m, nbytes):
        self.entries[key] = [m, nbytes, 1]
        self.trim()

    def unpin(self, key):
        self.entries[key][2] -= 1
        self.trim()

    def bytes(self):
        return sum(e[1] for e in self.entries.values())

    def trim(self):
        for k in list(self.entries):
            if self.bytes() <= self.budget:
                break
            m, _, pins = self.entries[k]
            if pins == 0:
                m.close()
                del self.entries[k]
                self.evictions += 1


class Server:
    def __init__(self, nslots=384, chunk=500, maxN=512, hbm_budget_gb=32.0, cache_gb=8.0, threads=256,
                 builds_per_launch=4):
        self.ctx_pool = H.Context(0)
        self.ctx_map = H.Context(0)
        self.nslots = nslots
        self.max_active = nslots // len(FUGACITIES_KPA)
        self.pool = G.Pool(self.ctx_pool, nslots, self.max_active, maxN=maxN, kmax_cap=0, max_sites=1,
                           precision=H.MH_PREC_FP32, threads=threads)
        self.chunk = chunk
        self.maxN = maxN
        self.hbm_budget = hbm_budget_gb * 1e9
        self.cache = MapCache(cache_gb * 1e9)
        self.spec = methane_spec()
        self.builds_per_launch = builds_per_launch

    def build_map(self, req):
        key = framework_key(req)
        m = self.cache.get(key)
        if m is None:
            c = self.ctx_map
            c.upload_framework(req["xyz"], req["type"], req["q"], req["L"], req["full"])
            c.set_lj(*synth.lb_table(req["eps"], req["sig"], METHANE[0], METHANE[1], HICUT))
            m = c.build_map(H.MH_MAP_LJ, H.MH_PREC_FP32, GRID, HIPOT_KJ)
            self.cache.put(key, m, map_nodes(req) * 68)
        return key, m

    def run(self, requests, arrivals, log=None):
        """Serve `requests` arriving at `arrivals` (s from the start).
        Returns per-request results and service statistics."""
        n = len(requests)
        order = np.argsort(arrivals, kind="stable")
        t0 = time.perf_counter()
        nxt = 0
        pending = collections.deque()
        prepared = collections.deque()
        active = {}  # ticket -> [index, launches left]
        results = [None] * n
        used = 0.0
        peak = 0.0
        mem_waits = 0
        launches = 0
        map_pts = 0
        steps_done = 0
        launches_needed = None
        while sum(r is not None for r in results) < n:
            now = time.perf_counter() - t0
            while nxt < n and arrivals[order[nxt]] <= now:
                pending.append(int(order[nxt]))
                nxt += 1
            launched = False
            if active:
                self.pool.launch(self.chunk)
                launched = True
                launches += 1
            # maps of waiting requests, built while the pool kernel runs
            nb = 0
            while pending and len(prepared) + len(active) < self.max_active and \
                    (nb < self.builds_per_launch or not active):
                i = pending[0]
                fp = footprint(requests[i], self.maxN)
                if used + fp > self.hbm_budget:
                    mem_waits += 1
                    break
                pending.popleft()
                miss0 = self.cache.misses
                key, m = self.build_map(requests[i])
                if self.cache.misses != miss0:
                    map_pts += map_nodes(requests[i])
                used += fp
                peak = max(peak, used)
                prepared.append((i, key, m, fp))
                nb += 1
            if launched:
                self.pool.wait()
                for t in list(active):
                    active[t][1] -= 1
                    steps_done += len(FUGACITIES_KPA) * min(self.chunk, active[t][2])
                    active[t][2] = max(0, active[t][2] - self.chunk)
                    if active[t][1] == 0:
                        i, _, _, key, fp = active.pop(t)
                        st = self.pool.status(t)
                        assert (st["remaining"] == 0).all()
                        req = requests[i]
                        results[i] = dict(rid=req["rid"], latency=time.perf_counter() - t0 - arrivals[i],
                                          N=st["N"].copy(), energy=st["energy"].copy(),
                                          mean_N=st["sumN"] / np.maximum(st["nsteps"], 1),
                                          att=st["att"].copy(), acc=st["acc"].copy(), err=st["err"].copy())
                        self.pool.release(t)
                        self.cache.unpin(key)
                        used -= fp
                        if log:
                            log("request %d done, latency %.1f ms" % (req["rid"], 1e3 * results[i]["latency"]))
            while prepared:
                i, key, m, fp = prepared[0]
                req = requests[i]
                t = self.pool.admit(self.spec, {0: (m, None)}, req["L"], req["full"], FUGACITIES_KPA,
                                    req["seeds"], req["steps"])
                if t is None:
                    break
                prepared.popleft()
                nl = -(-req["steps"] // self.chunk)
                active[t] = [i, nl, req["steps"], key, fp]
            if not active and not prepared and not pending and nxt < n:
                wait = arrivals[order[nxt]] - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
        makespan = time.perf_counter() - t0
        lat = np.array([r["latency"] for r in results])
        stats = dict(requests=n, makespan_s=makespan, p50_latency_ms=float(np.percentile(lat, 50) * 1e3),
                     p99_latency_ms=float(np.percentile(lat, 99) * 1e3), mean_latency_ms=float(lat.mean() * 1e3),
                     requests_per_s=n / makespan, trial_steps_per_s=steps_done / makespan,
                     map_pts_per_s=map_pts / makespan, launches=launches, chunk=self.chunk,
                     peak_admitted_bytes=peak, hbm_budget_bytes=self.hbm_budget, admission_mem_waits=mem_waits,
                     map_cache_hits=self.cache.hits, map_cache_misses=self.cache.misses,
                     map_cache_evictions=self.cache.evictions,
                     max_active_requests=self.max_active)
        return results, stats

    def close(self):
        for e in list(self.cache.entries.values()):
            e[0].close()
        self.cache.entries.clear()
        self.pool.close()
        self.ctx_map.close()
        self.ctx_pool.close()


def poisson_arrivals(n, rate, seed=4999):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))
