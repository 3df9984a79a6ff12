"""CPU baseline of the benchmark (test infrastructure: only bench.py's CPU
leg runs it).  The oracle -- the FP64 C restatement of the reference path,
built -O3 -march=native -- timed on the host's cores the way the reference
runs: one single-threaded process per map or per GCMC state point, P of
them side by side for the aggregate.

The worker pool is forked before the benchmark touches the GPU, so no
worker inherits device state; work items carry their inputs (numpy arrays)
and return (units of work, seconds, worker pid): the aggregate rate sums,
over the workers, each worker's units over its own compute seconds, so the
transfer of the inputs to the workers is not counted."""
import multiprocessing as mp
import os
import platform
import time

import numpy as np

from oracle import oracle as O


def host_info():
    model = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return dict(cpu_model=model, nproc=os.cpu_count(), cpus_available=avail)


def build_native():
    """Build (if stale) the -march=native oracle of this host."""
    O.lib(native=True)


def _native():
    L = O.lib(native=True)
    O._LIB = L
    return L


def map_nodes_work(args):
    """Map nodes m0..m1 of one map (kind 1 LJ, 2 Ewald) -> (nodes, seconds)."""
    kind, fw, npn, step, m0, m1 = args
    _native()
    t = time.perf_counter()
    O.build_nodes(kind, fw, npn, step, m0, m1)
    return m1 - m0, time.perf_counter() - t, os.getpid()


def gcmc_work(args):
    """One GCMC state point on the oracle: spec (maps as (npnts, step, ptr,
    vals)), fugacity, seed, start configuration, steps -> (attempted
    moves, steps, seconds)."""
    spec, fug, seed, cfg, steps = args
    _native()
    g = O.Gcmc(spec, fug, seed, maxN=2048)
    if cfg is not None and len(cfg[0]):
        g.set_config(cfg[0], cfg[1])
    a0 = sum(int(g.st.att[i]) for i in range(4))
    t = time.perf_counter()
    g.run(steps)
    dt = time.perf_counter() - t
    return sum(int(g.st.att[i]) for i in range(4)) - a0, steps, dt, os.getpid()


class Pool:
    """P forked workers (create before the first GPU call)."""

    def __init__(self, nproc):
        self.n = nproc
        self.pool = mp.get_context("fork").Pool(nproc)

    def run(self, fn, items):
        t = time.perf_counter()
        out = self.pool.map(fn, items, chunksize=1)
        return out, time.perf_counter() - t

    def close(self):
        self.pool.close()
        self.pool.join()


def aggregate_rate(res, unit=0, secs=-2):
    """Sum over the worker processes of (units / compute seconds) of the
    items each one ran (items: tuples with the units at index `unit`, the
    seconds at `secs` and the pid last)."""
    per = {}
    for r in res:
        u, t = per.get(r[-1], (0.0, 0.0))
        per[r[-1]] = (u + r[unit], t + r[secs])
    return sum(u / t for u, t in per.values() if t > 0), len(per)


def fifo_latencies(service_s, arrivals, servers):
    """Latency of each job (arrival -> done) under first-in first-out
    service by `servers` identical single-threaded servers."""
    free = np.zeros(servers)
    lat = np.zeros(len(arrivals))
    for i in np.argsort(arrivals, kind="stable"):
        k = int(np.argmin(free))
        start = max(free[k], arrivals[i])
        free[k] = start + service_s[i]
        lat[i] = free[k] - arrivals[i]
    return lat
