"""Benchmark: potential-map build of the charged-framework configuration on
one MI355X (libmusic_hip), with the CPU oracle timed beside it.

Workload (one step): the three maps a 3-site charged sorbate needs in a
50 A charged framework -- LJ pmap for the C probe, LJ pmap for the O probe
and the Ewald emap -- each at 0.15 A (334^3 = 37,259,704 nodes), computed in
FP32 pair arithmetic with FP64 Kahan accumulation, compacted on the device.
A map point is one grid node with all 8 values computed.

Prints ONE JSON line.  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "music-code_amd"))

METRIC = "GCMC trial moves/s + map pts/s, FP32-acc vs FP64; p50/p99 request latency"
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector, dense


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "map_ewald_real_fp32_charged_50A_hbm_traffic.txt")


def measured_traffic():
    """HBM bytes per launch of the Ewald real-space map kernel from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated)."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup(ctx, cfg):
    from musichip import synth
    ctx.upload_framework(cfg["xyz"], cfg["type"], cfg["q"], cfg["fund_ell"], cfg["full_ell"])
    nk = ctx.set_ewald(cfg["kappa_param"], cfg["kmax"], cfg["ewald_locut"])
    tabs = []
    for p in cfg["probes"]:
        tabs.append(synth.lb_table(cfg["fw_eps"], cfg["fw_sig"], p["eps"], p["sig"], cfg["hicut"]))
    return nk, tabs


def one_step(ctx, H, cfg, tabs, prec, keep=False):
    maps = []
    for tab, p in zip(tabs, cfg["probes"]):
        ctx.set_lj(*tab)
        maps.append(ctx.build_map(H.MH_MAP_LJ, prec, cfg["gridspace"], cfg["hipotcut"], keep_full=keep))
        log("  map %s built" % p["name"])
    maps.append(ctx.build_map(H.MH_MAP_EWALD, prec, cfg["gridspace"], cfg["hipotcut"], probe_q=1.0,
                              keep_full=keep))
    return maps


def cpu_baseline(cfg, tabs, seconds=15.0):
    """The oracle (-O3 -march=native, one thread) on a bounded sample of the
    same nodes: the first nodes of each of the three maps."""
    from oracle import oracle as O
    L = O.lib(native=True)
    O._LIB = L  # build_nodes uses the loaded library
    npn = (cfg["fund_ell"] / cfg["gridspace"]).astype(int) + 1
    step = cfg["fund_ell"] / (npn - 1)
    sf = O.ewald_sfactor(cfg["xyz"], cfg["q"], cfg["full_ell"], cfg["kappa_param"], cfg["kmax"],
                         cfg["ewald_locut"])
    fws = []
    for tab in tabs:
        A, B, h2, l2 = tab
        fws.append((1, dict(xyz=cfg["xyz"], type=cfg["type"], A=A, B=B, hicut2=h2, locut2=l2,
                            full_ell=cfg["full_ell"])))
    fws.append((2, dict(xyz=cfg["xyz"], q=cfg["q"], charge1=1.0, kappa=sf["kappa"], k=sf["k"],
                        s=sf["s"], si=sf["si"], full_ell=cfg["full_ell"])))
    per_node = []
    for kind, fw in fws:
        t = time.perf_counter()
        O.build_nodes(kind, fw, npn, step, 0, 64)
        per_node.append((time.perf_counter() - t) / 64)
    budget = seconds / 3.0
    total_t, total_n = 0.0, 0
    sample = []
    for (kind, fw), pn in zip(fws, per_node):
        n = int(max(64, min(200000, budget / max(pn, 1e-9))))
        t = time.perf_counter()
        O.build_nodes(kind, fw, npn, step, 0, n)
        dt = time.perf_counter() - t
        total_t += dt
        total_n += n
        sample.append(n)
    # one map point = one node of one map with its 8 values
    return dict(value=total_n / total_t, unit="map pts/s", cores=1, kind="port",
                sample="first %s nodes of the C/O pmaps and the emap (FP64 restatement, "
                       "one thread, -O3 -march=native, %.1f s)" % (sample, total_t))


def gcmc_leg(H, fp32_steps, fp64_steps, equil_steps, chains=(4, 32), n_single=8):
    """Batched isotherm: the 64 CO2-like state points advanced together in
    FP32 pair arithmetic (FP64 accumulation), one chain per state point and
    each count in `chains` of independent chains per state point (seeds
    3000 + i + 1000 c; workgroup size by the autotuner, capacity 256
    molecules per chain, so several chains share a CU), against the plain
    FP64 kernel run one state point at a time on the same equilibrated
    configurations."""
    from musichip import gcmc as G
    from musichip import workloads as W
    ctx = H.Context(0)
    cs = W.co2_isotherm_system(ctx)
    npt = len(cs["fugacities"])

    def batch(nchains, prec):
        fug = np.tile(cs["fugacities"], nchains)
        seeds = [cs["seeds"][i] + 1000 * c for c in range(nchains) for i in range(npt)]
        return G.GcmcBatch(ctx, cs["spec"], fug, seeds, maxN=256, maps=cs["maps"], precision=prec, threads=256)

    def timed(b, steps):
        s0 = b.state()
        t = time.perf_counter()
        b.run(steps)
        dt = time.perf_counter() - t
        s1 = b.state()
        return b.nrep * steps / dt, int((s1["att"] - s0["att"]).sum()) / dt, s1

    out = dict(workload="CO2-like 3-site sorbate, 64 state points 1-1e4 kPa, 300 K, C1 framework "
                        "(30 A, 768 atoms), pmaps C/O at 0.2 A, sorbate LJ 12.8 A + Ewald kappa 0.25/A "
                        "kmax 7 (nk %d half-space), moves 1:1:1:1" % (((2 * 7 + 1) ** 3 - 1) // 2))
    cfgs = None
    for nch in (1,) + tuple(chains):
        b = batch(nch, H.MH_PREC_FP32)
        t = time.perf_counter()
        b.run(equil_steps)
        log("gcmc: %d chains equilibrated %d steps in %.2f s" % (b.nrep, equil_steps, time.perf_counter() - t))
        tune = b.autotune(200)
        rate, arate, s1 = timed(b, fp32_steps)
        log("gcmc FP32 batched, %d chains (threads %d, xy tables %d): %.3e steps/s" % (b.nrep, tune[0], tune[1], rate))
        key = "fp32_batched_%dchains" % b.nrep
        out[key + "_steps_per_s"] = rate
        out[key + "_attempted_moves_per_s"] = arate
        out[key + "_tuned"] = dict(threads=tune[0], xy_tables=tune[1])
        if nch == 1:
            out["loadings"] = [int(x) for x in s1["N"]]
            cfgs = [b.config(r) for r in range(npt)]
            b32 = b
        else:
            b.close()
    # the same state points, FP64, one state point per launch
    picks = np.linspace(0, npt - 1, n_single).round().astype(int)
    tot_t, tot_steps, tot_att = 0.0, 0, 0
    tuned64 = []
    for r in picks:
        one = G.GcmcBatch(ctx, cs["spec"], cs["fugacities"][r:r + 1], cs["seeds"][r:r + 1], maxN=1024,
                          maps=cs["maps"], precision=H.MH_PREC_FP64, threads=256)
        one.set_config(0, cfgs[r][0], cfgs[r][1])
        tuned64.append(one.autotune(100)[0])
        a0 = one.state()["att"].sum()
        t = time.perf_counter()
        one.run(fp64_steps)
        tot_t += time.perf_counter() - t
        tot_steps += fp64_steps
        tot_att += int(one.state()["att"].sum() - a0)
        one.close()
    rate64 = tot_steps / tot_t
    log("gcmc FP64 single state point: %.3e steps/s" % rate64)
    out["fp64_single_state_point_steps_per_s"] = rate64
    out["fp64_single_state_point_attempted_moves_per_s"] = tot_att / tot_t
    out["fp64_single_state_point_threads"] = tuned64
    best = 0.0
    for nch in (1,) + tuple(chains):
        k = "fp32_batched_%dchains_steps_per_s" % (npt * nch)
        out["speedup_%dchains_vs_fp64_single" % (npt * nch)] = out[k] / rate64
        best = max(best, out[k])
    out["fp32_batched_steps_per_s"] = best
    # FP32 vs FP64 trial energies on the equilibrated configurations
    rng = np.random.default_rng(11)
    L = np.asarray(cs["spec"]["full_ell"])
    worst = 0.0
    b64 = G.GcmcBatch(ctx, cs["spec"], cs["fugacities"], cs["seeds"], maxN=1024, maps=cs["maps"],
                      precision=H.MH_PREC_FP64)
    for r in picks:
        b64.set_config(r, cfgs[r][0], cfgs[r][1])
        tc = rng.uniform(0, 1, (256, 3)) * L
        ta = np.stack([rng.uniform(-np.pi, np.pi, 256), np.arccos(rng.uniform(-1, 1, 256)),
                       rng.uniform(-np.pi, np.pi, 256)], 1)
        ok64, e64 = b64.trial_energy(r, tc, ta)
        ok32, e32 = b32.trial_energy(r, tc, ta)
        both = ok64 & ok32
        if both.any():
            worst = max(worst, float(np.max(np.abs(e32[both] - e64[both]) / np.maximum(np.abs(e64[both]), 1.0))))
    out["fp32_vs_fp64_trial_energy_max_rel_err"] = worst
    out["steps"] = dict(equil=equil_steps, fp32=fp32_steps, fp64_per_state_point=fp64_steps,
                        fp64_state_points=[int(x) for x in picks], chains_per_state_point=list(chains))
    b64.close()
    b32.close()
    for m in cs["map_objs"]:
        m.close()
    ctx.close()
    return out


def serving_leg(n=512, steps=10000, util=0.7, ncal=64):
    """Screening service: seeded requests (map at 0.2 A + 10k-step methane
    GCMC at 1/10/50 bar) with Poisson arrivals (seed 4999) at `util` of the
    capacity measured on a separate calibration set (all arriving at once)."""
    from musichip import serve
    cal = [serve.screening_request(100000 + r, steps) for r in range(ncal)]
    srv = serve.Server()
    _, cst = srv.run(cal, np.zeros(ncal))
    srv.close()
    rate = util * cst["requests_per_s"]
    log("serving: capacity %.1f requests/s (%d requests at once); arrival rate %.1f/s" %
        (cst["requests_per_s"], ncal, rate))
    reqs = [serve.screening_request(r, steps) for r in range(n)]
    arr = serve.poisson_arrivals(n, rate)
    srv = serve.Server()
    res, st = srv.run(reqs, arr)
    srv.close()
    st["arrival_rate_per_s"] = rate
    st["capacity_requests_per_s"] = cst["requests_per_s"]
    st["target_utilisation"] = util
    st["steps_per_request"] = steps * len(serve.FUGACITIES_KPA)
    st["failed_replicas"] = int(sum(int((r["err"] != 0).sum()) for r in res))
    log("serving: p50 %.1f ms, p99 %.1f ms, %.1f requests/s" % (st["p50_latency_ms"], st["p99_latency_ms"],
                                                             st["requests_per_s"]))
    return st


def hmc_leg(moves=20, nrep=256):
    """HMC bursts of flexible n-butane: 256 replicas, 50-200 chains each in
    the 40 A framework, one move = 10 NVE velocity-Verlet steps of 1 fs with
    Boltzmann velocities at 300 K; FP32 pair/map arithmetic with FP64
    integration against the FP64 kernel."""
    from musichip import hip as H
    from musichip import hmc as HM
    from musichip import workloads as W
    ctx = H.Context(0)
    w = W.butane_hmc_system(ctx, nrep=nrep, log=log)
    nch = np.array([len(c) for c in w["configs"]])
    out = dict(workload="flexible UA n-butane, %d replicas, 50-200 chains (placed %d..%d), 40 A framework "
                        "(768 atoms, 10 A cells), CH3/CH2 maps at 0.2 A, LJ 12.8 A, 10 VV steps x 1 fs, 300 K" %
                        (nrep, nch.min(), nch.max()), chains=[int(x) for x in nch])
    for prec, tag in ((H.MH_PREC_FP32, "fp32"), (H.MH_PREC_FP64, "fp64")):
        b = HM.HmcBatch(ctx, w["spec"], w["maps"], [5200 + i for i in range(nrep)], maxN=int(nch.max()),
                        precision=prec)
        for r, c in enumerate(w["configs"]):
            b.set_config(r, c)
        b.run(2)
        s0 = b.state()
        t = time.perf_counter()
        b.run(moves)
        dt = time.perf_counter() - t
        s1 = b.state()
        att = int((s1["att"] - s0["att"]).sum())
        acc = int((s1["acc"] - s0["acc"]).sum())
        out[tag + "_moves_per_s"] = att / dt
        out[tag + "_md_site_steps_per_s"] = float(nch.sum()) * 4 * 10 * moves / dt
        out[tag + "_acceptance"] = acc / max(att, 1)
        out[tag + "_failed_moves"] = int((s1["fail"] - s0["fail"]).sum())
        log("hmc %s: %.3e moves/s, acceptance %.3f" % (tag, att / dt, acc / max(att, 1)))
        if tag == "fp32":
            e32 = [b.energy_forces(r) for r in range(0, nrep, 32)]
            keep = [b.config(r)[0] for r in range(0, nrep, 32)]
        else:
            errs_e, errs_f = [], []
            for k, r in enumerate(range(0, nrep, 32)):
                b.set_config(r, keep[k])
                e64, F64, _ = b.energy_forces(r)
                errs_e.append(abs(e32[k][0] - e64) / max(1.0, abs(e64)))
                errs_f.append(np.abs(e32[k][1] - F64).max() / np.abs(F64).max())
            out["fp32_vs_fp64_energy_max_rel_err"] = float(max(errs_e))
            out["fp32_vs_fp64_force_max_rel_err"] = float(max(errs_f))
        b.close()
    out["speedup_fp32_vs_fp64"] = out["fp32_moves_per_s"] / out["fp64_moves_per_s"]
    for m in w["maps"].values():
        m.close()
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--gcmc-steps", type=int, default=4000)
    ap.add_argument("--gcmc-fp64-steps", type=int, default=1000)
    ap.add_argument("--gcmc-equil", type=int, default=30000)
    ap.add_argument("--no-gcmc", action="store_true")
    ap.add_argument("--gcmc-chains", default="4,32",
                    help="comma-separated chains per state point of the batched FP32 runs")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-hmc", action="store_true")
    ap.add_argument("--hmc-moves", type=int, default=20)
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this size (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the map build of one framework runs on one "
              "device (multi-GPU sharding is described in DESIGN.md, not run)", file=sys.stderr)
        sys.exit(2)
    import torch  # noqa: F401  (device memory/stream plumbing only)
    if args.copy_probe_mib:
        n = args.copy_probe_mib << 18  # float32 elements
        x = torch.ones(n, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps(dict(copy_probe_bytes=4 * n, launches=5)))
        return
    from musichip import hip as H
    from musichip import synth

    cfg = synth.config2()
    ctx = H.Context(0)
    nk, tabs = setup(ctx, cfg)
    npn = (cfg["fund_ell"] / cfg["gridspace"]).astype(int) + 1
    nodes = int(np.prod(npn))
    log("config2: %d atoms, nk=%d, grid %s = %d nodes" % (len(cfg["xyz"]), nk, list(npn), nodes))
    for w in range(args.warmup):
        for m in one_step(ctx, H, cfg, tabs, H.MH_PREC_FP32):
            m.close()
        log("warmup %d done" % w)
    ctx.set_timing(True)
    t0 = time.perf_counter()
    counts = None
    last32 = None
    for s in range(args.steps):
        maps = one_step(ctx, H, cfg, tabs, H.MH_PREC_FP32)
        counts = [list(map(int, m.counts)) for m in maps]
        if last32:
            for m in last32:
                m.close()
        last32 = maps
        log("step %d done" % s)
    t1 = time.perf_counter()
    ktimes = ctx.timings()
    ctx.set_timing(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = 3.0 * nodes / ((t1 - t0) / args.steps)
    log("FP32-acc: %.1f ms/step, %.3e map pts/s; kernel ms: %s" % (ms_per_step, value, ktimes))

    # dominant kernel roofline: Ewald real space in FP32 (documented cutoff)
    from musichip.hip import real_space_cutoff_default
    rc = real_space_cutoff_default(cfg["kappa_param"] / np.min(cfg["full_ell"]))
    vol = float(np.prod(cfg["full_ell"]))
    pairs_per_node = len(cfg["xyz"]) * (4.0 / 3.0) * np.pi * rc ** 3 / vol
    flops_launch = 85.0 * pairs_per_node * nodes
    kname = "nodes32_ewald_real"
    k_ms = ktimes[kname][0] / ktimes[kname][1]
    achieved = flops_launch / (k_ms * 1e-3) / 1e12
    roofline = dict(bound="valu", achieved=round(achieved, 2), peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=round(achieved / FP32_PEAK_TFLOPS, 4), traffic=measured_traffic(), kernel=kname,
                    kernel_ms=round(k_ms, 3),
                    algorithmic="85 FLOP x %.1f pairs/node (r < %.2f A) x %d nodes" % (pairs_per_node, rc, nodes))

    # FP64 mode: timing, then the stored FP32 values of the FP32-acc maps
    # against those of the FP64 maps (same compaction required)
    fp64 = None
    errs = {}
    if args.fp64_steps > 0:
        ctx.set_timing(True)
        t0 = time.perf_counter()
        m64 = None
        for s in range(args.fp64_steps):
            if m64:
                for m in m64:
                    m.close()
            m64 = one_step(ctx, H, cfg, tabs, H.MH_PREC_FP64)
        t1 = time.perf_counter()
        fp64 = 3.0 * nodes / ((t1 - t0) / args.fp64_steps)
        k64 = ctx.timings()
        ctx.set_timing(False)
        log("FP64: %.3e map pts/s; kernel ms: %s" % (fp64, k64))
        for nm, a, b in zip(["pmap_C", "pmap_O", "emap"], last32, m64):
            pa, va = a.get()
            pb, vb = b.get()
            both = (pa > 0) & (pb > 0)
            e = dict(same_compaction=bool(np.array_equal(pa, pb)),
                     nodes_tabulated_differently=int(np.count_nonzero((pa > 0) != (pb > 0))))
            ia, ib = pa[both] - 1, pb[both] - 1
            A = va[:, ia].astype(np.float64)
            B = vb[:, ib].astype(np.float64)
            d = np.abs(A - B)
            e["max_rel_err_V"] = float(np.max(d[0] / np.maximum(np.abs(B[0]), 1.0)))
            e["max_abs_err_V_kcal"] = float(d[0].max())
            # derivatives: relative error with the same floor of 1 (kcal/mol/A^n)
            e["max_rel_err_derivs"] = float(np.max(d[1:] / np.maximum(np.abs(B[1:]), 1.0)))
            errs[nm] = e
            log("FP32 vs FP64 %s: %s" % (nm, e))
            del pa, va, pb, vb, A, B, d
        for m in m64 + last32:
            m.close()

    gcmc = None
    if not args.no_gcmc:
        gcmc = gcmc_leg(H, args.gcmc_steps, args.gcmc_fp64_steps, args.gcmc_equil,
                        tuple(int(c) for c in str(args.gcmc_chains).split(",") if c))

    hmc = None
    if not args.no_hmc:
        hmc = hmc_leg(args.hmc_moves)

    serving = None
    if not args.no_serve:
        serving = serving_leg(args.serve_requests)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, tabs, args.cpu_seconds)
        log("CPU baseline: %s" % cpu)
    out = dict(metric=METRIC, value=value, unit="map pts/s", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=ms_per_step, higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype="f32 (f64 Kahan accumulation)", data="synthetic",
               config=dict(workload="charged-framework map build: 50 A cubic cell, 4096 atoms "
                                    "(seed 2002), pmap C + pmap O + Ewald emap at 0.15 A (334^3 nodes "
                                    "each), alpha 0.25/A, kmax 9",
                           nodes_per_map=nodes, maps_per_step=3, nk=nk, compaction_counts=counts),
               fp64_map_pts_per_s=fp64, fp32_vs_fp64=errs, kernel_ms_per_step={
                   k: v[0] / args.steps for k, v in ktimes.items()},
               gcmc=gcmc, gcmc_trial_moves_per_s=gcmc and gcmc["fp32_batched_steps_per_s"],
               p50_latency_ms=serving and serving["p50_latency_ms"],
               p99_latency_ms=serving and serving["p99_latency_ms"], serving=serving, hmc=hmc,
               roofline=roofline, cpu_baseline=cpu)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
