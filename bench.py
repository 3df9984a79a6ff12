"""Benchmark of the GCMC energy hot path on one MI355X (libmusic_hip), with
the CPU oracle timed beside it.

Headline (`value`): map points per second of the charged-framework map
build -- the three maps a 3-site charged sorbate needs in a 50 A charged
framework (LJ pmap for the C probe, LJ pmap for the O probe, Ewald emap),
each at 0.15 A (334^3 = 37,259,704 nodes), FP32 pair arithmetic with FP64
accumulation, compacted on the device.  A map point is one grid node with
all 8 values computed.  Beside it, one line carries:
  c1       plain FP64 single-site methane GCMC (8 state points, 20k steps);
  gcmc     the batched CO2 isotherm (64 state points), FP32 maps + FP32
           kernel against FP64, `gcmc_trial_moves_per_s` = its attempted
           moves per second, roofline by algorithmic bytes per trial;
  loading  the same kernel at seeded loadings up to the densest reached;
  hmc      flexible-butane hybrid MC (256 replicas);
  serving  512 screening requests, continuous batching, p50 / p99 latency;
  cpu_baseline  the oracle on the host's cores (single core and P-core
           aggregate; map points, C1 / C3 trial moves, a FIFO model of the
           serving stream).

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


HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def map_cpu_items(cfg, tabs, npn, step, nodes_each, parts=1):
    """Work items of the map CPU baseline: the first nodes of each of the
    three maps (LJ C, LJ O, Ewald), optionally split into `parts`."""
    from oracle import oracle as O
    sf = O.ewald_sfactor(cfg["xyz"], cfg["q"], cfg["full_ell"], cfg["kappa_param"], cfg["kmax"],
                         cfg["ewald_locut"])
    fws = []
    for tab in tabs:
        A, B, h2, l2 = tab
        fws.append((1, dict(xyz=cfg["xyz"], type=cfg["type"], A=A, B=B, hicut2=h2, locut2=l2,
                            full_ell=cfg["full_ell"])))
    fws.append((2, dict(xyz=cfg["xyz"], q=cfg["q"], charge1=1.0, kappa=sf["kappa"], k=sf["k"],
                        s=sf["s"], si=sf["si"], full_ell=cfg["full_ell"])))
    items = []
    for (kind, fw), n in zip(fws, nodes_each):
        for k in range(parts):
            m0 = k * n // parts
            items.append((kind, fw, npn, step, m0, (k + 1) * n // parts))
    return items


def gcmc_bytes_per_step(ns, nlook, nbar, nk, att, acc):
    """Algorithmic HBM bytes of trial moves: 288 B per site-map lookup
    (8 corners x (8 FP32 values + int32 ptr)), 12 B per other sorbate site
    (FP32 xyz), 16 B per k vector read (FP64 S(k)) and 16 B per k vector
    written on acceptance."""
    return att * (288.0 * nlook + 12.0 * ns * nbar + 16.0 * nk) + acc * 16.0 * nk


def c1_leg(H, steps):
    """Plain FP64 single-site methane GCMC: the 30 A framework (768 atoms),
    methane map at 0.2 A, 8 fugacities 100-5000 kPa, `steps` MC steps per
    state point, FP64 kernel; once as one batched launch of the 8 state
    points and once one state point per launch."""
    from musichip import gcmc as G
    from musichip import workloads as W
    ctx = H.Context(0)
    ms = W.methane_system(ctx)
    fug, seeds = ms["fugacities"], ms["seeds"]
    out = dict(workload="single-site methane (148 K, 3.73 A, LJ 12.8 A) in the 30 A framework (768 atoms, "
                        "seed 1001), map 0.2 A (151^3 nodes), 300 K, insert:delete:translate 1:1:1, "
                        "fugacities 100-5000 kPa (8), %d steps each, FP64" % steps)
    b = G.GcmcBatch(ctx, ms["spec"], fug, seeds, maxN=512, maps=ms["maps"], precision=H.MH_PREC_FP64)
    th = b.autotune(200)[0]
    s0 = b.state()
    t = time.perf_counter()
    b.run_cycles(steps // 100, 100)
    dt = time.perf_counter() - t
    s1 = b.state()
    out["fp64_batched_8pt_steps_per_s"] = 8 * steps / dt
    out["fp64_batched_8pt_attempted_moves_per_s"] = int((s1["att"] - s0["att"]).sum()) / dt
    out["fp64_batched_threads"] = th
    out["mean_loading"] = [float(x) for x in (s1["sumN"] - s0["sumN"]) / steps]
    cfgs = [b.config(r) for r in range(8)]
    b.close()
    # the same system with cavity-biased insertion / deletion (CAVITY_BIAS,
    # 20 trials per insertion): 1 A cubes tabulated from the methane map
    # below 0 kcal/mol, in every unit cell of the simulation cell
    full, fund = ms["spec"]["full_ell"], ms["spec"]["fund_ell"]
    ncell = tuple(int(round(a / b_)) for a, b_ in zip(full, fund))
    ncub = tuple(int(round(a)) // c * c for a, c in zip(full, ncell))
    tab = G.cavity_tab(ms["map_objs"][0], 0.0, ncub, ncell, full)
    cspec = dict(ms["spec"], moves=[("BINSERT", 1.0), ("BDELETE", 1.0), ("RTRANSLATE", 1.0)])
    cb = G.GcmcBatch(ctx, cspec, fug, seeds, maxN=512, maps=ms["maps"], precision=H.MH_PREC_FP64)
    cb.set_cavity(ncub, [20], [0], tab=tab, species_mask=1)
    cb.run(steps // 4)
    c0 = cb.state()
    t = time.perf_counter()
    cb.run(steps)
    dt = time.perf_counter() - t
    c1s = cb.state()
    cb.close()
    out["cavity_bias"] = dict(
        workload="the same 8 state points with CAVITY_BIAS BINSERT / BDELETE (20 trials), translate 1:1:1, "
                 "cubes %s tabulated from the map below 0 kcal/mol (%d of %d), %d steps after %d" %
                 ("x".join(str(v) for v in ncub), int(tab.sum()), tab.size, steps, steps // 4),
        fp64_batched_8pt_steps_per_s=8 * steps / dt,
        fp64_batched_8pt_attempted_moves_per_s=int((c1s["att"] - c0["att"]).sum()) / dt,
        mean_loading=[float(x) for x in (c1s["sumN"] - c0["sumN"]) / steps])
    log("c1 cavity bias (20 trials): batched %.3e moves/s" %
        out["cavity_bias"]["fp64_batched_8pt_attempted_moves_per_s"])
    tot, att = 0.0, 0
    for r in range(8):
        one = G.GcmcBatch(ctx, ms["spec"], fug[r:r + 1], seeds[r:r + 1], maxN=512, maps=ms["maps"],
                          precision=H.MH_PREC_FP64)
        a0 = one.state()["att"].sum()
        t = time.perf_counter()
        one.run(steps)
        tot += time.perf_counter() - t
        att += int(one.state()["att"].sum() - a0)
        one.close()
    out["fp64_single_state_point_steps_per_s"] = 8 * steps / tot
    out["fp64_single_state_point_attempted_moves_per_s"] = att / tot
    log("c1: batched %.3e moves/s, single state point %.3e moves/s" %
        (out["fp64_batched_8pt_attempted_moves_per_s"], out["fp64_single_state_point_attempted_moves_per_s"]))
    m = ms["map_objs"][0]
    ptr, vals = m.get()
    cpu = dict(spec=dict(ms["spec"], maps=[(m.npnts, m.step, ptr, vals)], site_pmap=[0], site_emap=[-1]),
               fug=list(fug), seeds=list(seeds), cfgs=[(c[0], c[1]) for c in cfgs])
    for mm in ms["map_objs"]:
        mm.close()
    ctx.close()
    return out, cpu


def gcmc_leg(H, fp32_steps, fp64_steps, equil_steps, chains=(4, 32), n_single=8):
    """Batched isotherm: the 64 CO2-like state points advanced together, FP32
    maps feeding the FP32 kernel (FP64 accumulation), one chain per state
    point (the configuration) and each count in `chains` of independent
    chains per state point (seeds 3000 + i + 1000 c; workgroup size by the
    autotuner, capacity 256 molecules per chain, so several chains share a
    CU), against FP64 maps and the plain FP64 kernel run one state point at
    a time on the same equilibrated configurations."""
    from musichip import gcmc as G
    from musichip import workloads as W
    ctx = H.Context(0)
    cs = W.co2_isotherm_system(ctx, map_precision=H.MH_PREC_FP32)
    cs64 = W.co2_isotherm_system(ctx, map_precision=H.MH_PREC_FP64)
    npt = len(cs["fugacities"])
    kmax = cs["spec"]["ewald"]["kmax"]
    nk = ((2 * kmax + 1) ** 3 - 1) // 2
    ns = len(cs["spec"]["ref"])

    def batch(nchains, prec, sys_):
        fug = np.tile(sys_["fugacities"], nchains)
        seeds = [sys_["seeds"][i] + 1000 * c for c in range(nchains) for i in range(npt)]
        return G.GcmcBatch(ctx, sys_["spec"], fug, seeds, maxN=256, maps=sys_["maps"], precision=prec,
                           threads=256)

    def timed(b, steps):
        s0 = b.state()
        t = time.perf_counter()
        b.run(steps)
        dt = time.perf_counter() - t
        s1 = b.state()
        att = (s1["att"] - s0["att"]).sum(1)
        acc = (s1["acc"] - s0["acc"]).sum(1)
        nbar = (s1["sumN"] - s0["sumN"]) / steps
        byt = float(sum(gcmc_bytes_per_step(ns, ns, nbar[r], nk, att[r], acc[r]) for r in range(len(att))))
        return b.nrep * steps / dt, int(att.sum()) / dt, s1, byt / dt

    out = dict(workload="CO2-like 3-site sorbate, 64 state points 1-1e4 kPa, 300 K, C1 framework "
                        "(30 A, 768 atoms), pmaps C/O at 0.2 A (FP32-built for the FP32 kernel, FP64-built "
                        "for the FP64 kernel), sorbate LJ 12.8 A + Ewald kappa 0.25/A kmax %d (nk %d "
                        "half-space), moves 1:1:1:1" % (kmax, nk))
    cfgs = None
    for nch in (1,) + tuple(chains):
        b = batch(nch, H.MH_PREC_FP32, cs)
        t = time.perf_counter()
        b.run(equil_steps)
        log("gcmc: %d chains equilibrated %d steps in %.2f s" % (b.nrep, equil_steps, time.perf_counter() - t))
        tune = b.autotune(200)
        rate, arate, s1, bps = timed(b, fp32_steps)
        log("gcmc FP32 batched, %d chains (threads %d, xy tables %d): %.3e steps/s, %.3e moves/s" %
            (b.nrep, tune[0], tune[1], rate, arate))
        key = "fp32_batched_%dchains" % b.nrep
        out[key + "_steps_per_s"] = rate
        out[key + "_attempted_moves_per_s"] = arate
        out[key + "_tuned"] = dict(threads=tune[0], xy_tables=tune[1])
        out[key + "_algorithmic_GBs"] = bps / 1e9
        if nch == 1:
            out["loadings"] = [int(x) for x in s1["N"]]
            cfgs = [b.config(r) for r in range(npt)]
            b32 = b
            out["roofline"] = dict(bound="hbm", kernel="k_gcmc", achieved=round(bps / 1e9, 2), peak=HBM_PEAK_GBS,
                                   unit="GB/s", frac=round(bps / 1e9 / HBM_PEAK_GBS, 5),
                                   algorithmic="per attempted move 288 B x 3 site-map lookups + 12 B x 3 sites "
                                               "x N other molecules + 16 B x %d k vectors (+16 B x nk per "
                                               "acceptance); N = each state point's mean loading" % nk)
        else:
            b.close()
    # the same state points, FP64 maps and kernel, one state point per launch
    picks = np.linspace(0, npt - 1, n_single).round().astype(int)
    tot_t, tot_steps, tot_att = 0.0, 0, 0
    tuned64 = []
    for r in picks:
        one = G.GcmcBatch(ctx, cs64["spec"], cs64["fugacities"][r:r + 1], cs64["seeds"][r:r + 1], maxN=1024,
                          maps=cs64["maps"], precision=H.MH_PREC_FP64, threads=256)
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
    for nch in (1,) + tuple(chains):
        k = "fp32_batched_%dchains_attempted_moves_per_s" % (npt * nch)
        out["speedup_%dchains_vs_fp64_single" % (npt * nch)] = out[k] / out[
            "fp64_single_state_point_attempted_moves_per_s"]
    out["trial_moves_per_s"] = out["fp32_batched_%dchains_attempted_moves_per_s" % npt]
    # end to end: FP32 maps + FP32 kernel against FP64 maps + FP64 kernel
    rng = np.random.default_rng(11)
    L = np.asarray(cs["spec"]["full_ell"])
    worst, nt = 0.0, 0
    b64 = G.GcmcBatch(ctx, cs64["spec"], cs64["fugacities"], cs64["seeds"], maxN=1024, maps=cs64["maps"],
                      precision=H.MH_PREC_FP64)
    for r in picks:
        b64.set_config(r, cfgs[r][0], cfgs[r][1])
        tc = rng.uniform(0, 1, (256, 3)) * L
        ta = np.stack([rng.uniform(-np.pi, np.pi, 256), np.arccos(rng.uniform(-1, 1, 256)),
                       rng.uniform(-np.pi, np.pi, 256)], 1)
        ok64, e64 = b64.trial_energy(r, tc, ta)
        ok32, e32 = b32.trial_energy(r, tc, ta)
        both = ok64 & ok32
        nt += int(both.sum())
        if both.any():
            worst = max(worst, float(np.max(np.abs(e32[both] - e64[both]) / np.maximum(np.abs(e64[both]), 1.0))))
    out["fp32_vs_fp64_trial_energy_max_rel_err"] = worst
    out["fp32_vs_fp64_trial_energy_note"] = ("%d insertion energies; FP32-built maps + FP32 kernel vs FP64-built "
                                             "maps + FP64 kernel; denominator floored at 1 kcal/mol" % nt)
    out["steps"] = dict(equil=equil_steps, fp32=fp32_steps, fp64_per_state_point=fp64_steps,
                        fp64_state_points=[int(x) for x in picks], chains_per_state_point=list(chains))
    # CPU baseline inputs: the FP64 maps and a few equilibrated configurations
    mlist = []
    pm_idx, seen = [], {}
    for a in range(ns):
        mp_ = cs64["maps"][a][0]
        if id(mp_) not in seen:
            ptr, vals = mp_.get()
            seen[id(mp_)] = len(mlist)
            mlist.append((mp_.npnts, mp_.step, ptr, vals))
        pm_idx.append(seen[id(mp_)])
    cpu = dict(spec=dict(cs64["spec"], maps=mlist, site_pmap=pm_idx, site_emap=[-1] * ns),
               fug=[float(cs["fugacities"][r]) for r in picks], seeds=[cs["seeds"][r] for r in picks],
               cfgs=[(cfgs[r][0], cfgs[r][1]) for r in picks])
    b64.close()
    b32.close()
    for m in cs["map_objs"] + cs64["map_objs"]:
        m.close()
    ctx.close()
    return out, cpu


def loading_leg(H, load_steps=100000, steps=400, per_load=16, targets=(0, 50, 100, 200)):
    """The batched CO2 kernel (FP32) at seeded loadings: a GCMC at 10^9 kPa
    packs the framework (maxN 640) and the densest configuration, cut to its
    first N molecules, starts `per_load` chains at each N of `targets` (and
    at the densest N reached); `steps` MC steps (moves 1:1:1:1 at 10^4 kPa,
    so N drifts little) time the kernel where the pair and LDS work grow
    with N."""
    from musichip import gcmc as G
    from musichip import workloads as W
    ctx = H.Context(0)
    cs = W.co2_isotherm_system(ctx, map_precision=H.MH_PREC_FP32, npoints=16)
    b = G.GcmcBatch(ctx, cs["spec"], [1.0e9] * 16, [7000 + i for i in range(16)], maxN=640, maps=cs["maps"],
                    precision=H.MH_PREC_FP32)
    t = time.perf_counter()
    b.run(load_steps)
    Ns = b.state()["N"]
    r = int(np.argmax(Ns))
    com, ang, _ = b.config(r)
    log("loading: packed %s molecules (max %d) in %.1f s" % (list(Ns), Ns.max(), time.perf_counter() - t))
    b.close()
    nmax = len(com)
    tg = sorted(set([n for n in targets if n < nmax] + [nmax]))
    out = dict(workload="CO2-like sorbate in the 30 A framework, FP32 maps and kernel, chains seeded with the "
                        "first N molecules of a configuration packed at 1e9 kPa (%d load steps), %d chains per N, "
                        "%d steps at 1e4 kPa" % (load_steps, per_load, steps), densest=int(nmax), points=[])
    for k, n in enumerate(tg):
        one = G.GcmcBatch(ctx, cs["spec"], [1.0e4] * per_load, [7100 + i for i in range(per_load)], maxN=640,
                          maps=cs["maps"], precision=H.MH_PREC_FP32)
        for j in range(per_load):
            one.set_config(j, com[:n], ang[:n])
        one.autotune(50)
        one.run(20)
        s0 = one.state()
        t = time.perf_counter()
        one.run(steps)
        dt = time.perf_counter() - t
        s1 = one.state()
        att = int((s1["att"] - s0["att"]).sum())
        out["points"].append(dict(N_start=int(n), N_mean=float(((s1["sumN"] - s0["sumN"]) / steps).mean()),
                                  steps_per_s=per_load * steps / dt, attempted_moves_per_s=att / dt,
                                  us_per_step_per_chain=1e6 * dt / steps))
        log("loading N=%d: %.3e moves/s over %d chains" % (n, att / dt, per_load))
        one.close()
    for m in cs["map_objs"]:
        m.close()
    ctx.close()
    return out


def serving_leg(n=512, steps=10000, util=0.7, ncal=64):
    """Screening service: seeded requests (map at 0.2 A + 10k-step methane
    GCMC at 1/10/50 bar; a quarter of them revisit 16 popular frameworks)
    with Poisson arrivals (seed 4999) at `util` of the capacity measured on
    a separate calibration set (all arriving at once)."""
    from musichip import serve
    cal = [serve.screening_request(100000 + r, steps) for r in range(ncal)]
    srv = serve.Server()
    _, cst = srv.run(cal, np.zeros(ncal))
    srv.close()
    rate = util * cst["requests_per_s"]
    log("serving: capacity %.1f requests/s (%d requests at once); arrival rate %.1f/s" %
        (cst["requests_per_s"], ncal, rate))
    fws = serve.request_frameworks(n)
    reqs = [serve.screening_request(r, steps, framework=f) for r, f in enumerate(fws)]
    arr = serve.poisson_arrivals(n, rate)
    srv = serve.Server()
    res, st = srv.run(reqs, arr)
    srv.close()
    st["arrival_rate_per_s"] = rate
    st["capacity_requests_per_s"] = cst["requests_per_s"]
    st["target_utilisation"] = util
    st["steps_per_request"] = steps * len(serve.FUGACITIES_KPA)
    st["failed_replicas"] = int(sum(int((r["err"] != 0).sum()) for r in res))
    st["distinct_frameworks"] = len(set(fws))
    log("serving: p50 %.1f ms, p99 %.1f ms, %.1f requests/s" % (st["p50_latency_ms"], st["p99_latency_ms"],
                                                             st["requests_per_s"]))
    sizes = [dict(nodes=serve.map_nodes(r), atoms=len(r["xyz"]), volume=float(np.prod(r["full"])),
                  steps=r["steps"] * len(serve.FUGACITIES_KPA), framework=r["framework"]) for r in reqs]
    return st, dict(sizes=sizes, arrivals=arr)


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
        if tag == "fp32":
            # algorithmic bytes: 11 force evaluations per move, each site
            # 288 B of map corners and 12 B per sorbate site within 12.8 A
            sites = 4.0 * nch
            nbr = (sites - 1.0) / 40.0 ** 3 * (4.0 / 3.0) * np.pi * 12.8 ** 3
            byt = moves * float(np.sum(11.0 * sites * (288.0 + 12.0 * nbr)))
            out["roofline"] = dict(bound="hbm", kernel="k_hmc", achieved=round(byt / dt / 1e9, 2),
                                   peak=HBM_PEAK_GBS, unit="GB/s", frac=round(byt / dt / 1e9 / HBM_PEAK_GBS, 5),
                                   algorithmic="per move 11 force evaluations x sites x (288 B map corners + "
                                               "12 B x sorbate sites within 12.8 A)")
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


def cpu_leg(pool, host, cfg, tabs, c1cpu, c3cpu, srv, seconds):
    """The oracle on the host: single-core rates (one single-threaded
    process per map / state point, as the reference runs) and the aggregate
    of P such processes side by side."""
    from oracle import cpu_baseline as CB
    P = pool.n
    npn = (cfg["fund_ell"] / cfg["gridspace"]).astype(int) + 1
    step = cfg["fund_ell"] / (npn - 1)
    probe, _ = pool.run(CB.map_nodes_work, map_cpu_items(cfg, tabs, npn, step, [64, 64, 64]))
    per_node = [t / n for n, t, _ in probe]
    nodes_each = [int(max(64, min(200000, seconds / 3.0 / max(pn, 1e-9)))) for pn in per_node]
    single, _ = pool.run(CB.map_nodes_work, map_cpu_items(cfg, tabs, npn, step, nodes_each))
    one = sum(n for n, _, _ in single) / sum(t for _, t, _ in single)
    agg_items = map_cpu_items(cfg, tabs, npn, step, nodes_each, parts=P)
    res, wall = pool.run(CB.map_nodes_work, agg_items)
    agg, nw = CB.aggregate_rate(res)
    log("CPU maps: one core %.3e pts/s, %d processes %.3e pts/s" % (one, P, agg))
    out = dict(value=agg, unit="map pts/s", cores=P, kind="port",
               sample="first %s nodes of the C / O pmaps and the emap of the charged-framework map build "
                      "(FP64 restatement, -O3 -march=native, one single-threaded process per work item, %d "
                      "processes; aggregate = sum over the %d busy processes of nodes / compute seconds)"
                      % (nodes_each, P, nw),
               single_core_map_pts_per_s=one, host=host)
    # C1 / C3 GCMC: one state point per process
    for tag, c, steps in (("c1", c1cpu, 200000), ("c3", c3cpu, 4000)):
        if c is None:
            continue
        items = [(c["spec"], f, sd, cfg_, steps) for f, sd, cfg_ in zip(c["fug"], c["seeds"], c["cfgs"])]
        res, wall = pool.run(CB.gcmc_work, items)
        att = sum(r[0] for r in res)
        st = sum(r[1] for r in res)
        tsum = sum(r[2] for r in res)
        agg, nw = CB.aggregate_rate(res, unit=0, secs=2)
        out[tag + "_gcmc"] = dict(single_core_attempted_moves_per_s=att / tsum, single_core_steps_per_s=st / tsum,
                                  aggregate_attempted_moves_per_s=agg, processes=nw,
                                  sample="%d state points x %d steps; aggregate = sum over the busy processes "
                                         "of attempted moves / compute seconds" % (len(items), steps),
                                  seconds_per_step=tsum / st)
        log("CPU %s GCMC: one core %.3e moves/s, aggregate %.3e over %d processes" % (tag, att / tsum, agg, nw))
    # serving stream on the CPU, modelled: per request the map cost (nodes x
    # LJ pairs within 12.8 A x the measured seconds per pair of the C pmap
    # sample) plus its GCMC steps x the measured C1 seconds per step, served
    # first in first out by P single-threaded processes on the same arrivals
    if srv is not None and "c1_gcmc" in out:
        rho2 = len(cfg["xyz"]) / float(np.prod(cfg["full_ell"]))
        vol_c = (4.0 / 3.0) * np.pi * cfg["hicut"] ** 3
        c_pair = per_node[0] / (rho2 * vol_c)
        c_step = out["c1_gcmc"]["seconds_per_step"]
        svc = np.array([z["nodes"] * min(z["atoms"], z["atoms"] / z["volume"] * vol_c) * c_pair + z["steps"] * c_step
                        for z in srv["sizes"]])
        lat = CB.fifo_latencies(svc, np.asarray(srv["arrivals"]), P)
        # the CPU's own operating point: Poisson arrivals at 70 % of the P
        # processes' capacity (the device run's target utilisation)
        cap = P / float(svc.mean())
        rng = np.random.default_rng(4)
        own = np.cumsum(rng.exponential(1.0 / (0.7 * cap), len(svc)))
        lat_own = CB.fifo_latencies(svc, own, P)
        out["c4_fifo_model"] = dict(p50_latency_ms=float(np.percentile(lat, 50) * 1e3),
                                    p99_latency_ms=float(np.percentile(lat, 99) * 1e3),
                                    capacity_requests_per_s=cap,
                                    at_70pct_capacity=dict(arrival_rate_per_s=0.7 * cap,
                                                           p50_latency_ms=float(np.percentile(lat_own, 50) * 1e3),
                                                           p99_latency_ms=float(np.percentile(lat_own, 99) * 1e3)),
                                    mean_service_s=float(svc.mean()), processes=P,
                                    note="modelled from the measured per-pair (map) and per-step (C1 GCMC) "
                                         "costs of this host; same request sizes and arrival times as the "
                                         "device run")
        log("CPU serving model: device arrivals p50 %.0f ms, p99 %.0f ms; capacity %.2f requests/s, at 70%%: "
            "p50 %.0f ms, p99 %.0f ms" % (out["c4_fifo_model"]["p50_latency_ms"],
                                         out["c4_fifo_model"]["p99_latency_ms"], cap,
                                         out["c4_fifo_model"]["at_70pct_capacity"]["p50_latency_ms"],
                                         out["c4_fifo_model"]["at_70pct_capacity"]["p99_latency_ms"]))
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
    ap.add_argument("--c1-steps", type=int, default=20000)
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-loading", action="store_true")
    ap.add_argument("--loading-steps", type=int, default=100000,
                    help="GCMC steps at 1e9 kPa that pack the framework for the loading sweep")
    ap.add_argument("--cpu-procs", type=int, default=12,
                    help="single-threaded oracle processes of the CPU aggregate (capped by the CPUs available)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every workload here runs on one device "
              "(multi-GPU sharding is described in DESIGN.md, not run)", file=sys.stderr)
        sys.exit(2)
    host, pool = None, None
    if not args.no_cpu and not args.copy_probe_mib:
        # the CPU worker processes are forked before anything touches the GPU
        from oracle import cpu_baseline as CB
        host = CB.host_info()
        CB.build_native()  # once, for this host, before the workers start
        pool = CB.Pool(max(1, min(args.cpu_procs, host["cpus_available"])))
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
    torch.cuda.synchronize()
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
    torch.cuda.synchronize()
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

    c1, c1cpu = None, None
    if not args.no_c1:
        c1, c1cpu = c1_leg(H, args.c1_steps)

    gcmc, c3cpu = None, None
    if not args.no_gcmc:
        gcmc, c3cpu = gcmc_leg(H, args.gcmc_steps, args.gcmc_fp64_steps, args.gcmc_equil,
                               tuple(int(c) for c in str(args.gcmc_chains).split(",") if c))

    loading = None
    if not args.no_loading:
        loading = loading_leg(H, args.loading_steps)

    hmc = None
    if not args.no_hmc:
        hmc = hmc_leg(args.hmc_moves)

    serving, srv = None, None
    if not args.no_serve:
        serving, srv = serving_leg(args.serve_requests)

    cpu = None
    if pool is not None:
        cpu = cpu_leg(pool, host, cfg, tabs, c1cpu, c3cpu, srv, args.cpu_seconds)
        pool.close()
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
               c1=c1, gcmc=gcmc, gcmc_trial_moves_per_s=gcmc and gcmc["trial_moves_per_s"], loading=loading,
               p50_latency_ms=serving and serving["p50_latency_ms"],
               p99_latency_ms=serving and serving["p99_latency_ms"], serving=serving, hmc=hmc,
               roofline=roofline, cpu_baseline=cpu)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
