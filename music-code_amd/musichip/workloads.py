"""GCMC workloads on the device: framework upload, probe maps and the GCMC
spec (as taken by musichip.gcmc.GcmcBatch) for

* the methane system: single-site methane (eps/k 148 K, sigma 3.73 A,
  LJ hicut 12.8 A) in the neutral 30 A framework of synth.config1, pmap at
  0.2 A (151^3 nodes), 300 K, insert:delete:translate 1:1:1, delta 0.4 A,
  fugacities 100-5000 kPa (8 log-spaced points), iseed 22111971;
* the batched isotherm: the 3-site CO2-like sorbate of synth.co2_like in the
  same framework, C and O pmaps at 0.2 A (no emap: the framework is neutral),
  sorbate-sorbate LJ (Lorentz-Berthelot, hicut 12.8 A) + Ewald (kappa
  0.25 / A, |m| <= 7 per axis, real-space cutoff 12.8 A), 300 K,
  insert:delete:translate:rotate 1:1:1:1, delta_trans 0.4 A, delta_rot 0.4,
  64 state points with fugacities log-spaced 1-10^4 kPa, seeds 3000 + i.

The map and move semantics follow the reference's GCMC driver
(gcmc.F90:224-258, mcmoveset.F90:280-335, rigidmoves.F90:691-1570).
"""
import numpy as np

from . import constants as C
from . import hip as H
from . import synth
from .inputs import lj_AB


def _self_table(types, lj, hicut):
    """Sorbate-sorbate LJ table over the site types (LB mixing)."""
    n = len(types)
    A, B = [], []
    for ti in types:
        for tj in types:
            e1, s1 = lj[ti]
            e2, s2 = lj[tj]
            a, b = lj_AB((s1 + s2) / 2.0, np.sqrt(e1 * e2))
            A.append(float(a))
            B.append(float(b))
    return (A, B, [hicut * hicut] * (n * n), [C.LOW_DIST_CUTOFF ** 2] * (n * n), [1] * (n * n))


def methane_system(ctx, map_precision=H.MH_PREC_FP64):
    cfg = synth.config1()
    L = cfg["full_ell"]
    ctx.upload_framework(cfg["xyz"], cfg["type"], cfg["q"], cfg["fund_ell"], L)
    p = cfg["probe"]
    ctx.set_lj(*synth.lb_table(cfg["fw_eps"], cfg["fw_sig"], p["eps"], p["sig"], cfg["hicut"]))
    pm = ctx.build_map(H.MH_MAP_LJ, map_precision, cfg["gridspace"], cfg["hipotcut"])
    spec = dict(full_ell=L, fund_ell=cfg["fund_ell"], ref=[[0.0, 0.0, 0.0]], stype=[0], sq=[0.0], nst=1,
                lj=_self_table(["CH4"], {"CH4": (p["eps"], p["sig"])}, cfg["hicut"]), T=300.0,
                moves=[("RINSERT", 1.0), ("RDELETE", 1.0), ("RTRANSLATE", 1.0)], dtrans=0.4)
    fug = np.logspace(2.0, np.log10(5000.0), 8)
    return dict(cfg=cfg, spec=spec, maps={0: (pm, None)}, map_objs=[pm], fugacities=fug,
                seeds=[22111971 + i for i in range(8)])


def co2_isotherm_system(ctx, map_precision=H.MH_PREC_FP64, npoints=64, kmax=7):
    cfg = synth.config1()
    L = cfg["full_ell"]
    ctx.upload_framework(cfg["xyz"], cfg["type"], cfg["q"], cfg["fund_ell"], L)
    sorb = synth.co2_like()
    pmaps = {}
    for name in ("O", "C"):
        e, s = sorb["lj"][name]
        ctx.set_lj(*synth.lb_table(cfg["fw_eps"], cfg["fw_sig"], e, s, cfg["hicut"]))
        pmaps[name] = ctx.build_map(H.MH_MAP_LJ, map_precision, cfg["gridspace"], cfg["hipotcut"])
    types = ["O", "C"]
    stype = [types.index(t) for t in sorb["types"]]
    spec = dict(full_ell=L, fund_ell=cfg["fund_ell"], ref=sorb["ref"], stype=stype, sq=sorb["q"],
                nst=len(types), lj=_self_table(types, sorb["lj"], cfg["hicut"]), T=300.0,
                ewald=dict(kappa=0.25, kmax=kmax, rc=cfg["hicut"]),
                moves=[("RINSERT", 1.0), ("RDELETE", 1.0), ("RTRANSLATE", 1.0), ("RROTATE", 1.0)],
                dtrans=0.4, drot=0.4)
    maps = {a: (pmaps[t], None) for a, t in enumerate(sorb["types"])}
    fug = np.logspace(0.0, 4.0, npoints)
    return dict(cfg=cfg, spec=spec, maps=maps, map_objs=list(pmaps.values()), fugacities=fug,
                seeds=[3000 + i for i in range(npoints)])


def butane_hmc_system(ctx, nrep=256, nmin=50, nmax=200, seed=5000, load_steps=1000000,
                      map_precision=H.MH_PREC_FP32, log=None):
    """HMC workload: flexible united-atom n-butane (synth.butane_ua) in the
    neutral 40 A framework of synth.config5 (768 atoms, seed 5000), maps for
    CH3 and CH2 at 0.2 A, nrep replicas with N ~ U[nmin, nmax] chains (seed
    + 1).  Start configurations are loaded on the device by a GCMC of rigid
    all-trans chains at 10^9 kPa (insert/delete/translate/rotate 2:1:1:1,
    seeds 5100 + i; 10^6 steps load 156-176 chains, the insertions of rigid
    chains into the dense fluid being rare) and truncated to each replica's N
    (fewer where the loading holds fewer).  HMC: 10 velocity-Verlet steps of
    1 fs, 300 K."""
    from . import gcmc as G
    from . import hmc as HM
    cfg = synth.config5(seed)
    L = cfg["full_ell"]
    ctx.upload_framework(cfg["xyz"], cfg["type"], cfg["q"], cfg["fund_ell"], L)
    model = synth.butane_ua()
    maps = {}
    for t, name in enumerate(model["types"]):
        e, s = model["lj"][name]
        ctx.set_lj(*synth.lb_table(cfg["fw_eps"], cfg["fw_sig"], e, s, cfg["hicut"]))
        maps[t] = ctx.build_map(H.MH_MAP_LJ, map_precision, cfg["gridspace"], cfg["hipotcut"])
    hs = HM.hmc_spec(model, L, cfg["fund_ell"], hicut=cfg["hicut"])
    # rigid-chain GCMC loading
    A, B, h2, l2, on = hs["lj"]
    gspec = dict(full_ell=L, fund_ell=cfg["fund_ell"], ref=model["chain"], stype=model["stype"],
                 sq=[0.0] * len(model["stype"]), nst=hs["nst"], lj=(A, B, h2, l2, on), T=300.0,
                 moves=[("RINSERT", 2.0), ("RDELETE", 1.0), ("RTRANSLATE", 1.0), ("RROTATE", 1.0)],
                 dtrans=0.4, drot=0.4)
    gmaps = {a: (maps[t], None) for a, t in enumerate(model["stype"])}
    b = G.GcmcBatch(ctx, gspec, [1.0e9] * nrep, [5100 + i for i in range(nrep)], maxN=512, maps=gmaps,
                    precision=H.MH_PREC_FP32)
    b.run(load_steps)
    if log:
        log("hmc: loaded start configurations (GCMC %d steps), N %s..%s" %
            (load_steps, b.state()["N"].min(), b.state()["N"].max()))
    rng = np.random.default_rng(seed + 1)
    target = rng.integers(nmin, nmax + 1, nrep)
    configs = []
    for r in range(nrep):
        _, _, sites = b.config(r)
        n = min(int(target[r]), len(sites))
        sites = sites[:n]
        d = sites - sites[:, :1, :]
        rp = sites[:, :1, :] + (d - np.round(d / L) * L)
        configs.append(rp)
    b.close()
    return dict(cfg=cfg, model=model, spec=hs, maps=maps, configs=configs, target=target)
