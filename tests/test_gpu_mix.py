"""GCMC of mixtures on the GPU against the oracle chain (mo_gcmc_run_mix):
a CO2-like 3-site sorbate (maps, Ewald) and united-atom methane (map) in a
charged 20 A framework; FP64 move sequences identical, energies within
1e-10, per-species statistics; the FP32 kernel's trial energies per
species within 1e-5 of FP64."""
import numpy as np
import pytest

from oracle import oracle as O
from musichip import hip as H
from musichip import gcmc as G
from musichip import inputs, synth

pytestmark = pytest.mark.gpu


def _mixture(hip, ewald=True, grid=0.3):
    rng = np.random.default_rng(33)
    xyz = synth.lattice_faces(2, 10.0, rng)
    eps, sig, typ = synth.fw_types(rng, len(xyz), 4)
    q = rng.uniform(-0.6, 0.6, len(xyz))
    q -= q.mean()
    L = np.array([20.0, 20.0, 20.0])
    hip.upload_framework(xyz, typ, q, L, L)
    sorb = synth.co2_like()
    lj = dict(sorb["lj"], M=(148.0, 3.73))
    pm = {}
    for name in ("O", "C", "M"):
        e, s_ = lj[name]
        hip.set_lj(*synth.lb_table(eps, sig, e, s_, 9.5))
        pm[name] = hip.build_map(H.MH_MAP_LJ, H.MH_PREC_FP64, grid, 100.0)
    em = None
    if ewald:
        hip.set_ewald(5.0, 7, 1e-10)
        em = hip.build_map(H.MH_MAP_EWALD, H.MH_PREC_FP64, grid, 100.0, probe_q=1.0)
    names = ("O", "C", "M")
    A, B = [], []
    for ti in names:
        for tj in names:
            e1, s1 = lj[ti]
            e2, s2 = lj[tj]
            a, b = inputs.lj_AB((s1 + s2) / 2, np.sqrt(e1 * e2))
            A.append(a)
            B.append(b)
    shared = dict(full_ell=L, fund_ell=L, nst=3, lj=(A, B, [9.5 ** 2] * 9, [0.5 ** 2] * 9, [1] * 9), T=300.0)
    if ewald:
        shared["ewald"] = dict(kappa=0.30, kmax=6, rc=9.5)
    co2 = dict(shared, ref=sorb["ref"], stype=[0, 1, 0], sq=sorb["q"],
               moves=[("RINSERT", 1.0), ("RDELETE", 1.0), ("RTRANSLATE", 1.0), ("RROTATE", 1.0)],
               dtrans=0.4, drot=0.4)
    ch4 = dict(shared, ref=[[0.0, 0.0, 0.0]], stype=[2], sq=[0.0],
               moves=[("RINSERT", 2.0), ("RDELETE", 2.0), ("RTRANSLATE", 1.0)], dtrans=0.5)
    maps = [{a: (pm[t], em) for a, t in enumerate(sorb["types"])}, {0: (pm["M"], None)}]
    return [co2, ch4], maps, list(pm.values()) + ([em] if em else [])


def _oracle(specs, maps, fugs, seed, nsteps):
    mspecs = []
    for sp, mp in zip(specs, maps):
        mlist, idx, pmi, emi = [], {}, [], []
        for a in range(len(sp["ref"])):
            p_, e_ = mp[a]
            for m, dst in ((p_, pmi), (e_, emi)):
                if m is None:
                    dst.append(-1)
                    continue
                if id(m) not in idx:
                    ptr, vals = m.get()
                    idx[id(m)] = len(mlist)
                    mlist.append((m.npnts, m.step, ptr, vals))
                dst.append(idx[id(m)])
        mspecs.append(dict(sp, maps=mlist, site_pmap=pmi, site_emap=emi))
    g = O.GcmcMix(mspecs, fugs, seed, trace=nsteps)
    g.run(nsteps)
    return g


@pytest.mark.parametrize("ewald", [True, False])
def test_mixture_fp64_trajectory_matches_oracle(hip, ewald):
    print("case: CO2 + CH4 mixture, ewald %s, FP64 trajectory parity" % ewald, flush=True)
    specs, maps, keep = _mixture(hip, ewald)
    fugs = [[3000.0, 3.0e4], [1.0e4, 1.0e5]]
    seeds = [4101, 4102]
    nsteps = 3000
    b = G.GcmcMix(hip, specs, fugs, seeds, maxN=256, maps=maps)
    b.set_trace(nsteps)
    b.run(nsteps)
    for r in range(2):
        g = _oracle(specs, maps, fugs[r], seeds[r], nsteps)
        to = g.trace_array(nsteps)
        tg = b.trace(r)
        assert np.array_equal(tg[:, :3].astype(int), to[:, :3].astype(int)), r
        scale = np.maximum(np.abs(to[:, 3]), 1.0)
        assert np.max(np.abs(tg[:, 3] - to[:, 3]) / scale) < 1e-10
        for s in range(2):
            b.select(s)
            st = b.state()
            assert st["N"][r] == g.N()[s]
            assert st["sumN"][r] == g.ch[s].st.sumN
            assert list(st["att"][r]) == list(g.ch[s].st.att) and list(st["acc"][r]) == list(g.ch[s].st.acc)
        assert abs(b.state()["energy"][r] - g.energy) < 1e-9 * max(1.0, abs(g.energy))
        spc = tg[:, 0].astype(int) // 16
        assert (spc == 0).sum() > 1000 and (spc == 1).sum() > 1000
    b.select(0)
    n0 = b.state()["N"]
    b.select(1)
    n1 = b.state()["N"]
    # both species present, so the cross-species pairs take part
    assert min(n0.max(), n1.max()) >= 2 and n0.max() + n1.max() >= 10, (n0, n1)
    tot = b.total_energy()
    assert np.allclose(tot, b.state()["energy"], rtol=1e-10, atol=1e-9)
    # FP32 kernel on the same configurations: trial energies per species
    b32 = G.GcmcMix(hip, specs, fugs, seeds, maxN=256, maps=maps, precision=H.MH_PREC_FP32)
    rng = np.random.default_rng(8)
    for s in range(2):
        b.select(s)
        b32.select(s)
        for r in range(2):
            com, ang, _ = b.config(r)
            b32.set_config(r, com, ang)
    for s in range(2):
        b.select(s)
        b32.select(s)
        tc = rng.uniform(0, 20.0, (1500, 3))
        ta = np.stack([rng.uniform(-np.pi, np.pi, 1500), np.arccos(rng.uniform(-1, 1, 1500)),
                       rng.uniform(-np.pi, np.pi, 1500)], 1)
        ok64, e64 = b.trial_energy(0, tc, ta)
        ok32, e32 = b32.trial_energy(0, tc, ta)
        assert np.count_nonzero(ok64 != ok32) <= 4
        both = ok64 & ok32
        assert both.sum() > 100
        assert np.max(np.abs(e32[both] - e64[both]) / np.maximum(np.abs(e64[both]), 1.0)) <= 1e-5
    b32.close()
    b.close()
    for m in keep:
        m.close()
