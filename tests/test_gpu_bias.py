"""Biased insertion / deletion on the device (mh_bmap_*, BINSERT / BDELETE
in mh_gcmc): the bias map against the oracle's (bmap_cubeinit), the
device search against the reference's bisection (bmap_getbiasindex), FP64
trajectories move for move against oracle chains, and the grand canonical
mean of non-interacting molecules from many device replicas."""
import numpy as np
import pytest

from oracle import oracle as O
from musichip import hip as H
from musichip import gcmc as G
from musichip import synth
import hmc_systems as HS

pytestmark = pytest.mark.gpu
KB = 8.314e-3 / 4.184


def _framework_maps(hip, types):
    """Maps of LJ probe types {name: (eps, sig)} in the 16 A test framework."""
    fw = HS.framework()
    L = HS.L16
    hip.upload_framework(fw["xyz"], fw["type"], None, L, L)
    maps = {}
    for name, (e, s) in types.items():
        hip.set_lj(*synth.lb_table(fw["eps"], fw["sig"], e, s, HS.HICUT))
        maps[name] = hip.build_map(H.MH_MAP_LJ, H.MH_PREC_FP64, HS.GRID, 100.0)
    return maps


def _oracle_map(m):
    ptr, vals = m.get()
    return (m.npnts, m.step, ptr, vals)


@pytest.mark.parametrize("T", [300.0, 20000.0])
def test_device_bias_map_matches_oracle(hip, T):
    print("case: bias map on the device vs the oracle, T_bias %g" % T, flush=True)
    m = _framework_maps(hip, {"P": (148.0, 3.73)})["P"]
    npn, step, ptr, vals = _oracle_map(m)
    ob = O.Bmap(npn, step, HS.L16, ptr, vals, T)
    db = H.BiasMap(m, T)
    cube, wt, cum = db.get()
    assert db.numtab == ob.numtab and db.ncubelets == ob.ncubelets
    assert np.array_equal(cube, ob.cube)
    np.testing.assert_allclose(wt, ob.wt, rtol=4e-16, atol=0)
    np.testing.assert_allclose(cum, ob.cum, rtol=2e-15, atol=0)
    assert db.cum_wt == pytest.approx(ob.cum_wt, rel=1e-14)
    assert (np.isinf(db.clamp_energy) and np.isinf(ob.clamp_energy)) or \
        db.clamp_energy == pytest.approx(ob.clamp_energy, rel=1e-13)
    # the kernel's guide-table search = the reference's bisection
    rng = np.random.default_rng(3)
    keys = np.concatenate([rng.random(50000), cum[rng.integers(0, db.numtab, 5000)], [0.0]])
    keys = keys[keys < 1.0]
    got = db.search(keys)
    assert np.array_equal(got, np.searchsorted(cum, keys, side="left"))
    ref = np.array([ob.index(k) for k in keys[:20000]])
    assert np.array_equal(got[:20000], ref)
    db.close()
    m.close()


def _co2_spec(bias):
    sorb = synth.co2_like()
    types = ["O", "C"]
    from musichip.inputs import lj_AB
    A, B = [], []
    for ti in types:
        for tj in types:
            e1, s1 = sorb["lj"][ti]
            e2, s2 = sorb["lj"][tj]
            a, b = lj_AB((s1 + s2) / 2, np.sqrt(e1 * e2))
            A.append(a)
            B.append(b)
    return dict(full_ell=list(HS.L16), fund_ell=list(HS.L16), ref=sorb["ref"], stype=[0, 1, 0], sq=sorb["q"], nst=2,
                lj=(A, B, [HS.HICUT ** 2] * 4, [0.25] * 4, [1] * 4), T=300.0,
                ewald=dict(kappa=0.32, kmax=5, rc=7.9),
                moves=[("BINSERT", 1.0), ("BDELETE", 1.0), ("RTRANSLATE", 1.0), ("RROTATE", 1.0)],
                dtrans=0.25, drot=0.3, bias_map=bias)


def _oracle_spec(spec, site_maps, bias):
    """The device spec with downloaded maps and the oracle's bias map."""
    mlist, idx, pm = [], {}, []
    for m in site_maps:
        if id(m) not in idx:
            idx[id(m)] = len(mlist)
            mlist.append(_oracle_map(m))
        pm.append(idx[id(m)])
    sp = dict(spec, maps=mlist, site_pmap=pm, site_emap=[-1] * len(site_maps), bias=bias)
    sp.pop("bias_map", None)
    return sp


@pytest.mark.parametrize("system", ["probe", "co2"])
def test_biased_gcmc_fp64_matches_oracle(hip, system):
    """BINSERT / BDELETE (+ translate, + rotate and Ewald for the 3-site
    sorbate) on the device, replica by replica against oracle chains:
    identical counts and N at every checkpoint, energies to 1e-10."""
    print("case: biased GCMC FP64 vs oracle, %s" % system, flush=True)
    if system == "probe":
        maps = _framework_maps(hip, {"P": (148.0, 3.73)})
        bm = maps["P"]
        from musichip.inputs import lj_AB
        A, B = lj_AB(3.73, 148.0)
        spec = dict(full_ell=list(HS.L16), fund_ell=list(HS.L16), ref=[[0.0, 0.0, 0.0]], stype=[0], sq=[0.0], nst=1,
                    lj=([A], [B], [HS.HICUT ** 2], [0.25], [1]), T=300.0,
                    moves=[("BINSERT", 1.0), ("BDELETE", 1.0), ("RTRANSLATE", 1.0)], dtrans=0.25)
        site_maps = [maps["P"]]
        fug = [200.0, 2000.0, 20000.0, 2000.0]
    else:
        sorb = synth.co2_like()
        maps = _framework_maps(hip, {t: sorb["lj"][t] for t in ("O", "C")})
        bm = maps["C"]
        spec = _co2_spec(None)
        site_maps = [maps[t] for t in sorb["types"]]
        fug = [100.0, 1000.0, 5000.0, 1000.0]
    dbias = H.BiasMap(bm, 300.0)
    npn, step, ptr, vals = _oracle_map(bm)
    obias = O.Bmap(npn, step, HS.L16, ptr, vals, 300.0)
    spec["bias_map"] = dbias
    seeds = [71, 72, 73, 74]
    b = G.GcmcBatch(hip, spec, fug, seeds, maxN=256, maps={a: (m, None) for a, m in enumerate(site_maps)})
    chains = [O.Gcmc(_oracle_spec(spec, site_maps, obias), f, s) for f, s in zip(fug, seeds)]
    for step_to in (2000, 6000, 12000):
        b.run(step_to - int(b.state()["nsteps"][0]))
        st = b.state()
        for r, g in enumerate(chains):
            g.run(step_to - g.st.nsteps)
            assert list(st["att"][r]) == [g.st.att[i] for i in range(4)], (r, step_to)
            assert list(st["acc"][r]) == [g.st.acc[i] for i in range(4)], (r, step_to)
            assert st["N"][r] == g.st.N
            assert st["energy"][r] == pytest.approx(g.st.energy, rel=1e-10, abs=1e-10)
    assert st["acc"][:, 0].sum() > 50 and st["acc"][:, 1].sum() > 20
    b.close()
    dbias.close()


def test_device_biased_ensemble_mean(hip):
    """Non-interacting molecules: the mean N of 32 replicas with random and
    with biased insertion / deletion against exp(B) <exp(-U/kT)>."""
    print("case: biased vs random GCMC ensemble mean on the device", flush=True)
    maps = _framework_maps(hip, {"P": (148.0, 3.73)})
    m = maps["P"]
    npn, step, ptr, vals = _oracle_map(m)
    rng = np.random.default_rng(11)
    pts = rng.random((100000, 3)) * HS.L16
    ok, e = m.interp(pts)
    boltz = np.where(ok, np.exp(-np.where(ok, e, 0.0) / (KB * 300.0)), 0.0)
    fug = 10000.0
    eB = np.exp(O.adams_B(fug, float(np.prod(HS.L16)), 300.0))
    lam, lam_err = eB * boltz.mean(), eB * boltz.std() / np.sqrt(len(boltz))
    bias = H.BiasMap(m, 300.0)
    res = {}
    for name, moves in (("random", [("RINSERT", 1.0), ("RDELETE", 1.0)]),
                        ("biased", [("BINSERT", 1.0), ("BDELETE", 1.0)])):
        spec = dict(full_ell=list(HS.L16), fund_ell=list(HS.L16), ref=[[0.0, 0.0, 0.0]], stype=[0], sq=[0.0], nst=1,
                    lj=([0.0], [0.0], [0.0], [0.0], [0]), T=300.0, moves=moves, bias_map=bias)
        b = G.GcmcBatch(hip, spec, [fug] * 32, list(range(500, 532)), maxN=128, maps={0: (m, None)})
        b.run(20000)
        s0 = b.state()
        b.run(100000)
        s1 = b.state()
        means = (s1["sumN"] - s0["sumN"]) / (s1["nsteps"] - s0["nsteps"])
        res[name] = (means.mean(), means.std(ddof=1) / np.sqrt(len(means)))
        b.close()
    for k, (mean, se) in res.items():
        assert abs(mean - lam) < 4.0 * np.hypot(se, lam_err) + 0.01 * lam, (k, mean, se, lam, lam_err)
    bias.close()
