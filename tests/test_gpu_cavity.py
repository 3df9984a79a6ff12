"""Cavity-biased insertion / deletion on the device (mh_gcmc_set_cavity,
CAVITY_BIAS / MAP_CAVITY_BIAS in mh_gcmc): the tabulation from a pmap
(mh_cavity_tab) against a restatement of cavitylist_initArrPmap and its
unit-cell periodicity, and FP64 trajectories move for move against oracle
chains, with the cav_prob tables and the cube counts of every replica equal
to the oracle's."""
import numpy as np
import pytest

from oracle import oracle as O
from musichip import hip as H
from musichip import gcmc as G
from musichip import synth
import hmc_systems as HS
from test_gpu_bias import _framework_maps, _oracle_map, _oracle_spec, _co2_spec

pytestmark = pytest.mark.gpu


def _tab_restated(m, ucutoff, n, ncell, full):
    return tab_from_arrays(*_oracle_map(m), ucutoff, n, ncell, full)


def tab_from_arrays(npn, step, ptr, vals, ucutoff, n, ncell, full):
    """cavitylist_initArrPmap: cube of each cubelet centre whose lower-corner
    node is tabulated below ucutoff, in every unit cell."""
    ptr = np.asarray(ptr).reshape(npn)
    n = np.asarray(n)
    per = n // np.asarray(ncell)
    d = np.asarray(full) / n
    tab = np.zeros(n, dtype=np.uint8)
    for i in range(npn[0] - 1):
        for j in range(npn[1] - 1):
            for k in range(npn[2] - 1):
                q = ptr[i, j, k]
                if q == 0 or not float(vals[0][q - 1]) < ucutoff:
                    continue
                x = np.array([i, j, k]) * np.asarray(step) + np.asarray(step) / 2
                ic = (x / d).astype(int)
                for u in range(ncell[0]):
                    for v in range(ncell[1]):
                        for w in range(ncell[2]):
                            tab[ic[0] + per[0] * u, ic[1] + per[1] * v, ic[2] + per[2] * w] = 1
    return tab


def test_cavity_tab_from_pmap(hip):
    print("case: cavity tabulation from a pmap", flush=True)
    m = _framework_maps(hip, {"P": (148.0, 3.73)})["P"]
    L = np.asarray(HS.L16)
    for ucut, n, ncell in ((0.0, (10, 10, 10), (1, 1, 1)), (-1.0, (16, 8, 12), (2, 1, 3))):
        full = L * np.asarray(ncell)
        tab = G.cavity_tab(m, ucut, n, ncell, full)
        ref = _tab_restated(m, ucut, n, ncell, full)
        assert np.array_equal(tab, ref)
        assert 0 < tab.sum() < tab.size
        per = np.asarray(n) // np.asarray(ncell)
        if ncell[0] == 2:
            assert np.array_equal(tab[: per[0]], tab[per[0]:])
        if ncell[2] == 3:
            assert np.array_equal(tab[:, :, : per[2]], tab[:, :, per[2]: 2 * per[2]])
    m.close()


def _oracle_cav(n, tab, mask=1):
    return dict(n=n, tab=None if tab is None else tab.astype(np.int8), mask=mask)


@pytest.mark.parametrize("system", ["probe", "co2"])
def test_cavity_gcmc_fp64_matches_oracle(hip, system):
    """CAVITY_BIAS (1-site probe, tabulation from the pmap, 20 trials) and
    MAP_CAVITY_BIAS (3-site, Ewald, translate + rotate, uniform tabulation,
    100 trials over two waves) replica by replica against oracle chains."""
    print("case: cavity-biased GCMC FP64 vs oracle, %s" % system, flush=True)
    n = (10, 10, 10)
    if system == "probe":
        maps = _framework_maps(hip, {"P": (148.0, 3.73)})
        from musichip.inputs import lj_AB
        A, B = lj_AB(3.73, 148.0)
        spec = dict(full_ell=list(HS.L16), fund_ell=list(HS.L16), ref=[[0.0, 0.0, 0.0]], stype=[0], sq=[0.0], nst=1,
                    lj=([A], [B], [HS.HICUT ** 2], [0.25], [1]), T=300.0,
                    moves=[("BINSERT", 1.0), ("BDELETE", 1.0), ("RTRANSLATE", 1.0)], dtrans=0.25)
        site_maps = [maps["P"]]
        fug = [200.0, 2000.0, 20000.0, 2000.0]
        tab = G.cavity_tab(maps["P"], 0.0, n, (1, 1, 1), HS.L16)
        T, cmap, dbias, obias = 20, 0, None, None
    else:
        sorb = synth.co2_like()
        maps = _framework_maps(hip, {t: sorb["lj"][t] for t in ("O", "C")})
        spec = _co2_spec(None)
        site_maps = [maps[t] for t in sorb["types"]]
        fug = [100.0, 1000.0, 5000.0, 1000.0]
        tab = None
        T, cmap = 100, 1
        bm = maps["C"]
        dbias = H.BiasMap(bm, 300.0)
        npn, step, ptr, vals = _oracle_map(bm)
        obias = O.Bmap(npn, step, HS.L16, ptr, vals, 300.0)
        spec["bias_map"] = dbias
    seeds = [81, 82, 83, 84]
    b = G.GcmcBatch(hip, spec, fug, seeds, maxN=256, maps={a: (m, None) for a, m in enumerate(site_maps)})
    b.set_cavity(n, [T], [cmap], tab=tab, species_mask=1)
    ospec = dict(_oracle_spec(spec, site_maps, obias), cavity=_oracle_cav(n, tab), cav_ntrials=T, cav_map=cmap)
    chains = [O.Gcmc(ospec, f, s_, maxN=256) for f, s_ in zip(fug, seeds)]
    for step_to in (1500, 4000, 8000):
        b.run(step_to - int(b.state()["nsteps"][0]))
        st = b.state()
        for r, g in enumerate(chains):
            g.run(step_to - g.st.nsteps)
            assert list(st["att"][r]) == [g.st.att[i] for i in range(4)], (r, step_to)
            assert list(st["acc"][r]) == [g.st.acc[i] for i in range(4)], (r, step_to)
            assert st["N"][r] == g.st.N
            assert st["energy"][r] == pytest.approx(g.st.energy, rel=1e-10, abs=1e-10)
            # the device table holds loadings up to nspc maxN + 1
            cap = 258
            avg, cnt, cubes = b.cavity(r, cap)
            assert np.array_equal(cnt, g.cav_n[:cap]), (r, step_to)
            np.testing.assert_allclose(avg, g.cav_avg[:cap], rtol=1e-15, atol=0)
            assert np.array_equal(cubes, g.cube_cnt), (r, step_to)
    assert st["acc"][:, 0].sum() > 50 and st["acc"][:, 1].sum() > 20
    # a configuration set from the host (half of replica 2's molecules)
    # rebuilds the cube counts
    c = b.config(2)
    k = len(c[0]) // 2
    b.set_config(0, c[0][:k], c[1][:k])
    chains[0].set_config(c[0][:k], c[1][:k])
    assert np.array_equal(b.cavity(0, 4)[2], chains[0].cube_cnt)
    b.close()
    if dbias is not None:
        dbias.close()


def test_cavity_mixture_fp64_matches_oracle(hip):
    """A CO2-like sorbate with CAVITY_BIAS (40 trials, no bias map) beside
    methane with random moves, both in the cavity list: FP64 move sequences,
    energies, cav_prob tables and cube counts against the oracle's mixture."""
    print("case: cavity-biased CO2 + CH4 mixture, FP64 trajectory parity", flush=True)
    from test_gpu_mix import _mixture, _oracle
    specs, maps, keep = _mixture(hip, ewald=True)
    specs[0] = dict(specs[0], moves=[("BINSERT", 1.0), ("BDELETE", 1.0), ("RTRANSLATE", 1.0), ("RROTATE", 1.0)])
    n = (10, 10, 10)
    ospecs = [dict(specs[0], cavity=_oracle_cav(n, None, mask=3), cav_ntrials=40, cav_map=0),
              dict(specs[1], cavity=_oracle_cav(n, None, mask=3))]
    fugs = [[3000.0, 3.0e4], [1.0e4, 1.0e5]]
    seeds = [4301, 4302]
    nsteps = 3000
    b = G.GcmcMix(hip, specs, fugs, seeds, maxN=256, maps=maps)
    b.set_cavity(n, [40, 0], [0, 0], tab=None, species_mask=3)
    b.set_trace(nsteps)
    b.run(nsteps)
    cap = 2 * 256 + 2
    for r in range(2):
        g = _oracle(ospecs, maps, fugs[r], seeds[r], nsteps)
        to = g.trace_array(nsteps)
        tg = b.trace(r)
        assert np.array_equal(tg[:, :3].astype(int), to[:, :3].astype(int)), r
        scale = np.maximum(np.abs(to[:, 3]), 1.0)
        assert np.max(np.abs(tg[:, 3] - to[:, 3]) / scale) < 1e-10
        for s in range(2):
            b.select(s)
            st = b.state()
            assert st["N"][r] == g.N()[s]
            assert list(st["att"][r]) == list(g.ch[s].st.att) and list(st["acc"][r]) == list(g.ch[s].st.acc)
        b.select(0)
        avg, cnt, cubes = b.cavity(r, cap)
        assert np.array_equal(cnt, g.ch[0].cav_n[:cap])
        np.testing.assert_allclose(avg, g.ch[0].cav_avg[:cap], rtol=1e-15, atol=0)
        assert np.array_equal(cubes, g.ch[0].cube_cnt)
        assert cubes.sum() == 3 * g.N()[0] + g.N()[1]
    assert b.state()["acc"][:, 0].sum() > 20
    b.close()


def test_cavity_graph_replay_and_fp32(hip):
    """Cycles replayed from one captured hipGraph equal eager launches for a
    cavity-biased batch (cube counts, cav_prob tables, counters); the FP32
    kernel's mean loading agrees with FP64's."""
    print("case: cavity-biased GCMC, graph replay and FP32", flush=True)
    maps = _framework_maps(hip, {"P": (148.0, 3.73)})
    from musichip.inputs import lj_AB
    A, B = lj_AB(3.73, 148.0)
    spec = dict(full_ell=list(HS.L16), fund_ell=list(HS.L16), ref=[[0.0, 0.0, 0.0]], stype=[0], sq=[0.0], nst=1,
                lj=([A], [B], [HS.HICUT ** 2], [0.25], [1]), T=300.0,
                moves=[("BINSERT", 1.0), ("BDELETE", 1.0), ("RTRANSLATE", 1.0)], dtrans=0.25)
    n = (10, 10, 10)
    fug = [3000.0] * 32
    seeds = list(range(900, 932))
    mp = {0: (maps["P"], None)}
    eager = G.GcmcBatch(hip, spec, fug, seeds, maxN=256, maps=mp)
    graph = G.GcmcBatch(hip, spec, fug, seeds, maxN=256, maps=mp)
    for b in (eager, graph):
        b.set_cavity(n, [30], [0], tab=None, species_mask=1)
    eager.run(2000)
    graph.run_cycles(500, 4)
    se, sg = eager.state(), graph.state()
    for k in ("N", "att", "acc", "energy", "nsteps"):
        assert np.array_equal(se[k], sg[k]), k
    for r in range(0, 32, 7):
        for a, b in zip(eager.cavity(r, 258), graph.cavity(r, 258)):
            assert np.array_equal(a, b)
    means = {}
    for prec in (H.MH_PREC_FP64, H.MH_PREC_FP32):
        b = G.GcmcBatch(hip, spec, fug, seeds, maxN=256, maps=mp, precision=prec)
        b.set_cavity(n, [30], [0], tab=None, species_mask=1)
        b.run(5000)
        s0 = b.state()
        b.run(40000)
        s1 = b.state()
        m = (s1["sumN"] - s0["sumN"]) / (s1["nsteps"] - s0["nsteps"])
        means[prec] = (m.mean(), m.std(ddof=1) / np.sqrt(len(m)))
        b.close()
    (m64, e64), (m32, e32) = means[H.MH_PREC_FP64], means[H.MH_PREC_FP32]
    assert m64 > 1.0
    assert abs(m64 - m32) < 4 * np.hypot(e64, e32) + 0.01 * m64, means
    eager.close()
    graph.close()
