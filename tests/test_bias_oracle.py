"""The oracle's biased insertion / deletion (BINSERT / BDELETE with a bias
map, bmap.F90:181-636, rigidmoves.F90:1386-1442, 1578-1654, 1938-1969),
pinned without the reference (its GCMC fixtures with bias maps need map
files the reference does not ship):

* the bias map against a numpy restatement of bmap_cubeinit (cubelet list,
  weights, running cumulative probabilities, the clamp branch);
* bmap_getbiasindex's bisection returns the first cumulative probability
  >= key (numpy searchsorted);
* ensemble: for molecules that do not interact with each other, N in the
  grand canonical ensemble is Poisson with mean exp(B) <exp(-U/kT)>_cell;
  chains with random and with biased insertion / deletion (two bias
  temperatures) reproduce that mean, computed independently by integrating
  the map over the cell.  A missing or inverted bias factor moves the biased
  chains' mean by a factor of several.
"""
import numpy as np
import pytest

from oracle import oracle as O
from musichip import synth
import hmc_systems as HS

KB = 8.314e-3 / 4.184


@pytest.fixture(scope="module")
def probe_map():
    """Methane-like probe (148 K, 3.73 A) in the 16 A test framework, 0.4 A grid."""
    fw = HS.framework()
    model = dict(types=["P"], lj={"P": (148.0, 3.73)})
    (npn, step, ptr, vals), = HS.oracle_maps(fw, model)
    return dict(npn=npn, step=step, ptr=ptr, vals=vals, eff=HS.L16)


def _numpy_bmap(m, T):
    """bmap_cubeinit in numpy: cubelets (x outer, z inner) whose lower-corner
    node is tabulated with V <= 30; clamp when min/sum <= 1e-6."""
    npn = m["npn"]
    nx1, ny1, nz1 = npn - 1
    ptr3 = m["ptr"].reshape(npn[0], npn[1], npn[2])[:nx1, :ny1, :nz1].reshape(-1)
    V = m["vals"][0]
    ok = ptr3 != 0
    u = np.full(len(ptr3), np.nan)
    u[ok] = V[ptr3[ok] - 1].astype(np.float64)
    ok &= u <= 30.0
    rti = 1.0 / (T * KB)
    cube = np.nonzero(ok)[0]
    uu = u[cube]
    wt = np.exp(np.clip(-rti * uu, -30, 30))
    tot = np.add.accumulate(wt)[-1]
    clamp = np.inf
    if not (wt.min() / tot > 1e-6):
        clamp = -np.log(tot * 1e-6) / rti
        wt = np.exp(np.clip(-rti * np.minimum(uu, clamp), -30, 30))
    run = np.add.accumulate(wt)
    return cube, wt, run / run[-1], run[-1], clamp


@pytest.mark.parametrize("T", [300.0, 20000.0])
def test_bias_map_matches_numpy_restatement(probe_map, T):
    m = probe_map
    b = O.Bmap(m["npn"], m["step"], m["eff"], m["ptr"], m["vals"], T)
    cube, wt, cum, tot, clamp = _numpy_bmap(m, T)
    assert b.numtab == len(cube) and b.ncubelets == int(np.prod(m["npn"] - 1))
    assert np.array_equal(b.cube, cube)
    np.testing.assert_allclose(b.wt, wt, rtol=4e-16, atol=0)
    np.testing.assert_allclose(b.cum, cum, rtol=2e-15, atol=0)
    assert b.cum[-1] == 1.0
    assert b.cum_wt == pytest.approx(tot, rel=1e-14)
    if np.isinf(clamp):
        assert np.isinf(b.clamp_energy)
    else:
        assert b.clamp_energy == pytest.approx(clamp, rel=1e-13)
    # 300 K takes the clamp branch (min / sum <= 1e-6), 20000 K does not
    assert np.isinf(clamp) == (T == 20000.0)


def test_bias_index_is_first_cumulative_ge_key(probe_map):
    m = probe_map
    b = O.Bmap(m["npn"], m["step"], m["eff"], m["ptr"], m["vals"], 300.0)
    rng = np.random.default_rng(7)
    keys = np.concatenate([rng.random(20000), b.cum[rng.integers(0, b.numtab, 2000)],
                           np.nextafter(b.cum[rng.integers(0, b.numtab - 1, 2000)], 0.0), [0.0]])
    keys = keys[keys < 1.0]
    want = np.searchsorted(b.cum, keys, side="left")
    got = np.array([b.index(k) for k in keys])
    assert np.array_equal(got, want)


def _ideal_spec(m, moves, bias=None):
    """Non-interacting single-site molecules in the probe map (LJ off)."""
    return dict(full_ell=list(m["eff"]), fund_ell=list(m["eff"]), ref=[[0.0, 0.0, 0.0]], stype=[0], sq=[0.0],
                nst=1, lj=([0.0], [0.0], [0.0], [0.0], [0]), T=300.0, moves=moves, dtrans=0.3,
                maps=[(m["npn"], m["step"], m["ptr"], m["vals"])], site_pmap=[0], site_emap=[-1], bias=bias)


def _mean_N(g, steps, nblk=20):
    g.run(steps // 10)  # equilibration
    means = []
    for _ in range(nblk):
        s0, n0 = g.st.sumN, g.st.nsteps
        g.run(steps // nblk)
        means.append((g.st.sumN - s0) / (g.st.nsteps - n0))
    means = np.array(means)
    return means.mean(), means.std(ddof=1) / np.sqrt(nblk)


def test_biased_moves_sample_the_grand_canonical_ensemble(probe_map):
    m = probe_map
    fug = 10000.0
    # exp(B) <exp(-U/kT)>: the map integrated over the cell at 2e5 uniform points
    rng = np.random.default_rng(11)
    pts = rng.random((200000, 3)) * m["eff"]
    ok, e = O.pothermite(m["npn"], m["step"], m["ptr"], m["vals"], pts, fund_ell=m["eff"])
    boltz = np.where(ok, np.exp(-np.where(ok, e, 0.0) / (KB * 300.0)), 0.0)
    vol = float(np.prod(m["eff"]))
    lam = np.exp(O.adams_B(fug, vol, 300.0)) * boltz.mean()
    lam_err = np.exp(O.adams_B(fug, vol, 300.0)) * boltz.std() / np.sqrt(len(boltz))
    assert 3.0 < lam < 30.0
    steps = 400000
    res = {}
    g = O.Gcmc(_ideal_spec(m, [("RINSERT", 1.0), ("RDELETE", 1.0)]), fug, 101)
    res["random"] = _mean_N(g, steps)
    for Tb in (300.0, 600.0):
        b = O.Bmap(m["npn"], m["step"], m["eff"], m["ptr"], m["vals"], Tb)
        g = O.Gcmc(_ideal_spec(m, [("BINSERT", 1.0), ("BDELETE", 1.0)], bias=b), fug, 202)
        res["biased %g K" % Tb] = _mean_N(g, steps)
        # biased insertions are accepted far more often than random ones
        assert g.st.acc[0] / g.st.att[0] > 0.3
    for k, (mean, se) in res.items():
        tol = 4.0 * np.hypot(se, lam_err) + 0.01 * lam
        assert abs(mean - lam) < tol, (k, mean, se, lam, lam_err)
