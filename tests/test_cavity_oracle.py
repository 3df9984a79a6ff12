"""The oracle's cavity-biased insertion and deletion (CAVITY_BIAS /
MAP_CAVITY_BIAS: rigidmoves_getConfigFromCavBias, rigidmoves.F90:1987-2126,
rigidmoves_bdelete :1632-1643, cavity list cavitylist.F90).

Pinned by what the mathematics fixes: the incremental cube occupancy
against a recount from the final positions after trajectories with every
move type (and a mixture whose second species is not in the cavity list),
the running average cav_prob against its closed form when every trial
succeeds, insertion points confined to the tabulated cubes (a transposed
cube index puts them elsewhere), and the grand canonical mean of an ideal
gas confined by the tabulation to a quarter of the cell: (1 - (3/4)^T)
e^B / 4 only when the acceptance carries cav_prob / ntrials on insertion and
its inverse on deletion (without it the mean is 4 times larger, inverted 4
times smaller), and MAP_CAVITY_BIAS for an ideal gas in a map field at the
unbiased mean e^B <exp(-U/kT)>."""
import numpy as np

from oracle import oracle as O
from musichip import inputs

L = np.array([18.0, 18.0, 18.0])


def _lj(nst, eps, sig, hicut=8.0):
    A, B = [], []
    for i in range(nst):
        for j in range(nst):
            a, b = inputs.lj_AB((sig[i] + sig[j]) / 2, np.sqrt(eps[i] * eps[j]))
            A.append(a)
            B.append(b)
    return (A, B, [hicut ** 2] * nst * nst, [0.5 ** 2] * nst * nst, [1] * nst * nst)


def _recount(g, cav, ns):
    n = np.asarray(cav["n"])
    d = L / n
    cnt = np.zeros(int(np.prod(n)), dtype=np.int64)
    for m in range(g.st.N):
        for a in range(ns):
            k = np.minimum((g.r[m, a] / d).astype(int), n - 1)
            cnt[(k[0] * n[1] + k[1]) * n[2] + k[2]] += 1
    return cnt


def _xy_spec(cav, ntrials=12):
    return dict(full_ell=L, fund_ell=L, ref=[[0.0, 0.0, -0.6], [0.0, 0.0, 0.6]], stype=[0, 1], sq=[0.0, 0.0], nst=2,
                lj=_lj(2, [70.0, 40.0], [3.3, 3.1]), T=300.0, cavity=cav, cav_ntrials=ntrials,
                moves=[("BINSERT", 1.0), ("BDELETE", 1.0), ("RTRANSLATE", 1.0), ("RROTATE", 1.0)],
                dtrans=0.7, drot=0.5)


def test_cube_occupancy_matches_recount():
    cav = dict(n=(9, 12, 6), tab=None, mask=1)
    g = O.Gcmc(_xy_spec(cav), 5.0e4, 13, maxN=256)
    for _ in range(4):
        g.run(5000)
        assert np.array_equal(g.cube_cnt, _recount(g, cav, 2))
    assert g.st.N > 20 and g.st.acc[1] > 100 and g.st.acc[2] > 100 and g.st.acc[3] > 100
    # set_config rebuilds the list
    rng = np.random.default_rng(2)
    g.set_config(rng.uniform(0, 1, (15, 3)) * L, rng.uniform(-1, 1, (15, 3)))
    assert np.array_equal(g.cube_cnt, _recount(g, cav, 2))
    assert abs(g.total_energy() - g.st.energy) <= 1e-9 * max(1.0, abs(g.st.energy))


def test_mixture_species_outside_the_list_do_not_occupy_cubes():
    cav = dict(n=(9, 9, 9), tab=None, mask=1)
    ch4 = dict(full_ell=L, fund_ell=L, ref=[[0.0, 0.0, 0.0]], stype=[2], sq=[0.0], T=300.0,
               moves=[("RINSERT", 1.0), ("RDELETE", 1.0), ("RTRANSLATE", 1.0)], dtrans=0.6)
    xy = _xy_spec(cav)
    xy["stype"] = [0, 1]
    glob = dict(nst=3, lj=_lj(3, [70.0, 40.0, 148.0], [3.3, 3.1, 3.73]), cavity=cav)
    mix = O.GcmcMix([dict(xy, **glob), dict(ch4, **glob)], [2.0e4, 2.0e4], 5, maxN=256)
    mix.run(20000)
    assert min(mix.N()) > 5
    assert np.array_equal(mix.ch[0].cube_cnt, _recount(mix.ch[0], cav, 2))
    tot = mix.total_energy()
    assert abs(tot - mix.energy) <= 1e-8 * max(1.0, abs(tot))


def test_cav_prob_running_average_closed_form():
    """Every trial succeeds in an empty, fully tabulated cell and every
    insertion is rejected (max_nrg below zero): after k attempts cav_prob(1)
    = (T / 2 + k T) / (k + 1) over k + 1 updates."""
    T = 7
    cav = dict(n=(6, 6, 6), tab=None, mask=1)
    spec = dict(full_ell=L, fund_ell=L, ref=[[0.0, 0.0, 0.0]], stype=[0], sq=[0.0], nst=1,
                lj=([0.0], [0.0], [0.0], [0.0], [0]), T=300.0, cavity=cav, cav_ntrials=T, max_nrg=-1.0,
                moves=[("BINSERT", 1.0), ("BDELETE", 1.0)])
    g = O.Gcmc(spec, 1.0e3, 3, maxN=16)
    g.run(4000)
    k = g.st.att[0]
    assert k > 1500 and g.st.N == 0
    assert g.cav_n[1] == k + 1
    assert abs(g.cav_avg[1] - (T / 2.0 + k * T) / (k + 1)) <= 1e-12 * T
    # untouched loadings keep their start value
    assert g.cav_n[2] == 1 and g.cav_avg[2] == T / 2.0


def _quarter_gas(fug, seed, T=10):
    n = (40, 40, 40)
    tab = np.zeros(n, dtype=np.int8)
    tab[: n[0] // 4] = 1  # x below L / 4
    cav = dict(n=n, tab=tab, mask=1)
    spec = dict(full_ell=L, fund_ell=L, ref=[[0.0, 0.0, 0.0]], stype=[0], sq=[0.0], nst=1,
                lj=([0.0], [0.0], [0.0], [0.0], [0]), T=300.0, cavity=cav, cav_ntrials=T,
                moves=[("BINSERT", 1.0), ("BDELETE", 1.0)])
    return O.Gcmc(spec, fug, seed, maxN=512), cav


def test_insertions_stay_in_tabulated_cubes():
    g, cav = _quarter_gas(2.0e5, 21)
    g.run(3000)
    assert g.st.N > 20
    x = g.com[: g.st.N]
    assert np.all(x[:, 0] < L[0] / 4) and x[:, 1].max() > L[1] / 2 and x[:, 2].max() > L[2] / 2
    assert np.array_equal(g.cube_cnt, _recount(g, cav, 1))


def test_ideal_gas_in_a_quarter_of_the_cell():
    fug = 1.2e5
    eB = np.exp(O.adams_B(fug, float(np.prod(L)), 300.0))
    means = []
    for seed in range(8):
        g, _ = _quarter_gas(fug, 100 + seed)
        g.run(20000)
        s0, n0 = g.st.sumN, g.st.nsteps
        g.run(200000)
        means.append((g.st.sumN - s0) / (g.st.nsteps - n0))
        # the running average approaches the tabulated fraction
        t = int(round(means[-1]))
        assert abs(g.cav_avg[t] / 10.0 - 0.25) < 0.02
    m, se = np.mean(means), np.std(means, ddof=1) / np.sqrt(len(means))
    # an insertion finds no cavity with probability (3/4)^T and then fails
    # (the reference does not fall back to a plain insertion): the loading
    # is Poisson with mean (1 - (3/4)^T) e^B / 4
    lam = (1.0 - 0.75 ** 10) * eB / 4
    assert lam > 20
    assert abs(m - lam) < 4 * se + 0.01 * lam, (m, se, lam)


def _wave_map(n=25, amp=1.0):
    """A tabulated map over the cell with V = amp cos(2 pi x / L) and its x
    derivative (the other columns zero), in the map-file layout."""
    npn = (n, n, n)
    step = L / (n - 1)
    nf = n ** 3
    ptr = np.arange(1, nf + 1, dtype=np.int32)
    x = (np.arange(nf) // (n * n)) * step[0]
    vals = np.zeros((8, nf), dtype=np.float32)
    vals[0] = amp * np.cos(2 * np.pi * x / L[0])
    vals[1] = -amp * (2 * np.pi / L[0]) * np.sin(2 * np.pi * x / L[0])
    return npn, step, ptr, vals


def test_map_cavity_ideal_gas_mean():
    """MAP_CAVITY_BIAS for non-interacting molecules in a map field: nearly
    every insertion finds a cavity (small cubes, every cube tabulated), so
    the mean loading is e^B <exp(-U/kT)> as for unbiased insertion — only
    when the map bias of the chosen trial enters with cav_prob / ntrials
    (dropping the map bias shifts the mean by the field's Boltzmann
    weights)."""
    npn, step, ptr, vals = _wave_map()
    bias = O.Bmap(npn, step, L, ptr, vals, 300.0)
    fug = 1.0e5
    eB = np.exp(O.adams_B(fug, float(np.prod(L)), 300.0))
    rng = np.random.default_rng(8)
    ok, e = O.pothermite(npn, step, ptr, vals, rng.random((20000, 3)) * L, fund_ell=L)
    assert ok.all()
    lam = eB * np.mean(np.exp(-e / (8.314e-3 / 4.184 * 300.0)))
    cav = dict(n=(40, 40, 40), tab=None, mask=1)
    spec = dict(full_ell=L, fund_ell=L, ref=[[0.0, 0.0, 0.0]], stype=[0], sq=[0.0], nst=1,
                lj=([0.0], [0.0], [0.0], [0.0], [0]), T=300.0, maps=[(npn, step, ptr, vals)], site_pmap=[0],
                site_emap=[-1], bias=bias, cavity=cav, cav_ntrials=10, cav_map=1,
                moves=[("BINSERT", 1.0), ("BDELETE", 1.0)])
    means = []
    for seed in range(6):
        g = O.Gcmc(spec, fug, 300 + seed, maxN=512)
        g.run(20000)
        s0, n0 = g.st.sumN, g.st.nsteps
        g.run(150000)
        means.append((g.st.sumN - s0) / (g.st.nsteps - n0))
    m, se = np.mean(means), np.std(means, ddof=1) / np.sqrt(len(means))
    assert lam > 20
    assert abs(m - lam) < 4 * se + 0.02 * lam, (m, se, lam)
