"""The oracle's GCMC of mixtures (gcmc_dosim, gcmc.F90:224-258): one species
drawn per step, its move set, its molecules against those of every species.

Pinned by what the mathematics fixes: the total energy of a two-species
configuration against a direct double sum over all site pairs (LJ + plain
Coulomb, cross-species pairs through the common site-type table), the
running energy against that recomputation after a trajectory, the species
draw frequencies, and a species split into two identical halves at half the
fugacity each reproducing the single-species loading.  The single-species
entry points are the mixture routines with one species (test007 and the
other single-species tests pin that path against the reference)."""
import numpy as np

from oracle import oracle as O
from musichip import inputs

L = np.array([18.0, 18.0, 18.0])


def _tables():
    # site types: 0 = CH4 united atom, 1 = X (two-site), 2 = Y (two-site)
    eps = np.array([148.0, 70.0, 40.0])
    sig = np.array([3.73, 3.30, 3.10])
    A, B = [], []
    for i in range(3):
        for j in range(3):
            a, b = inputs.lj_AB((sig[i] + sig[j]) / 2, np.sqrt(eps[i] * eps[j]))
            A.append(a)
            B.append(b)
    return A, B


def _specs():
    A, B = _tables()
    lj = (A, B, [8.0 ** 2] * 9, [0.5 ** 2] * 9, [1] * 9)
    coul = dict(on=[0, 0, 0, 0, 1, 1, 0, 1, 1], hicut=8.0, locut=0.5, shield=1.0)
    common = dict(full_ell=L, fund_ell=L, nst=3, lj=lj, coul=coul, T=300.0)
    ch4 = dict(common, ref=[[0.0, 0.0, 0.0]], stype=[0], sq=[0.0],
               moves=[("RINSERT", 1.0), ("RDELETE", 1.0), ("RTRANSLATE", 1.0)], dtrans=0.6)
    xy = dict(common, ref=[[0.0, 0.0, -0.6], [0.0, 0.0, 0.6]], stype=[1, 2], sq=[0.35, -0.35],
              moves=[("RINSERT", 1.0), ("RDELETE", 1.0), ("RTRANSLATE", 1.0), ("RROTATE", 1.0)],
              dtrans=0.5, drot=0.4)
    return ch4, xy


def _direct_total(specs, coms, angs):
    """Double sum over molecule pairs (each once) of site-site LJ + Coulomb."""
    A, B, h2, l2, on = specs[0]["lj"]
    cl = specs[0]["coul"]
    nst = specs[0]["nst"]
    sites = []
    for sp, com, ang in zip(specs, coms, angs):
        for c, a in zip(com, ang):
            r = np.zeros((len(sp["ref"]), 3))
            O.lib()
            r = O.euler_xyz(np.asarray(sp["ref"], dtype=float), c, a)
            sites.append([(r[k], sp["stype"][k], sp["sq"][k]) for k in range(len(sp["ref"]))])
    e = 0.0
    for m in range(len(sites)):
        for n in range(m):
            for ra, ta, qa in sites[m]:
                for rb, tb, qb in sites[n]:
                    d = ra - rb
                    d -= L * np.round(d / L)
                    r2 = float(d @ d)
                    t = ta * nst + tb
                    if on[t] and r2 <= h2[t]:
                        e += A[t] / r2 ** 6 - B[t] / r2 ** 3
                    if cl["on"][t] and r2 <= cl["hicut"] ** 2:
                        e += 331.5 * cl["shield"] * qa * qb * (1.0 / np.sqrt(r2) - 1.0 / cl["hicut"])
    return e


def test_mixture_total_energy_equals_direct_sum():
    specs = _specs()
    rng = np.random.default_rng(4)
    coms = [rng.uniform(0, 1, (9, 3)) * L, rng.uniform(0, 1, (7, 3)) * L]
    angs = [np.zeros((9, 3)), np.stack([rng.uniform(-np.pi, np.pi, 7), np.arccos(rng.uniform(-1, 1, 7)),
                                        rng.uniform(-np.pi, np.pi, 7)], 1)]
    # keep the random configuration free of overlaps (locut and LJ walls)
    g = O.GcmcMix(list(specs), [100.0, 100.0], 5, maxN=64)
    g.set_config(coms, angs)
    e = g.total_energy()
    ref = _direct_total(specs, coms, angs)
    assert np.isfinite(e)
    assert abs(e - ref) <= 1e-9 * max(1.0, abs(ref)), (e, ref)


def test_mixture_running_energy_and_species_draws():
    specs = _specs()
    g = O.GcmcMix(list(specs), [3.0e3, 1.0e4], 77, maxN=256, trace=6000)
    g.run(6000)
    N = g.N()
    assert min(N) >= 3, N
    tot = g.total_energy()
    assert abs(tot - g.energy) <= 1e-8 * max(1.0, abs(tot))
    tr = g.trace_array(6000)
    spc = tr[:, 0].astype(int) // 16
    mv = tr[:, 0].astype(int) % 16
    # each species about half of the steps; CH4 has no rotation move
    assert abs((spc == 0).mean() - 0.5) < 0.03
    assert not np.any((spc == 0) & (mv == 3))
    # accepted moves of both species
    assert tr[(spc == 0) & (tr[:, 1] == 1)].shape[0] > 50 and tr[(spc == 1) & (tr[:, 1] == 1)].shape[0] > 50


def test_identical_halves_reproduce_single_species_loading():
    ch4, _ = _specs()
    f = 2.0e4
    steps, burn = 60000, 10000
    one = O.Gcmc(ch4, f, 91, maxN=512)
    one.run(burn)
    s0 = one.st.sumN
    one.run(steps)
    n_one = (one.st.sumN - s0) / steps
    mix = O.GcmcMix([ch4, ch4], [f / 2, f / 2], 91, maxN=512)
    mix.run(2 * burn)
    a0 = [c.st.sumN for c in mix.ch]
    mix.run(2 * steps)
    n_mix = sum((c.st.sumN - a) for c, a in zip(mix.ch, a0)) / (2 * steps)
    # half of the steps move each half: the same chemical potential overall
    assert n_one > 10
    assert abs(n_mix - n_one) < 0.08 * n_one, (n_mix, n_one)
