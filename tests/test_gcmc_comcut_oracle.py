"""Centre-of-mass cutoffs of the sorbate pair terms in the oracle (ssbasic
COMCUT@r / COMTRUNC@r: ssbasic_mmint, ssbasic.F90:745-756, skips a molecule
pair's LJ or Coulomb term when the minimum-image distance of the molecules'
geometric centres, config_getMolecCTR at config.F90:1788-1802, exceeds r).

Pinned by what the definition fixes: the total energy against a direct
double sum that applies the cutoff to the mean of the unfolded site
positions; a pair placed so that its centre-of-mass distance and its
geometric-centre distance straddle the cutoff (a cutoff on the centre of
mass, or on the folded sites, gives the other answer); an infinite cutoff
equal to none and a vanishing one removing every pair term; the running
energy of a trajectory equal to the recomputation; per-species-pair radii
in a mixture."""
import numpy as np
import pytest

from oracle import oracle as O
from musichip import inputs

L = np.array([16.0, 16.0, 16.0])
# a bent 3-site molecule on its principal axes (masses 16, 1, 1): its
# geometric centre lies 0.54 A from the centre of mass
BENT = [[0.0, 0.065, 0.0], [0.8165, -0.5128, 0.0], [-0.8165, -0.5128, 0.0]]


def _tables():
    eps = np.array([78.2, 10.0, 148.0])
    sig = np.array([3.166, 1.2, 3.73])
    A, B = [], []
    for i in range(3):
        for j in range(3):
            a, b = inputs.lj_AB((sig[i] + sig[j]) / 2, np.sqrt(eps[i] * eps[j]))
            A.append(a)
            B.append(b)
    return (A, B, [7.5 ** 2] * 9, [0.3 ** 2] * 9, [1] * 9)


def _bent(comcut=None, coul=True):
    sp = dict(full_ell=L, fund_ell=L, nst=3, lj=_tables(), T=300.0, ref=BENT, stype=[0, 1, 1],
              sq=[-0.8, 0.4, 0.4], moves=[("RINSERT", 1.0), ("RDELETE", 1.0), ("RTRANSLATE", 1.0),
                                           ("RROTATE", 1.0)], dtrans=0.5, drot=0.5)
    if coul:
        sp["coul"] = dict(on=[1, 1, 0, 1, 1, 0, 0, 0, 0], hicut=7.5, locut=0.3, shield=1.0)
    if comcut:
        sp["comcut"] = comcut
    return sp


def _ch4(comcut=None):
    sp = dict(full_ell=L, fund_ell=L, nst=3, lj=_tables(), T=300.0, ref=[[0.0, 0.0, 0.0]], stype=[2], sq=[0.0],
              moves=[("RINSERT", 1.0), ("RDELETE", 1.0), ("RTRANSLATE", 1.0)], dtrans=0.6)
    if comcut:
        sp["comcut"] = comcut
    return sp


def _radius(cc, key, s2):
    v = (cc or {}).get(key, 0.0)
    return v[s2] if np.ndim(v) else v


def _direct(specs, coms, angs):
    """Sum over molecule pairs (each once) of site-site LJ + plain Coulomb,
    a pair's LJ / Coulomb term dropped when the minimum-image distance of
    the two means of unfolded site positions exceeds its radius."""
    A, B, h2, l2, on = specs[0]["lj"]
    cl = specs[0].get("coul")
    nst = specs[0]["nst"]
    mols = []
    for s, (sp, com, ang) in enumerate(zip(specs, coms, angs)):
        for c, a in zip(com, ang):
            rp = O.euler_xyz(np.asarray(sp["ref"], dtype=float), c, a)
            mols.append((s, rp, rp.mean(0)))
    e = 0.0
    for m in range(len(mols)):
        for n in range(m):
            sm, rm, cm = mols[m]
            sn, rn, cn = mols[n]
            d = cm - cn
            d -= L * np.round(d / L)
            c2 = float(d @ d)
            rl = _radius(specs[sm].get("comcut"), "lj", sn)
            rc = _radius(specs[sm].get("comcut"), "coul", sn)
            lj_on = not (rl and c2 > rl * rl)
            cl_on = not (rc and c2 > rc * rc)
            for a, ra in enumerate(rm):
                for b, rb in enumerate(rn):
                    d = ra - rb
                    d -= L * np.round(d / L)
                    r2 = float(d @ d)
                    t = specs[sm]["stype"][a] * nst + specs[sn]["stype"][b]
                    if lj_on and on[t] and r2 <= h2[t]:
                        e += A[t] / r2 ** 6 - B[t] / r2 ** 3
                    if cl is not None and cl_on and cl["on"][t] and r2 <= cl["hicut"] ** 2:
                        e += 331.5 * cl["shield"] * specs[sm]["sq"][a] * specs[sn]["sq"][b] * (
                            1.0 / np.sqrt(r2) - 1.0 / cl["hicut"])
    return e


def _config(rng, n):
    while True:
        com = rng.uniform(0, 1, (n, 3)) * L
        d = com[:, None] - com[None]
        d -= L * np.round(d / L)
        r = np.sqrt((d ** 2).sum(-1)) + 100 * np.eye(n)
        if r.min() > 3.2:
            break
    ang = np.stack([rng.uniform(-np.pi, np.pi, n), np.arccos(rng.uniform(-1, 1, n)),
                    rng.uniform(-np.pi, np.pi, n)], 1)
    return com, ang


@pytest.mark.parametrize("cc", [dict(lj=5.0), dict(coul=6.0), dict(lj=5.5, coul=4.5)])
def test_total_energy_equals_direct_sum_with_centre_cutoff(cc):
    rng = np.random.default_rng(12)
    com, ang = _config(rng, 14)
    sp = _bent(cc)
    g = O.Gcmc(sp, 100.0, 3, maxN=64)
    g.set_config(com, ang)
    e = g.total_energy()
    ref = _direct([sp], [com], [ang])
    none = _direct([_bent()], [com], [ang])
    assert abs(e - ref) <= 1e-9 * max(1.0, abs(ref)), (e, ref)
    assert abs(ref - none) > 1e-3  # the cutoff removed terms


def test_cutoff_is_on_the_geometric_centre():
    """Two bent molecules, x axis apart: centres of mass 5.0 A apart, their
    geometric centres (0.54 A below the centre of mass along the body y
    axis, turned along x) 5.0 +- 1.08 A; a 5.5 A radius keeps the pair in
    one orientation and drops it in the other."""
    half_pi = np.pi / 2
    # psi rotates body y onto -x (first molecule) and +x (second) or back
    out = np.array([[-half_pi, 0.0, 0.0], [half_pi, 0.0, 0.0]])
    inn = np.array([[half_pi, 0.0, 0.0], [-half_pi, 0.0, 0.0]])
    com = np.array([[5.0, 8.0, 8.0], [10.0, 8.0, 8.0]])
    res = {}
    for name, ang in (("out", out), ("in", inn)):
        rp = [O.euler_xyz(np.asarray(BENT), c, a) for c, a in zip(com, ang)]
        res[name] = (np.linalg.norm(rp[1].mean(0) - rp[0].mean(0)), ang)
    assert res["out"][0] > 5.5 > 5.0 > res["in"][0] or res["in"][0] > 5.5 > 5.0 > res["out"][0]
    for name, (dist, ang) in res.items():
        g = O.Gcmc(_bent(dict(lj=5.5, coul=5.5)), 100.0, 3, maxN=8)
        g.set_config(com, ang)
        e = g.total_energy()
        full = _direct([_bent()], [com], [ang])
        assert abs(full) > 1e-3
        if dist > 5.5:
            assert e == 0.0
        else:
            assert abs(e - full) <= 1e-12 * abs(full)


def test_infinite_and_vanishing_radius():
    rng = np.random.default_rng(8)
    com, ang = _config(rng, 12)
    base = O.Gcmc(_bent(), 100.0, 3, maxN=64)
    base.set_config(com, ang)
    big = O.Gcmc(_bent(dict(lj=1.0e3, coul=1.0e3)), 100.0, 3, maxN=64)
    big.set_config(com, ang)
    tiny = O.Gcmc(_bent(dict(lj=1.0e-3, coul=1.0e-3)), 100.0, 3, maxN=64)
    tiny.set_config(com, ang)
    assert big.total_energy() == base.total_energy()
    assert tiny.total_energy() == 0.0 and base.total_energy() != 0.0


def test_running_energy_with_centre_cutoff():
    g = O.Gcmc(_bent(dict(lj=6.0, coul=5.0)), 2.0e4, 41, maxN=256, trace=8000)
    g.run(8000)
    assert g.st.N > 8
    tot = g.total_energy()
    assert abs(tot - g.st.energy) <= 1e-8 * max(1.0, abs(tot))
    tr = g.trace_array(8000)
    assert (tr[:, 1] == 1).sum() > 500


def test_mixture_per_species_pair_radius():
    """LJ between the bent molecules and CH4 cut at 5 A of the centres,
    bent-bent and CH4-CH4 pairs without a cutoff."""
    rng = np.random.default_rng(21)
    com0, ang0 = _config(rng, 8)
    while True:
        com1 = rng.uniform(0, 1, (6, 3)) * L
        allc = np.concatenate([com0, com1])
        d = allc[:, None] - allc[None]
        d -= L * np.round(d / L)
        if (np.sqrt((d ** 2).sum(-1)) + 100 * np.eye(len(allc))).min() > 3.2:
            break
    specs = [_bent(dict(lj=[0.0, 5.0]), coul=False), _ch4(dict(lj=[5.0, 0.0]))]
    g = O.GcmcMix(specs, [100.0, 100.0], 9, maxN=32)
    g.set_config([com0, com1], [ang0, np.zeros((6, 3))])
    e = g.total_energy()
    ref = _direct(specs, [com0, com1], [ang0, np.zeros((6, 3))])
    none = _direct([_bent(coul=False), _ch4()], [com0, com1], [ang0, np.zeros((6, 3))])
    assert abs(e - ref) <= 1e-9 * max(1.0, abs(ref)), (e, ref)
    assert abs(ref - none) > 1e-3
    g.run(4000)
    tot = g.total_energy()
    assert abs(tot - g.energy) <= 1e-8 * max(1.0, abs(tot))
