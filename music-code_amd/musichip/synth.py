"""Seeded synthetic workloads (frameworks, sorbates, state points).

Framework recipe (porous cubic lattice): cubic cells of edge a with 12 atoms
on the x = 0, y = 0 and z = 0 faces at (0, a/2 +- a/4, a/2 +- a/4) and its
permutations, jittered by a Gaussian of sigma 0.3 A.  LJ atom types: 16 types
with eps/k ~ U[20, 80] K and sigma ~ U[2.5, 3.8] A (the reference allows at
most 20 atom types, defaults.F90:96); each framework atom draws a type.
Probe-framework pairs use Lorentz-Berthelot mixing (lj.F90:927-941).
"""
import numpy as np

from . import constants as C
from .inputs import lj_AB

NTYPES_FW = 16


def lattice_faces(ncell, a, rng, jitter=0.3):
    base = []
    q = [0.25 * a, 0.75 * a]
    for u in q:
        for v in q:
            base.append((0.0, u, v))
            base.append((u, 0.0, v))
            base.append((u, v, 0.0))
    base = np.array(base)
    pts = []
    for i in range(ncell):
        for j in range(ncell):
            for k in range(ncell):
                pts.append(base + np.array([i, j, k]) * a)
    xyz = np.concatenate(pts)
    xyz = xyz + rng.normal(0.0, jitter, xyz.shape)
    L = ncell * a
    return np.mod(xyz, L)


def fw_types(rng, n, ntypes=NTYPES_FW):
    eps = rng.uniform(20.0, 80.0, ntypes)
    sig = rng.uniform(2.5, 3.8, ntypes)
    typ = rng.integers(0, ntypes, n).astype(np.int32)
    return eps, sig, typ


def lb_table(fw_eps, fw_sig, probe_eps, probe_sig, hicut, locut=C.LOW_DIST_CUTOFF):
    """Per framework type A, B, hicut2, locut2 against one probe type."""
    sig = (fw_sig + probe_sig) / 2.0
    eps = np.sqrt(fw_eps * probe_eps)
    A, B = lj_AB(sig, eps)
    n = len(fw_eps)
    return (np.asarray(A, dtype=np.float64), np.asarray(B, dtype=np.float64),
            np.full(n, hicut * hicut), np.full(n, locut * locut))


def config1(seed=1001):
    """Neutral 30 A cubic framework, 768 atoms (4x4x4 cells of 7.5 A)."""
    rng = np.random.default_rng(seed)
    xyz = lattice_faces(4, 7.5, rng)
    eps, sig, typ = fw_types(rng, len(xyz))
    L = np.array([30.0, 30.0, 30.0])
    return dict(name="C1", xyz=xyz, type=typ, q=np.zeros(len(xyz)), fw_eps=eps, fw_sig=sig,
                fund_ell=L, full_ell=L.copy(), hicut=12.8, gridspace=0.2, hipotcut=100.0,
                probe=dict(name="CH4", eps=148.0, sig=3.73))


def config2(seed=2002):
    """Charged 50 A cubic framework treated as one cell, 4096 atoms on a
    jittered 16^3 simple-cubic lattice (3.125 A), charges U[-1, 1] minus the
    mean; Ewald alpha = 0.25 / A (KAPPA@12.5 on L = 50), kmax 9."""
    rng = np.random.default_rng(seed)
    n = 16
    a = 50.0 / n
    g = (np.arange(n) + 0.5) * a
    xyz = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    xyz = np.mod(xyz + rng.normal(0.0, 0.3, xyz.shape), 50.0)
    eps, sig, typ = fw_types(rng, len(xyz))
    q = rng.uniform(-1.0, 1.0, len(xyz))
    q -= q.mean()
    L = np.array([50.0, 50.0, 50.0])
    return dict(name="C2", xyz=xyz, type=typ, q=q, fw_eps=eps, fw_sig=sig, fund_ell=L,
                full_ell=L.copy(), hicut=12.8, gridspace=0.15, hipotcut=100.0,
                kappa_param=12.5, kmax=9, ewald_locut=1e-10,
                probes=[dict(name="C", eps=27.0, sig=2.80, q=0.70),
                        dict(name="O", eps=79.0, sig=3.05, q=-0.35)])


def co2_like():
    """3-site rigid linear sorbate: C at 0, O at +-1.16 A, q_C = +0.70,
    q_O = -0.35 (TraPPE-like LJ: C 27 K / 2.80 A, O 79 K / 3.05 A)."""
    return dict(name="CO2", ref=np.array([[0.0, 0.0, -1.16], [0.0, 0.0, 0.0], [0.0, 0.0, 1.16]]),
                types=["O", "C", "O"], q=np.array([-0.35, 0.70, -0.35]),
                lj={"C": (27.0, 2.80), "O": (79.0, 3.05)}, mass=44.0)


def butane_ua():
    """4-site united-atom n-butane (TraPPE-UA): CH3-CH2-CH2-CH3, LJ CH3
    98 K / 3.75 A, CH2 46 K / 3.95 A; harmonic bonds r0 = 1.54 A with
    k/kB = 452,900 K/A^2, harmonic angle theta0 = 114 deg with
    k/kB = 62,500 K/rad^2 (U = k/2 (x - x0)^2); torsion
    c1 [1 + cos phi] + c2 [1 - cos 2 phi] + c3 [1 + cos 3 phi] with
    c1, c2, c3 = 355.03, -68.19, 791.32 K (phi = 180 deg trans), rewritten
    as sum_n c[n] x^n in the reference's polymer convention x = -cos phi
    (cosexpansion.F90:142).  Energies in kcal/mol (k_B = 8.314e-3 / 4.184)."""
    kb = C.KCALMOLE_KB
    c1, c2, c3 = 355.03, -68.19, 791.32
    tc = [(c1 + 2 * c2 + c3) * kb, -(c1 - 3 * c3) * kb, -2 * c2 * kb, -4 * c3 * kb, 0.0, 0.0]
    th0 = np.radians(114.0)
    b = 1.54
    d1 = np.array([1.0, 0.0, 0.0])
    d2 = np.array([np.cos(np.pi - th0), np.sin(np.pi - th0), 0.0])
    chain = np.array([np.zeros(3), b * d1, b * d1 + b * d2, 2 * b * d1 + b * d2])
    chain -= chain.mean(0)
    return dict(name="butane", types=["CH3", "CH2"], stype=[0, 1, 1, 0], mass=[15.035, 14.027, 14.027, 15.035],
                lj={"CH3": (98.0, 3.75), "CH2": (46.0, 3.95)},
                bonds=[(0, 1, 452900.0 * kb, b), (1, 2, 452900.0 * kb, b), (2, 3, 452900.0 * kb, b)],
                angles=[(0, 1, 2, 62500.0 * kb, th0), (1, 2, 3, 62500.0 * kb, th0)],
                torsions=[(0, 1, 2, 3, tc)], chain=chain)


def config5(seed=5000):
    """Neutral 40 A cubic framework of the methane recipe with 10 A cells
    (4 x 4 x 4 cells, 768 atoms): open enough for 50-200 n-butane chains
    (the 8 A-cell variant with 1500 atoms holds about 55)."""
    rng = np.random.default_rng(seed)
    xyz = lattice_faces(4, 10.0, rng)
    eps, sig, typ = fw_types(rng, len(xyz))
    L = np.array([40.0, 40.0, 40.0])
    return dict(name="C5", xyz=xyz, type=typ, q=np.zeros(len(xyz)), fw_eps=eps, fw_sig=sig, fund_ell=L,
                full_ell=L.copy(), hicut=12.8, gridspace=0.2, hipotcut=100.0)


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def place_chains(model, n, L, rng, sites_ok, min_dist=3.0, max_tries=200000):
    """Principal coordinates [n][ns][3] of n rigid all-trans chains at random
    positions / orientations: every site passes sites_ok (framework check,
    called on [k][3] cell positions -> bool[k]) and is at least min_dist
    (minimum image) from the sites already placed."""
    L = np.asarray(L, dtype=np.float64)
    chain = np.asarray(model["chain"])
    placed = []
    flat = np.zeros((0, 3))
    tries = 0
    while len(placed) < n and tries < max_tries:
        batch = 256
        tries += batch
        cand = []
        for _ in range(batch):
            com = rng.uniform(0.0, 1.0, 3) * L
            cand.append(chain @ random_rotation(rng).T + com)
        cand = np.array(cand)
        ok = sites_ok(np.mod(cand, L).reshape(-1, 3)).reshape(batch, -1).all(1)
        for c, good in zip(cand, ok):
            if not good or len(placed) >= n:
                continue
            if len(flat):
                d = np.mod(c[:, None, :] - flat[None, :, :] + 0.5 * L, L) - 0.5 * L
                if (np.einsum("ijk,ijk->ij", d, d) < min_dist * min_dist).any():
                    continue
            placed.append(c)
            flat = np.concatenate([flat, c])
    if len(placed) < n:
        raise RuntimeError("placed only %d of %d chains" % (len(placed), n))
    return np.array(placed)
