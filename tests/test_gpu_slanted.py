"""Maps of slanted (monoclinic, triclinic) cells on the GPU against the
oracle, which is pinned bit for bit to the reference's fundcell / simcell /
lj / pmap modules (tests/test_oracle_ref_cells.py): the slanted single-fold
minimum image in the node kernel (fundcell_multMinImage), the grid over the
effective box, compaction, and interpolation folding into the slanted unit
cell without the corner wrap; both map precisions run the FP64 node kernel
for slanted cells; GCMC batches in slanted cells against the oracle chain
(without sorbate Ewald sums, which they refuse)."""
import numpy as np
import pytest

from oracle import oracle as O
from musichip import hip as H
from musichip import synth

pytestmark = pytest.mark.gpu

CELLS = [("monoclinic", [14.0, 13.0, 15.0], [90.0, 104.0, 90.0], [17.6, 13.0, 14.6], (1, 1, 1)),
         ("triclinic", [12.5, 13.5, 12.0], [84.0, 97.0, 108.0], [16.8, 13.0, 12.0], (1, 1, 1))]


def _system(hip, ell, ang, eff, nc, seed, n=90):
    rng = np.random.default_rng(seed)
    c = O.cell(ell, ang, eff, nc)
    # atoms uniform in the slanted cell: fractional coordinates times the lattice vectors
    U, S, lv = O.cell_matrices(c)
    frac = rng.uniform(0, 1, (n, 3))
    xyz = frac @ lv
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    typ = rng.integers(0, synth.NTYPES_FW, n).astype(np.int32)
    eps, sig = rng.uniform(20, 80, synth.NTYPES_FW), rng.uniform(2.5, 3.8, synth.NTYPES_FW)
    full = np.array(ell) * np.array(nc)
    hip.upload_framework(xyz, typ, q, np.array(ell, dtype=float), full)
    hip.set_cell(ang, eff)
    # hicut below half the shortest cell width keeps the single fold exact
    tab = synth.lb_table(eps, sig, 148.0, 3.73, 5.0)
    hip.set_lj(*tab)
    return c, xyz, typ, q, tab, full


@pytest.mark.parametrize("name,ell,ang,eff,nc", CELLS)
def test_slanted_lj_map_matches_oracle(hip, name, ell, ang, eff, nc):
    print("case: slanted LJ map, %s cell" % name, flush=True)
    c, xyz, typ, q, tab, full = _system(hip, ell, ang, eff, nc, 31)
    gs = 0.4
    m64 = hip.build_map(H.MH_MAP_LJ, H.MH_PREC_FP64, gs, 100.0, keep_full=True)
    npn = (np.array(eff) / gs).astype(int) + 1
    assert list(m64.npnts) == list(npn)
    step = np.array(eff) / (npn - 1)
    A, B, h2, l2 = tab
    ref = O.build_nodes(1, dict(xyz=xyz, type=typ, A=A, B=B, hicut2=h2, locut2=l2, full_ell=full, cell=c), npn,
                        step)
    _, f64 = m64.get_full()
    ok = ref[:, 0] < 1e9
    assert ok.sum() > 1000
    assert np.array_equal(f64[ok], ref[ok]), np.abs(f64[ok] - ref[ok]).max()
    # compaction on the reference's FP32-rounded potential
    ptr_o, _, counts = O.compact(npn, ref[:, 0].astype(np.float32), 100.0 / 4.184)
    ptr, vals = m64.get()
    assert np.array_equal(ptr, ptr_o) and list(m64.counts) == list(counts)
    # FP32 requests run the FP64 kernel for slanted cells: the same map
    m32 = hip.build_map(H.MH_MAP_LJ, H.MH_PREC_FP32, gs, 100.0)
    p32, v32 = m32.get()
    assert np.array_equal(p32, ptr) and np.array_equal(v32, vals)
    # interpolation: fold into the slanted cell, no corner wrap
    rng = np.random.default_rng(5)
    # without the corner wrap only points whose unslanted image lies inside
    # the node range interpolate (a few percent in the triclinic cell)
    pts = rng.uniform(-0.5, 1.5, (20000, 3)) @ O.cell_matrices(c)[2]
    oki, ei = m64.interp(pts, 1.0)
    oko, eo = O.pothermite(npn, step, ptr, vals, pts, cell=c)
    assert np.array_equal(oki, oko) and oki.sum() > 500
    assert np.array_equal(ei[oki], eo[oko])
    m32.close()
    m64.close()


def test_slanted_ewald_map_matches_oracle(hip):
    name, ell, ang, eff, nc = CELLS[1]
    print("case: slanted Ewald map, %s cell" % name, flush=True)
    c, xyz, typ, q, tab, full = _system(hip, ell, ang, eff, nc, 32)
    nk = hip.set_ewald(5.0, 5, 1e-10)
    sf = O.ewald_sfactor(xyz, q, full, 5.0, 5, 1e-10)
    assert nk == sf["nk"]
    gs = 0.6
    m = hip.build_map(H.MH_MAP_EWALD, H.MH_PREC_FP64, gs, 100.0, probe_q=0.7, keep_full=True)
    npn = (np.array(eff) / gs).astype(int) + 1
    step = np.array(eff) / (npn - 1)
    ref = O.build_nodes(2, dict(xyz=xyz, q=q, charge1=0.7, kappa=sf["kappa"], k=sf["k"], s=sf["s"], si=sf["si"],
                                full_ell=full, cell=c), npn, step)
    _, f64 = m.get_full()
    scale = np.abs(ref).max(0)
    err = np.abs(f64 - ref) / (np.abs(ref) + 1e-6 * scale)
    print("slanted Ewald FP64 max scaled rel err %.3e" % err.max(), flush=True)
    assert err.max() < 1e-10
    m.close()


def _gcmc_spec(hip, name, ell, ang, eff, nc, seed):
    """A 2-site rigid sorbate (sites 1.1 A apart, one LJ map) in a slanted
    framework cell, sorbate-sorbate LJ within 5 A and plain Coulomb."""
    from musichip import inputs
    # a sparse framework (30 atoms): room for the sorbate
    c, xyz, typ, q, tab, full = _system(hip, ell, ang, eff, nc, seed, n=30)
    m = hip.build_map(H.MH_MAP_LJ, H.MH_PREC_FP64, 0.3, 100.0)
    A, B = inputs.lj_AB(3.3, 36.0)
    spec = dict(full_ell=full, fund_ell=np.array(ell, dtype=float), ref=[[0.0, 0.0, -0.55], [0.0, 0.0, 0.55]],
                stype=[0, 0], sq=[0.4, -0.4], nst=1, lj=([A], [B], [5.0 ** 2], [0.5 ** 2], [1]),
                coul=dict(on=[1], hicut=5.0, locut=0.5, shield=1.0), T=300.0,
                moves=[("RINSERT", 1.0), ("RDELETE", 1.0), ("RTRANSLATE", 1.0), ("RROTATE", 1.0)],
                dtrans=0.5, drot=0.4, cell=c)
    return spec, {0: (m, None), 1: (m, None)}, m


@pytest.mark.parametrize("name,ell,ang,eff,nc", CELLS)
def test_gcmc_trajectory_in_slanted_cell_matches_oracle(hip, name, ell, ang, eff, nc):
    """Insertion by the lattice vectors (simcell_uniformRandPt), slanted pbc
    of the sites, slanted minimum image of sorbate pairs and map lookups
    folded into the slanted unit cell: the FP64 move sequence equals the
    oracle's, whose slanted pieces match the reference build bit for bit."""
    from musichip import gcmc as G
    from test_gpu_gcmc import _oracle_for
    print("case: GCMC of a 2-site sorbate in the %s cell, FP64 trajectory parity" % name, flush=True)
    spec, maps, m = _gcmc_spec(hip, name, ell, ang, eff, nc, 41)
    fugs, seeds, nsteps = [1.0e5, 1.0e6], [5151, 5152], 4000
    ref = []
    for r in range(2):
        g = _oracle_for(spec, maps, fugs[r], seeds[r])
        g.run(nsteps)
        ref.append((g.trace_array(nsteps), g.st.N, g.st.energy))
    b = G.GcmcBatch(hip, spec, fugs, seeds, maxN=256, maps=maps)
    b.set_trace(nsteps)
    b.run(nsteps)
    s = b.state()
    for r in range(2):
        to, No, eo = ref[r]
        tg = b.trace(r)
        assert np.array_equal(tg[:, :3].astype(int), to[:, :3].astype(int)), r
        scale = np.maximum(np.abs(to[:, 3]), 1.0)
        assert np.max(np.abs(tg[:, 3] - to[:, 3]) / scale) < 1e-10
        assert No == s["N"][r]
        assert abs(eo - s["energy"][r]) < 1e-9 * max(1.0, abs(eo))
    assert max(s["N"]) >= 5, list(s["N"])
    tot = b.total_energy()
    assert np.allclose(tot, s["energy"], rtol=1e-10, atol=1e-9)
    # FP32 kernel: per-trial energies of the same configurations within 1e-5
    b32 = G.GcmcBatch(hip, spec, fugs, seeds, maxN=256, maps=maps, precision=H.MH_PREC_FP32)
    rng = np.random.default_rng(3)
    U, S_, lv = O.cell_matrices(spec["cell"])
    n = 0
    for r in range(2):
        com, ang_, _ = b.config(r)
        b32.set_config(r, com, ang_)
        tc = rng.uniform(0, 1, (2000, 3)) @ lv
        ta = np.stack([rng.uniform(-np.pi, np.pi, 2000), np.arccos(rng.uniform(-1, 1, 2000)),
                       rng.uniform(-np.pi, np.pi, 2000)], 1)
        ok64, e64 = b.trial_energy(r, tc, ta)
        ok32, e32 = b32.trial_energy(r, tc, ta)
        assert np.count_nonzero(ok64 != ok32) <= 4
        both = ok64 & ok32
        n += int(both.sum())
        assert np.max(np.abs(e32[both] - e64[both]) / np.maximum(np.abs(e64[both]), 1.0)) <= 1e-5
    assert n > 200
    b32.close()
    b.close()
    m.close()


def test_gcmc_refuses_ewald_and_bias_in_slanted_cells(hip):
    from musichip import gcmc as G
    name, ell, ang, eff, nc = CELLS[0]
    spec, maps, m = _gcmc_spec(hip, name, ell, ang, eff, nc, 33)
    spec = dict(spec, coul=None, ewald=dict(kappa=0.3, kmax=5, rc=5.0))
    with pytest.raises(H.MusicHipError, match="slanted"):
        G.GcmcBatch(hip, spec, [100.0], [1], maxN=16, maps=maps)
    m.close()
