"""Maps of slanted (monoclinic, triclinic) cells on the GPU against the
oracle, which is pinned bit for bit to the reference's fundcell / simcell /
lj / pmap modules (tests/test_oracle_ref_cells.py): the slanted single-fold
minimum image in the node kernel (fundcell_multMinImage), the grid over the
effective box, compaction, and interpolation folding into the slanted unit
cell without the corner wrap; both map precisions run the FP64 node kernel
for slanted cells; GCMC batches refuse them."""
import numpy as np
import pytest

from oracle import oracle as O
from musichip import hip as H
from musichip import synth

pytestmark = pytest.mark.gpu

CELLS = [("monoclinic", [14.0, 13.0, 15.0], [90.0, 104.0, 90.0], [17.6, 13.0, 14.6], (1, 1, 1)),
         ("triclinic", [12.5, 13.5, 12.0], [84.0, 97.0, 108.0], [16.8, 13.0, 12.0], (1, 1, 1))]


def _system(hip, ell, ang, eff, nc, seed):
    rng = np.random.default_rng(seed)
    c = O.cell(ell, ang, eff, nc)
    n = 90
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


def test_gcmc_refuses_slanted_cells(hip):
    from musichip import gcmc as G
    name, ell, ang, eff, nc = CELLS[0]
    _system(hip, ell, ang, eff, nc, 33)
    spec = dict(full_ell=ell, fund_ell=ell, ref=[[0, 0, 0]], stype=[0], sq=[0.0], nst=1,
                lj=([1e6], [1e3], [25.0], [0.25], [1]), T=300.0, moves=[("RINSERT", 1.0), ("RDELETE", 1.0)])
    with pytest.raises(H.MusicHipError, match="slanted"):
        G.GcmcBatch(hip, spec, [100.0], [1], maxN=16)
