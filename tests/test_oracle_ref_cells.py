"""Slanted (monoclinic, triclinic) and orthorhombic cells in the oracle
against the reference's own fundcell / simcell / lj / pmap modules
(oracle/_ref, vectors in tests/golden/ref_vectors.json): the unslant matrix
from float32 angles and cosines (fundcell_initxform, fundcell.F90:378-412),
its LU inverse (matrix_inverse, matrix.F90:195-229), the lattice vectors,
the slanted single-fold minimum image (fundcell.F90:582-623), maptouc and
pbc (simcell.F90:460-484, 634-668), uniform insertion points
(simcell.F90:268-295), LJ map nodes in slanted cells and Hermite
interpolation without the periodic corner wrap (interpolate.F90:675-679),
bit for bit."""
import numpy as np

from oracle import oracle as O


def _cell(g):
    return O.cell(g["ell"], g["angles"], g["eff"], g["ncell"])


def test_cell_matrices_and_lattice_vectors(golden):
    for g in golden["fcell"]:
        c = _cell(g)
        assert bool(c.ortho) == g["ortho"]
        U, S, L = O.cell_matrices(c)
        assert np.array_equal(U.ravel(), np.array(g["U"])), g["angles"]
        assert np.array_equal(S.ravel(), np.array(g["S"])), g["angles"]
        assert np.array_equal(L.ravel(), np.array(g["L"])), g["angles"]
        assert O.cell_volume(c) == g["vol"], g["angles"]
    # the monoclinic / triclinic matrices are not the orthorhombic identity
    assert not golden["fcell"][1]["ortho"] and abs(golden["fcell"][1]["U"][2] + 0.1994) < 1e-3


def test_slanted_minimum_image_maptouc_pbc(golden):
    for g in golden["fcell"]:
        c = _cell(g)
        for pr, sp in zip(g["pairs"], g["sep"]):
            sep = O.cell_minimage(c, pr[:3], pr[3:])
            assert np.array_equal(sep, np.array(sp[:3])), g["angles"]
            d2 = 0.0
            for x in sep:
                d2 = d2 + x * x
            assert d2 == sp[3]
        for p, uc, pb in zip(g["points"], g["uc"], g["pbc"]):
            assert np.array_equal(O.cell_maptouc(c, p), np.array(uc)), g["angles"]
            assert np.array_equal(O.cell_pbc(c, p), np.array(pb)), g["angles"]


def test_uniform_insertion_points(golden):
    for g in golden["fcell"]:
        c = _cell(g)
        assert np.array_equal(O.cell_uniform(c, 5151, len(g["rand"])), np.array(g["rand"])), g["angles"]


def test_lj_nodes_in_slanted_cells(golden):
    for g in golden["ljnodex"]:
        ell, ang, eff, nc = g["cell"]
        c = O.cell(ell, ang, eff, nc)
        fw = dict(xyz=g["xyz"], type=g["type"], A=g["A"], B=g["B"], hicut2=g["hicut2"], locut2=g["locut2"],
                  full_ell=np.array(ell) * np.array(nc), cell=c)
        ok, vals = O.node_at(1, fw, g["nodes"])
        assert list(ok) == g["ok"]
        assert np.array_equal(vals, np.array(g["vals"])), ang
        assert sum(g["ok"]) > 20


def test_hermite_in_slanted_cell_without_wrap(golden):
    gc, gh = golden["compact"], golden["hermitex"]
    ptr, _, _ = O.compact(gc["npnts"], np.array(gc["vals"])[:, 0].astype(np.float32), gc["cutoff"])
    npn = gc["npnts"]
    n = int(np.prod(npn))
    src = O.compact(npn, np.array(gc["vals"])[:, 0].astype(np.float32), gc["cutoff"])[1]
    vals8 = np.array(gc["vals"], dtype=np.float32)[src].T.copy()
    eff = golden["hermite"]["eff"]
    step = [eff[i] / (npn[i] - 1) for i in range(3)]
    ell, ang, eff_c, nc = gh["cell"]
    c = O.cell(ell, ang, eff_c, nc)
    ok, V = O.pothermite(npn, step, ptr, vals8, gh["points"], cell=c)
    assert list(ok) == gh["ok"]
    assert np.array_equal(V[ok], np.array(gh["V"])[ok])
    assert 0 < sum(gh["ok"]) < len(gh["ok"])


def test_host_lattice_vectors_and_replication(golden, tmp_path):
    """The Python host's lattice vectors (fundcell_genlatticevects) equal the
    reference's bit for bit, and a framework is replicated along them
    (simcell.F90:200-216 -> molecule_expand): copy i along axis k is shifted
    by i lvec_k, so even a 90-degree cell carries the reference's
    truncated-pi components."""
    import os
    import shutil
    from musichip import inputs
    for g in golden["fcell"]:
        L = inputs.lattice_vectors(np.array(g["ell"]) * np.array(g["ncell"]), g["angles"])
        assert np.array_equal(L.ravel(), np.array(g["L"]))
    here = os.path.dirname(os.path.abspath(__file__))
    s = inputs.load_system(os.path.join(here, "golden", "test012-pmapgen", "test012.ctr"))
    fw = s["mols"]["Lattice"]
    n = len(fw["xyz"])
    assert list(s["nxyz"]) == [1, 1, 2]
    lv = inputs.lattice_vectors(s["fund_ell"], s["angles"])
    assert lv[2][0] != 0.0  # cos of the truncated-pi right angle
    assert np.array_equal(fw["full_xyz"][n:], fw["xyz"] + lv[2] * 1.0)
