"""GPU map build (libmusic_hip) against the CPU oracle and the reference's
golden values.  FP64 mode: node values within 1e-10 relative, integer
grid/compaction bit-exact.  FP32 mode: error against FP64 stated and bounded."""
import hashlib
import os
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from musichip import hip as H
from musichip import inputs, synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL64 = 1e-10
# FP32 map values: relative error with the denominator floored at 1 kcal/mol
# (|V| below 1 kcal/mol is judged by absolute error; the Boltzmann factor
# of a map energy at 300 K depends on absolute error in kcal/mol)
VFLOOR = 1.0
TOL32 = 1e-5
# Ewald real-space terms are individually large (332 q erfc(kr)/r) and
# cancel; their FP32 rounding leaves a few 1e-5 kcal/mol absolute error
TOL32_EWALD = 5e-5


def fp32_err(acc, ref, live, what):
    d = np.abs(acc[live, 0] - ref[live, 0])
    rel = d / np.maximum(np.abs(ref[live, 0]), VFLOOR)
    print("%s: FP32 V max rel err (|V| floor 1 kcal/mol) %.3e, max abs err %.3e kcal/mol, "
          "median rel err %.3e" % (what, rel.max(), d.max(), np.median(d / np.maximum(np.abs(ref[live, 0]), 1e-12))),
          flush=True)
    return rel.max()


def rel_close(got, ref, tol, what):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.max(np.abs(ref), axis=0, keepdims=True)
    err = np.abs(got - ref) / (np.abs(ref) + 1e-6 * scale + 1e-300)
    worst = float(np.max(err)) if err.size else 0.0
    print("%s: max scaled rel err %.3e" % (what, worst), flush=True)
    assert worst <= tol, what


def _system(name):
    return inputs.load_system(os.path.join(GOLD, name))


def _grid(s, job):
    npn = (s["fund_ell"] / job["gridspace"]).astype(int) + 1
    return npn, s["fund_ell"] / (npn - 1)


@pytest.fixture(scope="module")
def t012():
    s = _system("test012-pmapgen/test012.ctr")
    fw = s["mols"]["Lattice"]
    A, B, h2, l2, on = inputs.lj_table_for_probe(s, int(s["mols"]["Probe"]["type"][0]))
    job = s["ctrl"]["mapmaker"][0]
    npn, step = _grid(s, job)
    ofw = dict(xyz=fw["full_xyz"], type=fw["full_type"], A=A, B=B, hicut2=h2, locut2=l2,
               full_ell=s["full_ell"])
    ref = O.build_nodes(1, ofw, npn, step)
    return dict(s=s, fw=fw, lj=(A, B, h2, l2, on), job=job, npn=npn, step=step, ref=ref)


def test_lj_map_fp64_matches_oracle_and_goldens(hip, t012):
    s, fw, job = t012["s"], t012["fw"], t012["job"]
    print("case: test012 LJ map FP64", flush=True)
    hip.upload_framework(fw["full_xyz"], fw["full_type"], fw["full_q"], s["fund_ell"], s["full_ell"])
    hip.set_lj(*t012["lj"])
    m = hip.build_map(H.MH_MAP_LJ, H.MH_PREC_FP64, job["gridspace"], job["hipotcut"], keep_full=True)
    assert list(m.npnts) == [50, 50, 30]
    assert list(m.counts) == [28904, 20544, 49448, 29627]
    f32, f64 = m.get_full()
    rel_close(f64, t012["ref"], TOL64, "test012 LJ nodes FP64")
    assert np.array_equal(f32, t012["ref"].astype(np.float32)) or \
        np.max(np.abs(f32.view(np.int32).astype(np.int64) - t012["ref"].astype(np.float32).view(np.int32))) <= 1
    ptr_o, src_o, counts_o = O.compact(t012["npn"], t012["ref"][:, 0].astype(np.float32),
                                       job["hipotcut"] / 4.184)
    ptr, vals = m.get()
    assert np.array_equal(ptr, ptr_o)
    np.testing.assert_array_equal(vals, f32[src_o].T)
    m.close()


def test_lj_map_file_bytes_match_oracle_writer(hip, t012):
    s, fw, job = t012["s"], t012["fw"], t012["job"]
    hip.upload_framework(fw["full_xyz"], fw["full_type"], fw["full_q"], s["fund_ell"], s["full_ell"])
    hip.set_lj(*t012["lj"])
    m = hip.build_map(H.MH_MAP_LJ, H.MH_PREC_FP64, job["gridspace"], job["hipotcut"])
    ptr, vals = m.get()
    zat = [dict(name="Latom", symbol="L", charge=1.0, A12=t012["lj"][0][1], B6=t012["lj"][1][1],
                rcuthi=5.9, rcutlo=0.1)]
    with tempfile.TemporaryDirectory() as td:
        a, b = os.path.join(td, "gpu.map"), os.path.join(td, "ora.map")
        m.write(a, zat, job["hipotcut"])
        O.map_write(b, zat, job["hipotcut"], m.npnts, ptr, vals)
        da, db = open(a, "rb").read(), open(b, "rb").read()
        assert hashlib.sha256(da).hexdigest() == hashlib.sha256(db).hexdigest()
        m2 = hip.load_map(a, s["fund_ell"])
        p2, v2 = m2.get()
        assert np.array_equal(p2, ptr) and np.array_equal(v2, vals)


def test_lj_map_fp32_error_vs_fp64(hip, t012):
    s, fw, job = t012["s"], t012["fw"], t012["job"]
    hip.upload_framework(fw["full_xyz"], fw["full_type"], fw["full_q"], s["fund_ell"], s["full_ell"])
    hip.set_lj(*t012["lj"])
    m = hip.build_map(H.MH_MAP_LJ, H.MH_PREC_FP32, job["gridspace"], job["hipotcut"], keep_full=True)
    f32, acc = m.get_full()
    ref = t012["ref"]
    live = ref[:, 0] < job["hipotcut"] / 4.184  # tabulated region
    assert fp32_err(acc, ref, live, "test012 LJ") < TOL32
    assert list(m.counts) == [28904, 20544, 49448, 29627]


@pytest.fixture(scope="module")
def t013():
    s = _system("test013-emapgen/test013.ctr")
    fw = s["mols"]["Lattice"]
    e = s["ss"][("Probe", "Lattice", "COUL")]["ewald"]
    sf = O.ewald_sfactor(fw["full_xyz"], fw["full_q"], s["full_ell"], e["kappa"], e["kmax"], e["locut"])
    job = s["ctrl"]["mapmaker"][0]
    npn, step = _grid(s, job)
    d = dict(xyz=fw["full_xyz"], q=fw["full_q"], charge1=1.0, kappa=sf["kappa"], k=sf["k"],
             s=sf["s"], si=sf["si"], full_ell=s["full_ell"])
    ref = O.build_nodes(2, d, npn, step)
    return dict(s=s, fw=fw, e=e, sf=sf, job=job, npn=npn, step=step, ref=ref)


def test_ewald_map_fp64_matches_oracle(hip, t013):
    s, fw, e, job = t013["s"], t013["fw"], t013["e"], t013["job"]
    print("case: test013 Ewald map FP64", flush=True)
    hip.upload_framework(fw["full_xyz"], fw["full_type"], fw["full_q"], s["fund_ell"], s["full_ell"])
    nk = hip.set_ewald(e["kappa"], e["kmax"], e["locut"])
    assert nk == 158
    kk = hip.ewald_get(nk)
    np.testing.assert_allclose(kk["k"], t013["sf"]["k"], rtol=0, atol=0)
    np.testing.assert_allclose(kk["s"], t013["sf"]["s"], rtol=1e-12, atol=1e-15)
    m = hip.build_map(H.MH_MAP_EWALD, H.MH_PREC_FP64, job["gridspace"], job["hipotcut"],
                      probe_q=1.0, keep_full=True)
    assert list(m.counts) == [12240, 1360, 13600, 10200]
    f32, f64 = m.get_full()
    rel_close(f64, t013["ref"], TOL64, "test013 Ewald nodes FP64")
    ptr_o, _, _ = O.compact(t013["npn"], t013["ref"][:, 0].astype(np.float32), job["hipotcut"] / 4.184)
    ptr, _ = m.get()
    assert np.array_equal(ptr, ptr_o)


def test_ewald_map_fp32_error_vs_fp64(hip, t013):
    s, fw, e, job = t013["s"], t013["fw"], t013["e"], t013["job"]
    hip.upload_framework(fw["full_xyz"], fw["full_type"], fw["full_q"], s["fund_ell"], s["full_ell"])
    hip.set_ewald(e["kappa"], e["kmax"], e["locut"])
    m = hip.build_map(H.MH_MAP_EWALD, H.MH_PREC_FP32, job["gridspace"], job["hipotcut"],
                      probe_q=1.0, keep_full=True)
    f32, acc = m.get_full()
    ref = t013["ref"]
    live = ref[:, 0] < job["hipotcut"] / 4.184
    assert fp32_err(acc, ref, live, "test013 Ewald") < TOL32_EWALD


def test_buck_map_fp64_matches_oracle(hip):
    rng = np.random.default_rng(5)
    L = np.array([14.0, 15.0, 13.0])
    xyz = rng.uniform(0, 1, (60, 3)) * L
    typ = rng.integers(0, 3, 60).astype(np.int32)
    A = np.array([20000.0, 30000.0, 25000.0])
    rho = np.array([0.30, 0.28, 0.33])
    Cc = np.array([300.0, 250.0, 400.0])
    D = np.array([0.0, 50.0, 10.0])
    h2 = np.full(3, 6.5 ** 2)
    l2 = np.full(3, 0.5 ** 2)
    hip.upload_framework(xyz, typ, None, L, L)
    hip.set_buck(A, rho, Cc, D, h2, l2)
    m = hip.build_map(H.MH_MAP_BUCK, H.MH_PREC_FP64, 0.5, 100.0, keep_full=True)
    npn = (L / 0.5).astype(int) + 1
    step = L / (npn - 1)
    ref = O.build_nodes(3, dict(xyz=xyz, type=typ, bA=A, brho=rho, bC=Cc, bD=D, hicut2=h2, locut2=l2,
                                full_ell=L), npn, step)
    _, f64 = m.get_full()
    rel_close(f64, ref, 1e-12, "Buckingham nodes")


def test_interp_matches_oracle_hermite(hip, t012):
    s, fw, job = t012["s"], t012["fw"], t012["job"]
    hip.upload_framework(fw["full_xyz"], fw["full_type"], fw["full_q"], s["fund_ell"], s["full_ell"])
    hip.set_lj(*t012["lj"])
    m = hip.build_map(H.MH_MAP_LJ, H.MH_PREC_FP64, job["gridspace"], job["hipotcut"])
    ptr, vals = m.get()
    rng = np.random.default_rng(99)
    pts = rng.uniform(-0.3, 1.3, (4000, 3)) * s["full_ell"]
    ok, e = m.interp(pts, 1.0)
    ok_o, e_o = O.pothermite(m.npnts, m.step, ptr, vals, pts, fund_ell=s["fund_ell"])
    assert np.array_equal(ok, ok_o)
    assert ok.sum() > 1000
    np.testing.assert_array_equal(e[ok], e_o[ok])


def test_small_charged_config_fp64_and_fp32(hip):
    # a reduced version of the charged-framework configuration
    rng = np.random.default_rng(2002)
    L = np.array([20.0, 20.0, 20.0])
    xyz = rng.uniform(0, 1, (300, 3)) * L
    q = rng.uniform(-1, 1, 300)
    q -= q.mean()
    typ = rng.integers(0, synth.NTYPES_FW, 300).astype(np.int32)
    eps, sig = rng.uniform(20, 80, synth.NTYPES_FW), rng.uniform(2.5, 3.8, synth.NTYPES_FW)
    hip.upload_framework(xyz, typ, q, L, L)
    nk = hip.set_ewald(5.0, 9, 1e-10)
    sf = O.ewald_sfactor(xyz, q, L, 5.0, 9, 1e-10)
    assert nk == sf["nk"]
    m = hip.build_map(H.MH_MAP_EWALD, H.MH_PREC_FP64, 0.5, 100.0, probe_q=0.7, keep_full=True)
    npn = (L / 0.5).astype(int) + 1
    step = L / (npn - 1)
    ref = O.build_nodes(2, dict(xyz=xyz, q=q, charge1=0.7, kappa=sf["kappa"], k=sf["k"], s=sf["s"],
                                si=sf["si"], full_ell=L), npn, step)
    _, f64 = m.get_full()
    rel_close(f64, ref, TOL64, "charged box Ewald FP64")
    A, B, h2, l2 = synth.lb_table(eps, sig, 27.0, 2.80, 9.5)
    hip.set_lj(A, B, h2, l2)
    m2 = hip.build_map(H.MH_MAP_LJ, H.MH_PREC_FP64, 0.5, 100.0, keep_full=True)
    ref2 = O.build_nodes(1, dict(xyz=xyz, type=typ, A=A, B=B, hicut2=h2, locut2=l2, full_ell=L), npn, step)
    _, g64 = m2.get_full()
    rel_close(g64, ref2, TOL64, "charged box LJ FP64")
    m3 = hip.build_map(H.MH_MAP_LJ, H.MH_PREC_FP32, 0.5, 100.0, keep_full=True)
    _, g32 = m3.get_full()
    live = ref2[:, 0] < 100.0 / 4.184
    assert fp32_err(g32, ref2, live, "charged box LJ") < TOL32
    m4 = hip.build_map(H.MH_MAP_EWALD, H.MH_PREC_FP32, 0.5, 100.0, probe_q=0.7, keep_full=True)
    _, e32 = m4.get_full()
    live = ref[:, 0] < 100.0 / 4.184
    assert fp32_err(e32, ref, live, "charged box Ewald") < TOL32_EWALD


def _fd_columns(V, npn, h):
    """Central differences of node values V (x fastest) -> the 7 derivative
    columns at interior nodes, order x, y, z, xy, xz, yz, xyz."""
    v = V.reshape(npn[2], npn[1], npn[0]).transpose(2, 1, 0)  # [x][y][z]
    c = (slice(1, -1),) * 3

    def sh(dx, dy, dz):
        return v[1 + dx:npn[0] - 1 + dx, 1 + dy:npn[1] - 1 + dy, 1 + dz:npn[2] - 1 + dz]
    out = [(sh(1, 0, 0) - sh(-1, 0, 0)) / (2 * h[0]), (sh(0, 1, 0) - sh(0, -1, 0)) / (2 * h[1]),
           (sh(0, 0, 1) - sh(0, 0, -1)) / (2 * h[2])]
    for a, b in ((0, 1), (0, 2), (1, 2)):
        def e(sa, sb):
            d = [0, 0, 0]
            d[a], d[b] = sa, sb
            return sh(*d)
        out.append((e(1, 1) - e(1, -1) - e(-1, 1) + e(-1, -1)) / (4 * h[a] * h[b]))
    t = 0
    for sx in (1, -1):
        for sy in (1, -1):
            for sz in (1, -1):
                t = t + sx * sy * sz * sh(sx, sy, sz)
    out.append(t / (8 * h[0] * h[1] * h[2]))
    return out, c


def test_ewald_physics_mode_derivatives_are_those_of_the_potential(hip):
    """A framework that is not centrosymmetric about the origin: in physics
    mode (MH_EWALD_PHYSICS) every derivative column equals the finite
    differences of the map's own potential; in parity mode (the reference's
    signs, ewald.F90:987-1002) the potential belongs to the inverted
    framework, so some columns disagree with the differences of V; and both
    modes give the same real-space part, the parity mode staying
    bit-identical to the oracle."""
    rng = np.random.default_rng(808)
    L = np.array([12.0, 12.0, 12.0])
    n = 24
    xyz = rng.uniform(0, 1, (n, 3)) * L
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    hip.upload_framework(xyz, np.zeros(n, dtype=np.int32), q, L, L)
    h = 0.1
    npn = (L / h).astype(int) + 1
    step = L / (npn - 1)
    errs = {}
    maps = {}
    for mode in (H.MH_EWALD_PHYSICS, H.MH_EWALD_PARITY):
        print("case: Ewald map sign mode %d, finite differences" % mode, flush=True)
        hip.set_ewald(4.0, 7, 1e-12, sign_mode=mode)
        m = hip.build_map(H.MH_MAP_EWALD, H.MH_PREC_FP64, h, 1.0e6, probe_q=1.0, keep_full=True)
        _, f64 = m.get_full()
        maps[mode] = f64
        fd, _ = _fd_columns(f64[:, 0], npn, step)
        # nodes at least 2.5 A from every atom (smooth region)
        g = np.stack(np.meshgrid(*[np.arange(k) * s for k, s in zip(npn, step)], indexing="ij"), -1)[1:-1, 1:-1, 1:-1]
        dmin = np.full(g.shape[:3], np.inf)
        for a in xyz:
            d = g - a
            d -= L * np.round(d / L)
            dmin = np.minimum(dmin, np.sqrt((d * d).sum(-1)))
        far = dmin > 2.5
        cols = f64[:, 1:].reshape(npn[2], npn[1], npn[0], 7).transpose(2, 1, 0, 3)[1:-1, 1:-1, 1:-1]
        e = []
        for c in range(7):
            a, b = cols[..., c][far], fd[c][far]
            e.append(float(np.abs(a - b).max() / np.abs(b).max()))
        errs[mode] = e
        print("  max |analytic - FD| / max |FD| per column:", ["%.2e" % x for x in e], flush=True)
        m.close()
    assert max(errs[H.MH_EWALD_PHYSICS]) < 1e-2
    assert max(errs[H.MH_EWALD_PARITY]) > 0.1
    # parity mode = the oracle (reference arithmetic, pinned to the flang build)
    sf = O.ewald_sfactor(xyz, q, L, 4.0, 7, 1e-12)
    ref = O.build_nodes(2, dict(xyz=xyz, q=q, charge1=1.0, kappa=sf["kappa"], k=sf["k"], s=sf["s"],
                                si=sf["si"], full_ell=L), npn, step, 0, 20000)
    rel_close(maps[H.MH_EWALD_PARITY][:20000], ref, TOL64, "parity-mode Ewald nodes")
