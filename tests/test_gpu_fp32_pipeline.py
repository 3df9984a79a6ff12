"""The FP32 pipeline end to end against FP64: maps built in FP32 pair
arithmetic (FP64 accumulation) feeding the FP32 GCMC kernel, against FP64
maps feeding the FP64 kernel, which equals the oracle (checked here on a
sample of trial energies through oracle chains that use the same maps).

Bounds (floor 1 kcal/mol: below it absolute error in kcal/mol is what the
Boltzmann factor at 300 K sees):
* trial insertion energies of the 3-site CO2-like sorbate in the 30 A
  methane-isotherm framework: |E32 - E64| <= 1e-5 max(|E64|, 1);
* map derivative columns of order n (x, y, z: 1; xy, xz, yz: 2; xyz: 3),
  judged by what they add to a Hermite interpolant on a grid of spacing h:
  |d32 - d64| h^n <= 1e-5 max(|V|, 1) on every tabulated node, and the
  interpolated energy at random points within 1e-5 (floor 1);
* the FP32 Ewald real-space sum with its cutoff 4.32 / kappa (erfc below
  1e-9) in a 40 A cell, where the cutoff (17.3 A) is shorter than L / 2:
  against the oracle's FP64 sum over every atom, within 5e-5 (floor 1)."""
import numpy as np
import pytest

from oracle import oracle as O
from musichip import hip as H
from musichip import gcmc as G
from musichip import synth, workloads as W

pytestmark = pytest.mark.gpu
FLOOR = 1.0


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), FLOOR)


@pytest.fixture(scope="module")
def co2_maps(hip):
    """The batched-isotherm system with its maps built twice: FP32 and FP64."""
    s32 = W.co2_isotherm_system(hip, map_precision=H.MH_PREC_FP32, npoints=4)
    s64 = W.co2_isotherm_system(hip, map_precision=H.MH_PREC_FP64, npoints=4)
    yield s32, s64
    for m in s32["map_objs"] + s64["map_objs"]:
        m.close()


def test_fp32_maps_and_kernel_trial_energies_vs_fp64(hip, co2_maps):
    s32, s64 = co2_maps
    print("case: FP32 maps + FP32 GCMC vs FP64 maps + FP64 GCMC, CO2 in the 30 A framework", flush=True)
    fug = [1.0e3, 1.0e4, 1.0e5, 1.0e6]
    b64 = G.GcmcBatch(hip, s64["spec"], fug, [3000, 3001, 3002, 3003], maxN=512, maps=s64["maps"],
                      precision=H.MH_PREC_FP64)
    b64.run(20000)
    N = b64.state()["N"]
    print("loadings", list(N), flush=True)
    assert N.max() >= 40
    b32 = G.GcmcBatch(hip, s32["spec"], fug, [3000, 3001, 3002, 3003], maxN=512, maps=s32["maps"],
                      precision=H.MH_PREC_FP32)
    rng = np.random.default_rng(17)
    L = np.asarray(s64["spec"]["full_ell"])
    worst, n = 0.0, 0
    cfgs = []
    for r in range(4):
        com, ang, _ = b64.config(r)
        cfgs.append((com, ang))
        b32.set_config(r, com, ang)
        # about one random placement in ten lands on tabulated corners
        nt = 4096
        tc = rng.uniform(0, 1, (nt, 3)) * L
        ta = np.stack([rng.uniform(-np.pi, np.pi, nt), np.arccos(rng.uniform(-1, 1, nt)),
                       rng.uniform(-np.pi, np.pi, nt)], 1)
        ok64, e64 = b64.trial_energy(r, tc, ta)
        ok32, e32 = b32.trial_energy(r, tc, ta)
        # the same failures (untabulated corners, overlaps) on both sides
        assert np.count_nonzero(ok64 != ok32) <= 2
        both = ok64 & ok32
        n += int(both.sum())
        worst = max(worst, float(_rel(e32[both], e64[both]).max()))
    print("FP32 maps + kernel vs FP64: %d trial energies, max rel err %.3e (floor 1 kcal/mol)" % (n, worst),
          flush=True)
    assert n > 500
    assert worst <= 1e-5
    # the FP64 side is the oracle's: insertion energy = E(N + 1) - E(N) of an
    # oracle chain on the same (FP64) maps, for a sample of trials
    from test_gpu_gcmc import _oracle_for
    com, ang = cfgs[1]
    g = _oracle_for(s64["spec"], s64["maps"], fug[1], 5)
    g.set_config(com, ang)
    e0 = g.total_energy()
    # about one random placement in ten is tabulated: 160 placements
    tc = rng.uniform(0, 1, (160, 3)) * L
    ta = np.stack([rng.uniform(-np.pi, np.pi, 160), np.arccos(rng.uniform(-1, 1, 160)),
                   rng.uniform(-np.pi, np.pi, 160)], 1)
    ok64, e64 = b64.trial_energy(1, tc, ta)
    checked = 0
    for i in range(160):
        g.set_config(np.vstack([com, tc[i:i + 1]]), np.vstack([ang, ta[i:i + 1]]))
        e1 = g.total_energy()
        if not ok64[i]:
            assert not np.isfinite(e1) or e1 > 1e9
            continue
        assert e1 - e0 == pytest.approx(e64[i], rel=1e-9, abs=1e-9)
        checked += 1
    assert checked >= 8
    b32.close()
    b64.close()


def test_fp32_map_derivatives_and_interpolation_vs_fp64(hip):
    """Derivative columns of FP32 LJ and Ewald maps (charged 20 A box) judged
    by their contribution to the Hermite interpolant, and the interpolated
    energies themselves."""
    rng = np.random.default_rng(2002)
    L = np.array([20.0, 20.0, 20.0])
    n = 300
    xyz = rng.uniform(0, 1, (n, 3)) * L
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    typ = rng.integers(0, synth.NTYPES_FW, n).astype(np.int32)
    eps, sig = rng.uniform(20, 80, synth.NTYPES_FW), rng.uniform(2.5, 3.8, synth.NTYPES_FW)
    hip.upload_framework(xyz, typ, q, L, L)
    hip.set_lj(*synth.lb_table(eps, sig, 27.0, 2.80, 9.5))
    hip.set_ewald(5.0, 9, 1e-10)
    h = 0.25
    order = np.array([0, 1, 1, 1, 2, 2, 2, 3])
    for kind, name in ((H.MH_MAP_LJ, "LJ"), (H.MH_MAP_EWALD, "Ewald")):
        print("case: FP32 %s map derivative columns" % name, flush=True)
        m64 = hip.build_map(kind, H.MH_PREC_FP64, h, 100.0, probe_q=0.7, keep_full=True)
        m32 = hip.build_map(kind, H.MH_PREC_FP32, h, 100.0, probe_q=0.7, keep_full=True)
        _, a64 = m64.get_full()
        _, a32 = m32.get_full()
        live = a64[:, 0] < 100.0 / 4.184
        scale = np.maximum(np.abs(a64[live, 0]), FLOOR)
        for c in range(8):
            e = np.abs(a32[live, c] - a64[live, c]) * h ** order[c] / scale
            print("  column %d (order %d): max |d32 - d64| h^n / max(|V|, 1) = %.3e" % (c, order[c], e.max()),
                  flush=True)
            assert e.max() <= 1e-5, (name, c)
        pts = rng.uniform(0, 1, (20000, 3)) * L
        ok64, e64 = m64.interp(pts, 1.0)
        ok32, e32 = m32.interp(pts, 1.0)
        assert np.array_equal(ok64, ok32)
        ie = _rel(e32[ok64], e64[ok64]).max()
        print("  interpolated energies: max rel err %.3e over %d points" % (ie, ok64.sum()), flush=True)
        assert ok64.sum() > 5000 and ie <= 1e-5
        m64.close()
        m32.close()


def test_fp32_ewald_real_space_cutoff_shorter_than_half_cell(hip):
    """40 A cell, kappa 0.25 / A: the FP32 kernel's real-space cutoff
    4.32 / kappa = 17.28 A < 20 A cuts pairs the reference sums; the
    omitted terms are below erfc(4.32) = 1e-9 of q / r.  Against the
    oracle's FP64 nodes over every atom (first z-planes of the grid)."""
    print("case: FP32 Ewald map, 40 A cell, rc 17.28 A < L/2", flush=True)
    rng = np.random.default_rng(4040)
    L = np.array([40.0, 40.0, 40.0])
    n = 400
    xyz = rng.uniform(0, 1, (n, 3)) * L
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    hip.upload_framework(xyz, np.zeros(n, dtype=np.int32), q, L, L)
    kp = 0.25 * 40.0
    nk = hip.set_ewald(kp, 7, 1e-10)
    assert 4.32 / 0.25 < 20.0
    m = hip.build_map(H.MH_MAP_EWALD, H.MH_PREC_FP32, 0.5, 100.0, probe_q=1.0, keep_full=True)
    _, acc = m.get_full()
    npn = (L / 0.5).astype(int) + 1
    step = L / (npn - 1)
    sf = O.ewald_sfactor(xyz, q, L, kp, 7, 1e-10)
    assert sf["nk"] == nk
    nodes = int(npn[0] * npn[1] * 3)
    ref = O.build_nodes(2, dict(xyz=xyz, q=q, charge1=1.0, kappa=sf["kappa"], k=sf["k"], s=sf["s"],
                                si=sf["si"], full_ell=L), npn, step, 0, nodes)
    live = ref[:, 0] < 100.0 / 4.184
    e = _rel(acc[:nodes][live, 0], ref[live, 0]).max()
    print("FP32 Ewald (cutoff 17.28 A) vs FP64 all atoms: max rel err %.3e over %d nodes" % (e, live.sum()),
          flush=True)
    assert live.sum() > 5000 and e <= 5e-5
    m.close()
