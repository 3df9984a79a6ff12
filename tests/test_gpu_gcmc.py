"""Device-resident GCMC (mh_gcmc_*) against the reference's test007 run and
the CPU oracle: FP64 trajectories identical move for move, energies within
1e-10 relative; FP32 per-move energies within 1e-5 of FP64."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from musichip import hip as H
from musichip import gcmc as G
from musichip import inputs, synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF7 = json.load(open(os.path.join(GOLD, "reference_outputs.json")))["test007"]


def _cell(ctx, L):
    L = np.asarray(L, dtype=np.float64)
    ctx.upload_framework(np.zeros((1, 3)), np.zeros(1, dtype=np.int32), None, L, L)


def _spec007():
    aa = inputs.read_atom_atom(os.path.join(GOLD, "test007-gcmc-CH4", "atom_atom_file"))
    p = aa[("CH4", "CH4")]
    return dict(full_ell=[20.0] * 3, fund_ell=[20.0] * 3, ref=[[0, 0, 0]], stype=[0], sq=[0.0], nst=1,
                lj=([p["A"]], [p["B"]], [p["hicut2"]], [p["locut2"]], [1]), T=300.0,
                moves=[("RINSERT", 1.0), ("RDELETE", 1.0), ("RTRANSLATE", 1.0)], dtrans=0.4)


def test_gpu_reproduces_test007_counts(hip):
    print("case: test007 on the GPU, FP64, seed 22111971", flush=True)
    _cell(hip, [20.0] * 3)
    spec = _spec007()
    # replica 0: the reference run; replicas 1-3: other seeds (independent chains)
    b = G.GcmcBatch(hip, spec, [10000.0] * 4, [22111971, 1, 2, 3], maxN=256)
    for cp in REF7["checkpoints"]:
        b.run(cp["step"] - int(b.state()["nsteps"][0]))
        s = b.state()
        assert [s["acc"][0][0], s["att"][0][0]] == cp["insert"], cp["step"]
        assert [s["acc"][0][1], s["att"][0][1]] == cp["delete"], cp["step"]
        assert [s["acc"][0][2], s["att"][0][2]] == cp["translate"], cp["step"]
        assert s["N"][0] == int(cp["N"]), cp["step"]
    s = b.state()
    assert abs(s["energy"][0] * 4.184 - REF7["final_energy_kJ"]) < 5e-4
    # the final molecules, slot by slot, against the reference's restart file
    import music_gcmc_run as MG
    rp = MG.read_restart_rp(os.path.join(GOLD, "test007-gcmc-CH4", "100.res.1"))["CH4"]
    com, _, _ = b.config(0)
    assert len(com) == len(rp) == 21
    assert np.abs(com - rp[:, 0, :]).max() < 5.01e-6
    tot = b.total_energy()
    assert np.allclose(tot, s["energy"], rtol=1e-10, atol=1e-10)
    # the other seeds against the oracle
    for r, seed in ((1, 1), (2, 2), (3, 3)):
        g = O.Gcmc(spec, 10000.0, seed)
        g.run(20000)
        assert [g.st.att[i] for i in range(3)] == list(s["att"][r][:3])
        assert [g.st.acc[i] for i in range(3)] == list(s["acc"][r][:3])
        assert g.st.N == s["N"][r]


def small_charged_system(hip, grid=0.3):
    """20 A cubic framework (2x2x2 cells of 10 A, 96 atoms) with charges,
    LJ maps for the O and C sites of a CO2-like sorbate and an emap."""
    rng = np.random.default_rng(33)
    xyz = synth.lattice_faces(2, 10.0, rng)
    eps, sig, typ = synth.fw_types(rng, len(xyz), 4)
    q = rng.uniform(-0.6, 0.6, len(xyz))
    q -= q.mean()
    L = np.array([20.0, 20.0, 20.0])
    hip.upload_framework(xyz, typ, q, L, L)
    sorb = synth.co2_like()
    maps = {}
    pm = {}
    for name in ("O", "C"):
        e, s_ = sorb["lj"][name]
        hip.set_lj(*synth.lb_table(eps, sig, e, s_, 9.5))
        pm[name] = hip.build_map(H.MH_MAP_LJ, H.MH_PREC_FP64, grid, 100.0)
    hip.set_ewald(5.0, 7, 1e-10)
    em = hip.build_map(H.MH_MAP_EWALD, H.MH_PREC_FP64, grid, 100.0, probe_q=1.0)
    for a, t in enumerate(sorb["types"]):
        maps[a] = (pm[t], em)
    # sorbate-sorbate LJ table: types 0 = O, 1 = C
    tab = {}
    for i, ti in enumerate(("O", "C")):
        for j, tj in enumerate(("O", "C")):
            e1, s1 = sorb["lj"][ti]
            e2, s2 = sorb["lj"][tj]
            A, B = inputs.lj_AB((s1 + s2) / 2, np.sqrt(e1 * e2))
            tab[i * 2 + j] = (A, B)
    spec = dict(full_ell=L, fund_ell=L, ref=sorb["ref"], stype=[0, 1, 0], sq=sorb["q"], nst=2,
                lj=([tab[t][0] for t in range(4)], [tab[t][1] for t in range(4)], [9.5 ** 2] * 4,
                    [0.5 ** 2] * 4, [1] * 4),
                T=300.0, ewald=dict(kappa=0.30, kmax=6, rc=9.5),
                moves=[("RINSERT", 1.0), ("RDELETE", 1.0), ("RTRANSLATE", 1.0), ("RROTATE", 1.0)],
                dtrans=0.4, drot=0.4)
    return spec, maps, L


def _oracle_for(spec, maps, fug, seed):
    """Oracle chain using the same device maps (downloaded)."""
    mlist = []
    idx = {}
    pm_idx, em_idx = [], []
    for a in range(len(spec["ref"])):
        pm, em = maps[a]
        for m, dst in ((pm, pm_idx), (em, em_idx)):
            if m is None:
                dst.append(-1)
                continue
            if id(m) not in idx:
                ptr, vals = m.get()
                idx[id(m)] = len(mlist)
                mlist.append((m.npnts, m.step, ptr, vals))
            dst.append(idx[id(m)])
    sp = dict(spec, maps=mlist, site_pmap=pm_idx, site_emap=em_idx)
    return O.Gcmc(sp, fug, seed, trace=4000)


# kernel variants: (workgroup size, maxN) -> molecules in LDS with the k
# vectors in registers (tiles of 8, 16, 4) or S(k) in HBM (1024 lanes), and
# molecules in HBM (maxN too large for LDS)
VARIANTS = [(256, 512), (128, 512), (512, 512), (1024, 512), (256, 20000)]


def test_gpu_fp64_trajectory_matches_oracle_with_maps_and_ewald(hip):
    spec, maps, L = small_charged_system(hip)
    fugs = [300.0, 3000.0]
    seeds = [3001, 3002]
    ref = []
    for r in range(2):
        g = _oracle_for(spec, maps, fugs[r], seeds[r])
        g.run(3000)
        ref.append((g.trace_array(3000), g.st.N, g.st.energy))
    for threads, maxN in VARIANTS:
        print("case: CO2-like sorbate, maps + Ewald, FP64 trajectory parity, threads %d maxN %d"
              % (threads, maxN), flush=True)
        b = G.GcmcBatch(hip, spec, fugs, seeds, maxN=maxN, maps=maps, threads=threads)
        b.set_trace(3000)
        b.run(3000)
        s = b.state()
        for r in range(2):
            to, No, eo = ref[r]
            tg = b.trace(r)
            assert np.array_equal(tg[:, :3].astype(int), to[:, :3].astype(int)), (threads, maxN, r)
            scale = np.maximum(np.abs(to[:, 3]), 1.0)
            assert np.max(np.abs(tg[:, 3] - to[:, 3]) / scale) < 1e-10
            assert No == s["N"][r] and No > 5
            assert abs(eo - s["energy"][r]) < 1e-9 * max(1.0, abs(eo))
        tot = b.total_energy()
        assert np.allclose(tot, s["energy"], rtol=1e-10, atol=1e-9)
        b.close()


def test_gpu_fp32_trial_energies_vs_fp64(hip):
    spec, maps, L = small_charged_system(hip)
    b64 = G.GcmcBatch(hip, spec, [3000.0], [77], maxN=512, maps=maps)
    b64.run(4000)
    com, ang, _ = b64.config(0)
    assert len(com) >= 10
    b32 = G.GcmcBatch(hip, spec, [3000.0], [77], maxN=512, maps=maps, precision=H.MH_PREC_FP32)
    b32.set_config(0, com, ang)
    rng = np.random.default_rng(9)
    tc = rng.uniform(0, 1, (400, 3)) * L
    ta = np.stack([rng.uniform(-np.pi, np.pi, 400), np.arccos(rng.uniform(-1, 1, 400)),
                   rng.uniform(-np.pi, np.pi, 400)], 1)
    ok64, e64 = b64.trial_energy(0, tc, ta)
    ok32, e32 = b32.trial_energy(0, tc, ta)
    both = ok64 & ok32
    assert both.sum() > 100
    rel = np.abs(e32[both] - e64[both]) / np.maximum(np.abs(e64[both]), 1.0)
    print("FP32 trial energy max rel err (floor 1 kcal/mol) %.3e" % rel.max(), flush=True)
    assert rel.max() < 1e-5
    assert np.array_equal(ok32 | ok64, ok64)  # no extra failures in FP32
    # oracle insertion energies through the same configuration
    g = _oracle_for(spec, maps, 3000.0, 77)
    g.set_config(com, ang)
    assert abs(g.total_energy() - b64.total_energy()[0]) < 1e-9 * max(1.0, abs(g.total_energy()))


def test_gpu_checkpoint_resume_is_exact(hip):
    _cell(hip, [20.0] * 3)
    spec = _spec007()
    a = G.GcmcBatch(hip, spec, [10000.0], [4242], maxN=256)
    a.run(2000)
    st = a.state()
    com, ang, _ = a.config(0)
    a.run(2000)
    full = a.state()
    b = G.GcmcBatch(hip, spec, [10000.0], [4242], maxN=256)
    b.set_config(0, com, ang)
    b.set_rng(0, st["ix"][0], st["iy"][0])
    b.run(2000)
    res = b.state()
    c1, _, _ = a.config(0)
    c2, _, _ = b.config(0)
    assert res["N"][0] == full["N"][0]
    assert np.array_equal(c1, c2)
    assert np.array_equal(res["acc"][0] + st["acc"][0], full["acc"][0])


def test_gpu_graph_cycles_equal_eager_runs(hip):
    """mh_gcmc_run_cycles replays one captured launch per cycle: after
    graph replays interleaved with eager launches and state reads, the batch
    is bit for bit the one advanced by eager launches only (counts, N,
    energies, RNG, coordinates), for the charged 3-site system with maps and
    Ewald in FP32 and FP64."""
    print("case: GCMC hipGraph cycles vs eager launches", flush=True)
    for prec in (H.MH_PREC_FP64, H.MH_PREC_FP32):
        spec, maps, _ = small_charged_system(hip)
        fug = [1000.0, 5000.0, 20000.0, 50000.0]
        seeds = [611, 612, 613, 614]
        a = G.GcmcBatch(hip, spec, fug, seeds, maxN=128, maps=maps, precision=prec)
        b = G.GcmcBatch(hip, spec, fug, seeds, maxN=128, maps=maps, precision=prec)
        for _ in range(7):
            a.run(150)
            a.state()
        b.run_cycles(150, 2)
        b.state()
        b.run(150)
        b.run_cycles(150, 3)
        b.state()
        b.run_cycles(150, 1)
        sa, sb = a.state(), b.state()
        for k in ("N", "energy", "att", "acc", "sumN", "nsteps", "ix", "iy"):
            assert np.array_equal(sa[k], sb[k]), (prec, k)
        assert sa["acc"].sum() > 50
        for r in range(len(seeds)):
            assert np.array_equal(a.config(r)[0], b.config(r)[0])
        a.close()
        b.close()


def _coul_spec(spec, hicut):
    """The charged system's sorbate with plain Coulomb (coulomb_snglint)
    between sorbate sites instead of Ewald; the framework keeps its maps."""
    sp = dict(spec)
    sp.pop("ewald", None)
    sp["coul"] = dict(on=[1, 1, 1, 1], hicut=hicut, locut=0.5, shield=1.0)
    return sp


@pytest.mark.parametrize("hicut", [9.5, 0.0])
def test_gpu_plain_coulomb_trajectory_matches_oracle(hip, hicut):
    """Plain Coulomb sorbate-sorbate electrostatics (shifted at hicut, or
    every minimum-image pair): FP64 trajectories move for move against the
    oracle, running energy = recompute; FP32 trial energies within 1e-5."""
    print("case: plain Coulomb GCMC vs oracle, hicut %g" % hicut, flush=True)
    spec, maps, L = small_charged_system(hip)
    sp = _coul_spec(spec, hicut)
    fugs = [1000.0, 10000.0]
    seeds = [5001, 5002]
    b = G.GcmcBatch(hip, sp, fugs, seeds, maxN=512, maps=maps)
    b.set_trace(3000)
    b.run(3000)
    s = b.state()
    for r in range(2):
        g = _oracle_for(sp, maps, fugs[r], seeds[r])
        g.run(3000)
        to, tg = g.trace_array(3000), b.trace(r)
        assert np.array_equal(tg[:, :3].astype(int), to[:, :3].astype(int)), r
        scale = np.maximum(np.abs(to[:, 3]), 1.0)
        assert np.max(np.abs(tg[:, 3] - to[:, 3]) / scale) < 1e-10
        assert g.st.N == s["N"][r] and g.st.N > 5
        assert abs(g.st.energy - s["energy"][r]) < 1e-9 * max(1.0, abs(g.st.energy))
    assert np.allclose(b.total_energy(), s["energy"], rtol=1e-10, atol=1e-9)
    com, ang, _ = b.config(1)
    b32 = G.GcmcBatch(hip, sp, [10000.0], [7], maxN=512, maps=maps, precision=H.MH_PREC_FP32)
    b32.set_config(0, com, ang)
    rng = np.random.default_rng(19)
    tc = rng.uniform(0, 1, (400, 3)) * L
    ta = np.stack([rng.uniform(-np.pi, np.pi, 400), np.arccos(rng.uniform(-1, 1, 400)),
                   rng.uniform(-np.pi, np.pi, 400)], 1)
    ok64, e64 = b.trial_energy(1, tc, ta)
    ok32, e32 = b32.trial_energy(0, tc, ta)
    both = ok64 & ok32
    assert both.sum() > 100
    rel = np.abs(e32[both] - e64[both]) / np.maximum(np.abs(e64[both]), 1.0)
    print("plain Coulomb FP32 trial energy max rel err %.3e" % rel.max(), flush=True)
    assert rel.max() < 1e-5
    b.close()
    b32.close()


def test_gpu_adjustable_displacements_match_oracle(hip):
    """Adjustable translate / rotate displacements (rigidmoves_adjustdelta*,
    after every attempt): FP64 trajectories move for move against the
    oracle and the same final displacements, bit for bit; a checkpoint of
    the displacements restores them."""
    print("case: adjustable displacements vs oracle", flush=True)
    spec, maps, L = small_charged_system(hip)
    sp = dict(spec)
    sp.update(adjust_trans=1, adjust_rot=1, dtrans=1.5, drot=1.5)
    fugs = [3000.0, 30000.0]
    seeds = [6101, 6102]
    b = G.GcmcBatch(hip, sp, fugs, seeds, maxN=512, maps=maps)
    b.set_trace(3000)
    b.run(3000)
    d = b.deltas()
    for r in range(2):
        g = _oracle_for(sp, maps, fugs[r], seeds[r])
        g.run(3000)
        to, tg = g.trace_array(3000), b.trace(r)
        assert np.array_equal(tg[:, :3].astype(int), to[:, :3].astype(int)), r
        assert list(d[r]) == list(g.st.delta), (r, d[r], list(g.st.delta))
        assert d[r][0] != 1.5 and d[r][1] != 1.5
    b.set_deltas(0, 0.25, 0.125)
    assert list(b.deltas()[0]) == [0.25, 0.125]
    b.close()
