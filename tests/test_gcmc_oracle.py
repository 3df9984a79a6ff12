"""The GCMC move-loop oracle against the reference: the shipped test007
run (pure-fluid methane GCMC) is reproduced move for move, and the Ewald
extension is checked against closed forms."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from musichip import inputs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = json.load(open(os.path.join(GOLD, "reference_outputs.json")))["test007"]


def _test007_spec():
    """test007.ctr: 20 A box, CH4 LJ sig 3.166 eps 78.15 K hicut 9 locut 0.3
    (atom_atom_file:7), RINSERT/RDELETE/RTRANSLATE 1:1:1, 0.4 A, 300 K."""
    s = inputs.read_ctrlfile(os.path.join(GOLD, "test007-gcmc-CH4", "test007.ctr"))
    aa = inputs.read_atom_atom(os.path.join(GOLD, "test007-gcmc-CH4", "atom_atom_file"))
    p = aa[("CH4", "CH4")]
    g = s["gcmc"]
    sp = g["species"][0]
    moves = [(m["type"], m["weight"]) for m in sp["moves"]]
    dtrans = [m["delta"] for m in sp["moves"] if m["type"] == "RTRANSLATE"][0]
    L = s["simcell"]["fund"]["ell"] * np.array(s["simcell"]["nxyz"])
    spec = dict(full_ell=L, fund_ell=L, ref=[[0.0, 0.0, 0.0]], stype=[0], sq=[0.0], nst=1,
                lj=([p["A"]], [p["B"]], [p["hicut2"]], [p["locut2"]], [1]), T=g["tk"],
                moves=moves, dtrans=dtrans)
    return spec, sp["fugacity"][0], s["general"]["iseed"], s["general"]["niterations"]


def test_adams_B_matches_reference():
    # gcmc_genfuglist: B = 2.961 for 10^4 kPa, 8000 A^3, 300 K (output_reference:167)
    assert round(O.adams_B(REF["pressure_kPa"], REF["volume"], REF["T"]), 3) == REF["B"]


def test_test007_trajectory_counts_exact():
    spec, fug, seed, niter = _test007_spec()
    assert (fug, seed, niter) == (10000.0, 22111971, 20000)
    g = O.Gcmc(spec, fug, seed)
    for cp in REF["checkpoints"]:
        g.run(cp["step"] - g.st.nsteps)
        st = g.st
        assert [st.acc[0], st.att[0]] == cp["insert"], cp["step"]
        assert [st.acc[1], st.att[1]] == cp["delete"], cp["step"]
        assert [st.acc[2], st.att[2]] == cp["translate"], cp["step"]
        assert st.N == int(cp["N"]), cp["step"]
    assert abs(g.st.sumN / g.st.nsteps - REF["N_cum_step20000"]) < 0.005
    # final stored energy -6.435 kJ/mol (output_reference:351-356), drift check
    assert abs(g.st.energy * 4.184 - REF["final_energy_kJ"]) < 5e-4
    assert abs(g.total_energy() - g.st.energy) < 1e-9


def co2_spec(L=24.0, kappa=0.30, kmax=7, rc=11.5, lj=True):
    A_O, B_O = inputs.lj_AB(3.05, 79.0)
    A_C, B_C = inputs.lj_AB(2.80, 27.0)
    A_OC, B_OC = inputs.lj_AB((3.05 + 2.80) / 2, np.sqrt(79.0 * 27.0))
    # site types: 0 = O, 1 = C
    A = [A_O, A_OC, A_OC, A_C]
    B = [B_O, B_OC, B_OC, B_C]
    on = [1, 1, 1, 1] if lj else [0, 0, 0, 0]
    return dict(full_ell=[L, L, L], fund_ell=[L, L, L],
                ref=[[0, 0, -1.16], [0, 0, 0], [0, 0, 1.16]], stype=[0, 1, 0], sq=[-0.35, 0.70, -0.35],
                nst=2, lj=(A, B, [rc * rc] * 4, [0.5 ** 2] * 4, on), T=300.0,
                ewald=dict(kappa=kappa, kmax=kmax, rc=rc),
                moves=[("RINSERT", 1.0), ("RDELETE", 1.0), ("RTRANSLATE", 1.0), ("RROTATE", 1.0)],
                dtrans=0.4, drot=0.4)


def test_ewald_running_energy_equals_recompute():
    # every accepted delta (insert/delete/translate/rotate, incl. the
    # structure-factor updates) must add up to the full recompute
    g = O.Gcmc(co2_spec(), 3000.0, 3001)
    g.run(3000)
    assert g.st.N > 10
    tot = g.total_energy()
    assert abs(tot - g.st.energy) < 1e-9 * max(1.0, abs(tot)), (tot, g.st.energy)
    # S(k) running totals equal a recompute from the coordinates
    s0 = g.Sre.copy(), g.Sim.copy()
    O._gcmc_lib().mo_gcmc_recompute_sk(__import__("ctypes").byref(g.sys), __import__("ctypes").byref(g.st))
    assert np.max(np.abs(g.Sre - s0[0])) < 1e-10 and np.max(np.abs(g.Sim - s0[1])) < 1e-10


def _ewald_numpy(com, ang_ref, q, L, kappa, kmax, rc):
    """Independent Ewald energy (exact erfc, numpy) of rigid molecules with
    intramolecular pairs excluded: real + reciprocal (full k sphere) - self
    - intra erf correction, kcal/mol."""
    from math import erfc, erf, sqrt, pi
    pos = np.concatenate(ang_ref)
    nm, ns = len(ang_ref), len(q)
    qq = np.tile(q, nm)
    mol = np.repeat(np.arange(nm), ns)
    e_real = 0.0
    for i in range(len(pos)):
        for j in range(i + 1, len(pos)):
            d = pos[i] - pos[j]
            d -= L * np.round(d / L)
            r = np.linalg.norm(d)
            if mol[i] == mol[j]:
                e_real -= qq[i] * qq[j] * erf(kappa * r) / r
            elif r < rc:
                e_real += qq[i] * qq[j] * erfc(kappa * r) / r
    V = L ** 3
    m = np.arange(-kmax, kmax + 1)
    MX, MY, MZ = np.meshgrid(m, m, m, indexing="ij")
    kv = np.stack([MX, MY, MZ], -1).reshape(-1, 3) * (2 * pi / L)
    kv = kv[np.any(kv != 0, axis=1)]
    k2 = (kv ** 2).sum(1)
    S = (qq[None, :] * np.exp(1j * kv @ pos.T)).sum(1)
    e_rec = (2 * pi / V * np.exp(-k2 / (4 * kappa ** 2)) / k2 * np.abs(S) ** 2).sum()
    e_self = -kappa / sqrt(pi) * (qq ** 2).sum()
    return (e_real + e_rec + e_self) * 332.064


def test_ewald_total_is_kappa_independent():
    # a converged Ewald sum does not depend on the splitting parameter: the
    # oracle (A&S erfc, reference twopi) at kappa 0.30 against an exact-erfc
    # numpy Ewald at kappa 0.45 with a larger k cube
    rng = np.random.default_rng(11)
    L = 18.0
    sp = co2_spec(L=L, kappa=0.30, kmax=9, rc=8.99, lj=False)
    g = O.Gcmc(sp, 1000.0, 5)
    com = rng.uniform(0, L, (6, 3))
    ang = np.stack([rng.uniform(-np.pi, np.pi, 6), np.arccos(rng.uniform(-1, 1, 6)),
                    rng.uniform(-np.pi, np.pi, 6)], 1)
    g.set_config(com, ang)
    e_or = g.total_energy()
    mols = [g.r[m] for m in range(6)]
    e_np = _ewald_numpy(None, mols, np.array(sp["sq"]), L, 0.45, 14, 8.99)
    assert abs(e_or - e_np) < 2e-4 * max(1.0, abs(e_np)), (e_or, e_np)


def test_test007_final_configuration_equals_reference_restart():
    """The reference's restart file written at the end of test007
    (tests/test007-gcmc-CH4/100.res.1: 21 molecules, principal coordinates
    f13.5) against the oracle chain's final molecules, slot by slot."""
    import music_gcmc_run as MG
    spec, fug, seed, niter = _test007_spec()
    g = O.Gcmc(spec, fug, seed)
    g.run(niter)
    rp = MG.read_restart_rp(os.path.join(GOLD, "test007-gcmc-CH4", "100.res.1"))["CH4"]
    assert g.st.N == len(rp) == 21
    assert np.abs(g.rp[:g.st.N] - rp).max() < 5.01e-6


def _replay_deltas(tr, d0):
    """The displacements rigidmoves_adjustdeltatrans / _adjustdeltarot
    (rigidmoves.F90:2586-2617) leave after the attempts in a trace: every
    TRANSLATE (2) / ROTATE (3) step with molecules present is an attempt."""
    d = list(d0)
    att, acc = [0, 0], [0, 0]
    for mv, a, n, _ in tr:
        mv = int(mv)
        if mv not in (2, 3) or int(n) == 0:
            continue
        k = mv - 2
        att[k] += 1
        acc[k] += int(a)
        ratio = acc[k] / att[k]
        if ratio < 0.49:
            d[k] = max(d[k] * 0.95, 0.01)
        elif ratio > 0.51:
            d[k] = min(d[k] * 1.05, 2.0)
    return d, att


def test_adjustable_displacements_follow_acceptance():
    # both move away from their start, within [0.01, 2.0]
    sp = co2_spec()
    sp.update(adjust_trans=1, adjust_rot=1, dtrans=1.9, drot=1.9)
    g = O.Gcmc(sp, 3000.0, 3011, trace=4000)
    g.run(4000)
    tr = g.trace_array(4000)
    want, att = _replay_deltas(tr, (1.9, 1.9))
    assert min(att) > 300
    assert list(g.st.delta) == want
    assert want[0] != 1.9 and want[1] != 1.9 and all(0.01 <= x <= 2.0 for x in want)
    assert list(g.st.att[2:4]) == att
    # adjust off: the displacements stay as given
    sp0 = co2_spec()
    g0 = O.Gcmc(sp0, 3000.0, 3011)
    g0.run(500)
    assert list(g0.st.delta) == [0.4, 0.4]
