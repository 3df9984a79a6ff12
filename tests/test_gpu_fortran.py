"""The Fortran iso_c_binding layer (music-code_amd/fortran/music_hip_mod.f90)
and the Fortran drivers over it: test012's LJ map built through bind(C)
interfaces gives the reference's counts and the same map-file bytes as the
C-ABI called from Python; music_mapmaker reads the reference's ctrlfiles and
writes the same map files as the Python host; music_gcmc reproduces the
reference's printed test007 statistics and matches oracle chains on map and
Ewald systems, batched and chained."""
import hashlib
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

import music_gcmc_run as MG
from musichip import hip as H
from musichip import inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
EXE = os.path.join(os.path.dirname(HERE), "music-code_amd", "bin", "map_from_list")


def test_fortran_driver_builds_test012_map(hip):
    print("case: Fortran driver, test012 LJ map FP64", flush=True)
    assert os.path.exists(EXE), "build the Fortran drivers (make -C music-code_amd fortran)"
    s = inputs.load_system(os.path.join(GOLD, "test012-pmapgen", "test012.ctr"))
    fw = s["mols"]["Lattice"]
    A, B, h2, l2, on = inputs.lj_table_for_probe(s, int(s["mols"]["Probe"]["type"][0]))
    job = s["ctrl"]["mapmaker"][0]
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "fw.txt")
        with open(inp, "w") as f:
            f.write("%d %d\n" % (len(fw["full_xyz"]), len(A)))
            f.write(" ".join("%.17g" % v for v in list(s["fund_ell"]) + list(s["full_ell"])) + "\n")
            f.write("%.17g %.17g 0\n" % (job["gridspace"], job["hipotcut"]))
            for x, t in zip(fw["full_xyz"], fw["full_type"]):
                f.write("%.17g %.17g %.17g %d\n" % (x[0], x[1], x[2], t))
            for i in range(len(A)):
                f.write("%.17g %.17g %.17g %.17g\n" % (A[i], B[i], h2[i], l2[i]))
        fmap = os.path.join(td, "fortran.map")
        out = subprocess.run([EXE, inp, fmap], capture_output=True, text=True, timeout=120)
        print(out.stdout, out.stderr, flush=True)
        assert out.returncode == 0
        vals = [int(v) for v in out.stdout.split()]
        assert vals[:3] == [50, 50, 30]
        assert vals[4:8] == [28904, 20544, 49448, 29627]
        hip.upload_framework(fw["full_xyz"], fw["full_type"], fw["full_q"], s["fund_ell"], s["full_ell"])
        hip.set_lj(A, B, h2, l2, on)
        m = hip.build_map(H.MH_MAP_LJ, H.MH_PREC_FP64, job["gridspace"], job["hipotcut"])
        pmap = os.path.join(td, "python.map")
        m.write(pmap, [], job["hipotcut"])
        assert vals[3] == m.nfsize
        da, db = open(fmap, "rb").read(), open(pmap, "rb").read()
        assert hashlib.sha256(da).hexdigest() == hashlib.sha256(db).hexdigest()
        m.close()


MAPMAKER = os.path.join(os.path.dirname(HERE), "music-code_amd", "bin", "music_mapmaker")


@pytest.mark.parametrize("case", ["test012-pmapgen/test012.ctr", "test013-emapgen/test013.ctr"])
def test_fortran_music_mapmaker_matches_python_host(hip, case):
    """music_mapmaker (Fortran, ctrlfile in, map files out) against the
    Python host's mapmaker on the same ctrlfile: identical map files."""
    from musichip import mapmaker
    print("case: music_mapmaker %s" % case, flush=True)
    ctrl = os.path.join(GOLD, case)
    with tempfile.TemporaryDirectory() as tf, tempfile.TemporaryDirectory() as tp:
        out = subprocess.run([MAPMAKER, ctrl], cwd=tf, capture_output=True, text=True, timeout=300)
        print(out.stdout, out.stderr, flush=True)
        assert out.returncode == 0
        res = mapmaker.make_maps(ctrl, ctx=hip, precision=H.MH_PREC_FP64, outdir=tp)
        assert res
        for job, m, _ in res:
            name = os.path.basename(job["path"])
            fa, fb = os.path.join(tf, name), job["path"]
            assert os.path.exists(fa), name
            da, db = open(fa, "rb").read(), open(fb, "rb").read()
            assert hashlib.sha256(da).hexdigest() == hashlib.sha256(db).hexdigest(), name
            assert ("%12d" * 4) % tuple(m.counts) in out.stdout.replace("\n", " ") or \
                all(str(c) in out.stdout for c in m.counts)
            m.close()


# ----------------------------------------------------------- music_gcmc
REF7 = json.load(open(os.path.join(GOLD, "reference_outputs.json")))["test007"]


def test_music_gcmc_reproduces_test007_output(hip):
    """test007.ctr through the Fortran driver: at every printed step the
    reference's acceptance counts and N exactly, its block / cumulative
    averages and standard deviation of N to the printed 2 decimals, its final
    energy and loading (tests/test007-gcmc-CH4/output_reference:168-361)."""
    print("case: music_gcmc test007.ctr FP64", flush=True)
    with tempfile.TemporaryDirectory() as td:
        text = MG.run([os.path.join(GOLD, "test007-gcmc-CH4", "test007.ctr")], td)
    blocks, ends = MG.parse(text)
    assert len(blocks) == len(REF7["checkpoints"]) == 20
    for b, cp in zip(blocks, REF7["checkpoints"]):
        assert b["step"] == cp["step"]
        assert b["moves"]["INSERT"] == tuple(cp["insert"])
        assert b["moves"]["DELETE"] == tuple(cp["delete"])
        assert b["moves"]["TRANSLATE"] == tuple(cp["translate"])
        assert b["N"] == cp["N"]
        for k, rk in (("blk", "N_blk"), ("cum", "N_cum"), ("std", "N_std")):
            assert abs(b[k] - cp[rk]) <= 0.0101, (cp["step"], k, b[k], cp[rk])
    e = ends[1]
    assert abs(e["actual"] - REF7["final_energy_kJ"]) < 1.5e-3
    assert abs(e["stored"] - REF7["final_energy_kJ"]) < 1.5e-3
    assert e["dev"] < 1e-9
    assert abs(e["loading"] - REF7["N_cum_step20000"]) <= 0.0101


def _fixture_maps(hip, td):
    """Lattice.LJ.map and Lattice.EWALD.map by music_mapmaker from test012.ctr
    and test013.ctr, loaded back for the oracle."""
    for case in ("test012-pmapgen/test012.ctr", "test013-emapgen/test013.ctr"):
        out = subprocess.run([MG.MAPMAKER, os.path.join(GOLD, case)], cwd=td, capture_output=True, text=True,
                             timeout=300)
        print(out.stdout, flush=True)
        assert out.returncode == 0, out.stdout + out.stderr
    fund = [20.0, 20.0, 12.0]
    L = np.array([20.0, 20.0, 24.0])
    hip.upload_framework(np.zeros((1, 3)), np.zeros(1, dtype=np.int32), None, fund, L)
    pm = hip.load_map(os.path.join(td, "Lattice.LJ.map"), fund)
    em = hip.load_map(os.path.join(td, "Lattice.EWALD.map"), fund)
    mlist = []
    for m in (pm, em):
        ptr, vals = m.get()
        mlist.append((m.npnts, m.step, ptr, vals))
        m.close()
    return mlist


def _oracle_blocks(g, steps, every=1000, block=100):
    """(att/acc, N, block mean, cumulative mean) of an oracle chain at every
    print step, its counters as the reference's statistics (block of the
    last `block` steps)."""
    out = []
    base = g.st.nsteps
    sumN0 = g.st.sumN
    done = 0
    while done < steps:
        g.run(every - block)
        s0 = g.st.sumN
        g.run(block)
        done += every
        out.append(dict(att=[g.st.att[i] for i in range(4)], acc=[g.st.acc[i] for i in range(4)], N=g.st.N,
                        blk=(g.st.sumN - s0) / block, cum=(g.st.sumN - sumN0) / (g.st.nsteps - base)))
    return out


def _compare(blocks, oracle, tags=("INSERT", "DELETE", "TRANSLATE", "ROTATE")):
    assert len(blocks) == len(oracle)
    for b, o in zip(blocks, oracle):
        for i, t in enumerate(tags):
            if t in b["moves"]:
                assert b["moves"][t] == (o["acc"][i], o["att"][i]), (b["step"], t)
        assert b["N"] == o["N"]
        assert abs(b["blk"] - o["blk"]) <= 0.0051 and abs(b["cum"] - o["cum"]) <= 0.0051


def test_music_gcmc_maps_batched_and_chained(hip):
    """gcmc.ctr (probe in the charged lattice, LJ + Ewald maps from
    music_mapmaker, 3 state points): the batched run (independent chains,
    seeds iseed + p - 1) and the chained run (the reference's sequence:
    configuration and random stream carried, statistics reset) against
    oracle chains on the same maps, move for move at every printed step."""
    print("case: music_gcmc gcmc-maps/gcmc.ctr, batched and chain, FP64", flush=True)
    spec = MG.probe_spec()
    press = [10.0, 100.0, 1000.0]
    with tempfile.TemporaryDirectory() as td:
        mlist = _fixture_maps(hip, td)
        spec = dict(spec, maps=mlist, site_pmap=[0], site_emap=[1])
        ctrl = os.path.join(MG.FIX, "gcmc.ctr")
        blocks, ends = MG.parse(MG.run([ctrl], td))
        for p in range(3):
            g = O.Gcmc(spec, press[p], 1892134 + p)
            _compare([b for b in blocks if b["sim"] == p + 1], _oracle_blocks(g, 3000))
            assert abs(ends[p + 1]["stored"] - g.st.energy * 4.184) < 1.5e-3
            assert ends[p + 1]["dev"] < 1e-8
        # chained: one oracle chain through the three points
        cblocks, cends = MG.parse(MG.run([ctrl, "chain"], td))
        g = O.Gcmc(spec, press[0], 1892134)
        vol = float(np.prod(spec["full_ell"]))
        Ns = []
        for p in range(3):
            if p:
                g.st.B = O.adams_B(press[p], vol, spec["T"])
                for i in range(4):
                    g.st.att[i] = 0
                    g.st.acc[i] = 0
                g.st.sumN, g.st.nsteps = 0.0, 0
            _compare([b for b in cblocks if b["sim"] == p + 1], _oracle_blocks(g, 3000))
            Ns.append(g.st.N)
        assert [b["N"] for b in blocks if b["sim"] == 1] == [b["N"] for b in cblocks if b["sim"] == 1]
        assert Ns[-1] > 0


def test_music_gcmc_bent_sorbate_ewald(hip):
    """bent.ctr (3-site rigid sorbate on its principal axes, LBMIX pair,
    Ewald between sorbates, rotations) through the driver against the oracle
    chain built from the fixture values, move for move; the FP32 build runs
    the same input to within statistical agreement of the loading."""
    print("case: music_gcmc gcmc-maps/bent.ctr FP64 and FP32", flush=True)
    with tempfile.TemporaryDirectory() as td:
        ctrl = os.path.join(MG.FIX, "bent.ctr")
        blocks, ends = MG.parse(MG.run([ctrl], td))
        g = O.Gcmc(MG.bent_spec(), 5000.0, 424242)
        _compare(blocks, _oracle_blocks(g, 2000, every=500, block=250))
        assert abs(ends[1]["stored"] - g.st.energy * 4.184) < 1.5e-3
        assert ends[1]["dev"] < 1e-8
        b32, e32 = MG.parse(MG.run([ctrl, "fp32"], td))
        assert len(b32) == 4 and e32[1]["dev"] < 1e-3
        assert b32[-1]["moves"]["ROTATE"][1] > 0


def test_music_gcmc_bent_sorbate_plain_coulomb(hip):
    """bent.ctr with the sorbate electrostatics switched to plain Coulomb
    (spc_spc 'COUL BASIC', atom_atom 'COUL HICUT@9.0' lines): the driver
    against the oracle chain with the same Coulomb term, move for move."""
    import shutil
    print("case: music_gcmc bent sorbate with plain Coulomb, FP64", flush=True)
    with tempfile.TemporaryDirectory() as td:
        shutil.copytree(MG.FIX, os.path.join(td, "f"), copy_function=shutil.copyfile)
        d = os.path.join(td, "f")
        with open(os.path.join(d, "spc_spc_bent"), "w") as f:
            f.write("Bent     Bent     NCOUL BASIC LJ FAST\nBent     Bent     COUL  BASIC FAST\n")
        with open(os.path.join(d, "atom_atom_file"), "a") as f:
            f.write("Ow       Ow     COUL HICUT@9.0\nHw       Hw     COUL HICUT@9.0\nOw       Hw     COUL HICUT@9.0\n")
        blocks, ends = MG.parse(MG.run([os.path.join(d, "bent.ctr")], td))
        spec = MG.bent_spec()
        spec.pop("ewald")
        spec["coul"] = dict(on=[1] * 4, hicut=9.0, locut=1.0e-3, shield=1.0)
        g = O.Gcmc(spec, 5000.0, 424242)
        _compare(blocks, _oracle_blocks(g, 2000, every=500, block=250))
        assert abs(ends[1]["stored"] - g.st.energy * 4.184) < 1.5e-3
        assert ends[1]["dev"] < 1e-8
        assert blocks[-1]["N"] > 3


def test_music_gcmc_biased_moves(hip):
    """gcmc_bias.ctr: BINSERT (bias map Lattice.LJ.map at 300 K) / BDELETE
    for the probe in the charged lattice through the driver, three state
    points, against oracle chains with the oracle's bias map, move for move."""
    print("case: music_gcmc gcmc-maps/gcmc_bias.ctr FP64", flush=True)
    spec = dict(MG.probe_spec(), moves=[("BINSERT", 1.0), ("BDELETE", 1.0)])
    press = [10.0, 100.0, 1000.0]
    with tempfile.TemporaryDirectory() as td:
        mlist = _fixture_maps(hip, td)
        npn, step, ptr, vals = mlist[0]
        bias = O.Bmap(npn, step, [20.0, 20.0, 12.0], ptr, vals, 300.0)
        spec = dict(spec, maps=mlist, site_pmap=[0], site_emap=[1], bias=bias)
        blocks, ends = MG.parse(MG.run([os.path.join(MG.FIX, "gcmc_bias.ctr")], td))
        for p in range(3):
            g = O.Gcmc(spec, press[p], 1892134 + p)
            _compare([b for b in blocks if b["sim"] == p + 1], _oracle_blocks(g, 3000))
            assert abs(ends[p + 1]["stored"] - g.st.energy * 4.184) < 1.5e-3
        assert sum(b["moves"]["INSERT"][0] for b in blocks if b["it"] == 3000) > 20


def test_music_gcmc_cavity_bias(hip):
    """gcmc_cavity.ctr: MAP_CAVITY_BIAS (bias map Lattice.LJ.map at 300 K, 20
    trials, a 10 x 10 x 12 cube list tabulated from the map below 0
    kcal/mol) through the driver, three state points, against oracle
    chains with the oracle's bias map and cavity list, move for move."""
    print("case: music_gcmc gcmc-maps/gcmc_cavity.ctr FP64", flush=True)
    from test_gpu_cavity import tab_from_arrays
    spec = dict(MG.probe_spec(), moves=[("BINSERT", 1.0), ("BDELETE", 1.0)])
    press = [10.0, 100.0, 1000.0]
    with tempfile.TemporaryDirectory() as td:
        mlist = _fixture_maps(hip, td)
        npn, step, ptr, vals = mlist[0]
        bias = O.Bmap(npn, step, [20.0, 20.0, 12.0], ptr, vals, 300.0)
        tab = tab_from_arrays(npn, step, ptr, vals, 0.0, (10, 10, 12), (1, 1, 2), [20.0, 20.0, 24.0])
        assert 0 < tab.sum() < tab.size
        cav = dict(n=(10, 10, 12), tab=tab.astype(np.int8), mask=1)
        spec = dict(spec, maps=mlist, site_pmap=[0], site_emap=[1], bias=bias, cavity=cav, cav_ntrials=20, cav_map=1)
        blocks, ends = MG.parse(MG.run([os.path.join(MG.FIX, "gcmc_cavity.ctr")], td))
        for p in range(3):
            g = O.Gcmc(spec, press[p], 1892134 + p)
            _compare([b for b in blocks if b["sim"] == p + 1], _oracle_blocks(g, 3000))
            assert abs(ends[p + 1]["stored"] - g.st.energy * 4.184) < 1.5e-3
        assert sum(b["moves"]["INSERT"][0] for b in blocks if b["it"] == 3000) > 20


def _t007_tree(td, **edits):
    """test007's inputs in td (ctrlfile edited line by line: key -> (old, new))."""
    import shutil
    # contents only: the checked-out tree may be read-only
    shutil.copytree(os.path.join(GOLD, "test007-gcmc-CH4"), os.path.join(td, "t7"), dirs_exist_ok=True,
                    copy_function=shutil.copyfile)
    if not os.path.exists(os.path.join(td, "potentials")):
        shutil.copytree(os.path.join(GOLD, "potentials"), os.path.join(td, "potentials"),
                        copy_function=shutil.copyfile)
    ctrl = os.path.join(td, "t7", "test007.ctr")
    txt = open(os.path.join(GOLD, "test007-gcmc-CH4", "test007.ctr")).read()
    for old, new in edits.values():
        assert old in txt, old
        txt = txt.replace(old, new, 1)
    open(ctrl, "w").write(txt)
    return ctrl


ICRASH = ("100000        # Steps between writes to restart file", "%d        # Steps between writes to restart file")
NITER = ("20000         # No. of MD/MCiterations", "%d         # No. of MD/MCiterations")


def test_music_gcmc_restart_file_and_resume(hip):
    """Restart files: test007 with a restart written at step 20000 is the
    reference's own restart file (tests/test007-gcmc-CH4/100.res.1) byte for
    byte; a run split at 10000 and resumed from its restart (full-precision
    state and random-number state) ends in the same file."""
    print("case: music_gcmc restart files and resume (test007)", flush=True)
    gold = open(os.path.join(GOLD, "test007-gcmc-CH4", "100.res.1"), "rb").read()
    with tempfile.TemporaryDirectory() as td:
        ctrl = _t007_tree(td, icrash=(ICRASH[0], ICRASH[1] % 20000))
        MG.run([ctrl], td)
        assert open(os.path.join(td, "temp.res.1"), "rb").read() == gold
    with tempfile.TemporaryDirectory() as td:
        ctrl = _t007_tree(td, icrash=(ICRASH[0], ICRASH[1] % 10000), niter=(NITER[0], NITER[1] % 10000))
        MG.run([ctrl], td)
        assert os.path.exists(os.path.join(td, "temp.res.1.state"))
        ctrl2 = _t007_tree(td, icrash=(ICRASH[0], ICRASH[1] % 10000), niter=(NITER[0], NITER[1] % 10000),
                           src=("GCMC NULL # Source of configs", "RESTARTFILE ../temp.res.1 # Source of configs"),
                           res=("temp.res   # Name of Restart File", "temp2.res   # Name of Restart File"))
        out = MG.run([ctrl2, "resume"], td)
        assert "started from" in out
        got = open(os.path.join(td, "temp2.res.1"), "rb").read().split(b"\n")
        want = gold.split(b"\n")
        assert got[1].strip() == b"10000" and want[1].strip() == b"20000"
        assert got[:1] + got[2:] == want[:1] + want[2:]


ICONFIG = ("500000         # Steps between writes to config. file", "%d         # Steps between writes to config. file")


def test_music_gcmc_configuration_datafile(hip):
    """Configuration files (datafile.F90 layout): test007 with a
    configuration every 1000 steps writes temp.con.1 whose header holds the
    ctrlfile, whose extra-info lines are datafile_writeExtraInfo's, and whose
    20 configurations hold the N printed at the same steps, the stored
    energy, and at step 20000 the molecules of the reference's own restart
    file (tests/test007-gcmc-CH4/100.res.1) to its print precision."""
    from musichip import datafile as D
    print("case: music_gcmc configuration datafile (test007)", flush=True)
    with tempfile.TemporaryDirectory() as td:
        ctrl = _t007_tree(td, iconfig=(ICONFIG[0], ICONFIG[1] % 1000))
        blocks, ends = MG.parse(MG.run([ctrl], td))
        d = D.read(os.path.join(td, "temp.con.1"))
        txt = open(ctrl).read().split("\n")
    if txt and txt[-1] == "":
        txt = txt[:-1]
    assert [x.rstrip() for x in d["ctrl_lines"]] == [x[:96].rstrip() for x in txt]
    assert d["simno"] == 1 and d["flags"] == dict(tot_nrg=True, pair_nrg=True, intra_nrg=False, posns=True,
                                                   vels=False, time=False)
    assert d["extra"] == D.extra_info_lines(1, 10000.0, 300.0, "GCMC", ["CH4"], [False], 20000, 1000)
    cf = d["configs"]
    assert len(cf) == 20 and len(blocks) == 20
    for b, c in zip(blocks, cf):
        assert c["species"][0]["nmoles"] == b["N"], b["step"]
        assert c["pair"][0] == (c["total"], 0.0) and c["kin_intra"][0] == (0.0, 0.0)
    assert abs(cf[-1]["total"] - ends[1]["stored"]) < 1e-3
    want = MG.read_restart_rp(os.path.join(GOLD, "test007-gcmc-CH4", "100.res.1"))["CH4"]
    got = cf[-1]["species"][0]["rp"]
    assert got.shape == want.shape
    assert np.abs(got - want).max() < 6e-6


def test_music_post_isotherm_from_driver_datafiles(hip, monkeypatch):
    """music_post over configuration files music_gcmc wrote: test007 at three
    pressures (100 .. 10000 kPa, linear spacing as gcmc_genpressures gives
    for two decades) with a configuration every 1000 steps; the isotherm
    holds each file's pressure and the mean N over the configurations the
    skips and blocks keep."""
    from musichip import datafile as D
    from musichip import post as P
    print("case: music_post isotherm over music_gcmc datafiles (test007, 3 pressures)", flush=True)
    with tempfile.TemporaryDirectory() as td:
        ctrl = _t007_tree(td, iconfig=(ICONFIG[0], ICONFIG[1] % 1000),
                          npress=("1                    # No. of pressure values",
                                  "3                    # No. of pressure values"),
                          fug=("10000             # Fugacity Range (kPa)", "100, 10000        # Fugacity Range (kPa)"))
        MG.run([ctrl], td)
        monkeypatch.chdir(td)
        open("post.ctr", "w").write(
            "-- Post Processor Information --\nGCMC\ntemp.con\n1, 3\nregen.ctr\npost.out\n20, 0\n"
            "-- Post : Energy Average Info --\n4\n-- Post : Loading Average Info --\n4\n")
        run = P.PostRun("post.ctr")
        run.run()
        # 20 configurations, the first 4 skipped: blocks of Int(15/4)+1 = 4
        assert (run.firstconf, run.lastconf) == (5, 20)
        iso = open("isotherm.CH4").read().split("\n")
        zs = []
        for q in range(3):
            d = D.read("temp.con.%d" % (q + 1))
            p = float(d["extra"][1].split()[0])
            assert abs(p - (100.0, 5050.0, 10000.0)[q]) < 1e-9
            N = np.array([c["species"][0]["nmoles"] for c in d["configs"]], dtype=float)
            assert len(N) == 20
            zs.append(N[4:20].mean())
            assert iso[4 + q] == "%-15s%10s%12.7f" % (P.real2str(p, 15)[:15], "", zs[-1])
        assert zs[0] < zs[2]
        assert iso[7].startswith("# Results of a Langmuir Fit")


ADJ_T = ("0.4, 0              # Del-R Ang., adjust delta?(0=NO, 1=YES)", "1.9, 1              # Del-R Ang., adjust delta?(0=NO, 1=YES)")
ADJ_R = ("0.4, 0              # Del-Theta Rad., adjust delta?(0=NO, 1=YES)", "0.4, 1              # Del-Theta Rad., adjust delta?(0=NO, 1=YES)")


PRESS_1E5 = ("10000             # Fugacity Range (kPa)", "100000            # Fugacity Range (kPa)")


def test_music_gcmc_adjustable_displacements_resume(hip):
    """test007 at 10^5 kPa (about 100 molecules, translate acceptance near
    1/2) with an adjustable translate displacement: the printed "Max disp"
    leaves its start and stays inside (0.01, 2.0), and a run split at 10000
    steps and resumed from its restart (.state holds the displacements and
    the attempt / acceptance counts that drive them) prints the same
    displacements from step 10000 on and ends in the restart file of the
    unbroken run."""
    import re
    print("case: music_gcmc adjustable displacements, unbroken vs resumed (test007 at 1e5 kPa)", flush=True)
    with tempfile.TemporaryDirectory() as td:
        ctrl = _t007_tree(td, icrash=(ICRASH[0], ICRASH[1] % 20000), adjt=ADJ_T, adjr=ADJ_R, p=PRESS_1E5)
        out = MG.run([ctrl], td)
        disp = [float(x) for x in re.findall(r"Max disp:\s+([0-9.]+)", out)]
        # test007 moves: insert, delete, translate (one "Max disp" per block)
        assert len(disp) >= 20 and all(0.01 <= d <= 2.0 for d in disp)
        assert any(abs(d - 1.9) > 1e-3 for d in disp)
        assert 0.02 < disp[-1] < 1.99, disp[-5:]
        full = open(os.path.join(td, "temp.res.1"), "rb").read().split(b"\n")
    with tempfile.TemporaryDirectory() as td:
        ctrl = _t007_tree(td, icrash=(ICRASH[0], ICRASH[1] % 10000), niter=(NITER[0], NITER[1] % 10000),
                          adjt=ADJ_T, adjr=ADJ_R, p=PRESS_1E5)
        MG.run([ctrl], td)
        ctrl2 = _t007_tree(td, icrash=(ICRASH[0], ICRASH[1] % 10000), niter=(NITER[0], NITER[1] % 10000),
                           src=("GCMC NULL # Source of configs", "RESTARTFILE ../temp.res.1 # Source of configs"),
                           res=("temp.res   # Name of Restart File", "temp2.res   # Name of Restart File"),
                           adjt=ADJ_T, adjr=ADJ_R, p=PRESS_1E5)
        out2 = MG.run([ctrl2, "resume"], td)
        disp2 = [float(x) for x in re.findall(r"Max disp:\s+([0-9.]+)", out2)]
        assert disp2[-len(disp) // 2:] == disp[-len(disp) // 2:]
        got = open(os.path.join(td, "temp2.res.1"), "rb").read().split(b"\n")
        assert got[:1] + got[2:] == full[:1] + full[2:]


def _oracle_mix_blocks(g, names, steps, every=500, block=250):
    """Per print step and species of a GcmcMix chain: att/acc, N, block and
    cumulative means of N (the species' sums over the common step count)."""
    out = []
    st0 = g.ch[0].st
    base = st0.nsteps
    sum0 = [c.st.sumN for c in g.ch]
    done = 0
    while done < steps:
        g.run(every - block)
        s0 = [c.st.sumN for c in g.ch]
        g.run(block)
        done += every
        out.append({n: dict(att=[c.st.att[i] for i in range(4)], acc=[c.st.acc[i] for i in range(4)], N=c.st.N,
                            blk=(c.st.sumN - s0[k]) / block, cum=(c.st.sumN - sum0[k]) / (st0.nsteps - base))
                    for k, (n, c) in enumerate(zip(names, g.ch))})
    return out


def _compare_mix(blocks, oracle, tags=("INSERT", "DELETE", "TRANSLATE", "ROTATE")):
    assert len(blocks) == len(oracle)
    for b, o in zip(blocks, oracle):
        assert sorted(b["spc"]) == sorted(o)
        for n, ob in o.items():
            bs = b["spc"][n]
            for i, t in enumerate(tags):
                if t in bs["moves"]:
                    assert bs["moves"][t] == (ob["acc"][i], ob["att"][i]), (b["step"], n, t)
            assert bs["N"] == ob["N"], (b["step"], n)
            assert abs(bs["blk"] - ob["blk"]) <= 0.0051 and abs(bs["cum"] - ob["cum"]) <= 0.0051


def test_music_gcmc_mixture(hip):
    """mix.ctr (the bent sorbate with Ewald and a one-site LJ sorbate, two
    state points): the batched run against oracle mixture chains
    (mo_gcmc_run_mix) per point, species by species and move for move at
    every printed step; the chained run against one oracle chain whose
    fugacities change between the points; a run split at a restart and
    resumed ends in the unbroken run's restart file."""
    print("case: music_gcmc gcmc-maps/mix.ctr batched, chain, restart, FP64", flush=True)
    names = ["Bent", "Neutral"]
    fug = [[5000.0, 2000.0], [8000.0, 1000.0]]
    specs = MG.mix_specs()
    with tempfile.TemporaryDirectory() as td:
        ctrl = os.path.join(MG.FIX, "mix.ctr")
        blocks, ends = MG.parse_mix(MG.run([ctrl], td))
        for p in range(2):
            g = O.GcmcMix(specs, fug[p], 424242 + p)
            _compare_mix([b for b in blocks if b["sim"] == p + 1], _oracle_mix_blocks(g, names, 2000))
            assert abs(ends[p + 1]["stored"] - g.energy * 4.184) < 1.5e-3
            assert ends[p + 1]["dev"] < 1e-8
            assert [ends[p + 1]["loading"][n][0] for n in names] == fug[p]
        assert min(b["spc"][n]["N"] for b in blocks[-2:] for n in names) > 0
        # chained: the configuration of both species and the random stream
        # carry over, the statistics restart
        cblocks, _ = MG.parse_mix(MG.run([ctrl, "chain"], td))
        g = O.GcmcMix(specs, fug[0], 424242)
        for p in range(2):
            if p:
                for k, c in enumerate(g.ch):
                    c.st.B = O.adams_B(fug[p][k], 8000.0, 300.0)
                    for i in range(4):
                        c.st.att[i] = 0
                        c.st.acc[i] = 0
                    c.st.sumN = 0.0
                g.ch[0].st.nsteps = 0
            _compare_mix([b for b in cblocks if b["sim"] == p + 1], _oracle_mix_blocks(g, names, 2000))
    # restart files: one block per species; split at 1000 and resumed
    with tempfile.TemporaryDirectory() as td:
        txt = open(os.path.join(MG.FIX, "mix.ctr")).read()
        ic = "100000               # No. of steps between writes to crash file"
        assert ic in txt
        full = txt.replace(ic, "2000                 # No. of steps between writes to crash file")
        import shutil
        shutil.copytree(MG.FIX, os.path.join(td, "f"), copy_function=shutil.copyfile)
        shutil.copytree(os.path.join(GOLD, "potentials"), os.path.join(td, "potentials"),
                        copy_function=shutil.copyfile)
        open(os.path.join(td, "f", "full.ctr"), "w").write(full)
        MG.run([os.path.join(td, "f", "full.ctr")], td)
        want = open(os.path.join(td, "temp.res.1")).read().split("\n")
        rp = MG.read_restart_rp(os.path.join(td, "temp.res.1"))
        assert sorted(rp) == names and rp["Bent"].shape[1] == 3 and rp["Neutral"].shape[1] == 1
        half = full.replace("2000                 # No. of iterations", "1000                 # No. of iterations")
        half = half.replace("2000                 # No. of steps between writes to crash file",
                            "1000                 # No. of steps between writes to crash file")
        assert half != full
        open(os.path.join(td, "f", "half.ctr"), "w").write(half)
        MG.run([os.path.join(td, "f", "half.ctr")], td)
        src = "------ Configuration Initialization -------------------------------------\n" \
              "Bent\nRESTARTFILE ../temp.res.1\nNeutral\nRESTARTFILE ../temp.res.1\n"
        res = half.replace("temp.res             # Restart File to write to",
                           "temp2.res            # Restart File to write to").rstrip("\n") + "\n" + src
        open(os.path.join(td, "f", "res.ctr"), "w").write(res)
        out = MG.run([os.path.join(td, "f", "res.ctr"), "resume"], td)
        assert out.count("started from") == 2
        got = open(os.path.join(td, "temp2.res.1")).read().split("\n")
        assert got[1].strip() == "1000" and want[1].strip() == "2000"
        assert got[:1] + got[2:] == want[:1] + want[2:]
