"""The post-processor (post.F90, drivers/music_post.F90) on configuration
files in the reference's datafile layout: the block logic against plain
means over the configurations it covers, the Langmuir fit against exact
Langmuir data, real2str against hand-derived Fortran edits, and the output
text layout."""
import os

import numpy as np
import pytest

from musichip import datafile as D
from musichip import post as P

HERE = os.path.dirname(os.path.abspath(__file__))
T007 = os.path.join(HERE, "golden", "test007-gcmc-CH4", "test007.ctr")


def test_real2str_fortran_edits():
    # f15.9: 5 digits before the point, 9 after
    assert P.real2str(10000.0, 15) == "10000.000000000"
    # zero: log taken as 1, two digits before the point, f15.12
    assert P.real2str(0.0, 15) == "0.000000000000"
    assert P.real2str(-3.5, 15) == "-3.500000000000"
    # f15.13 keeps one significant digit of 1e-12
    assert P.real2str(1.0e-12, 15) == "0.0000000000010"
    # below that the exponential form e7.1
    assert P.real2str(1.0e-15, 15) == "0.1E-14"
    # too many digits before the point: e15.9
    assert P.real2str(1.5e20, 15) == "0.150000000E+21"
    assert P.real2str(25.0) == "25.000000000"


def test_langmuir_fit_recovers_exact_isotherm():
    zmax, K = 4.0, 50.0
    p = np.array([10.0, 30.0, 100.0, 300.0, 1000.0])
    z = zmax * p / (p + K)
    zm, k, err = P.langmuir_fit(p, z)
    assert abs(zm - zmax) < 1e-9 and abs(k - K) < 1e-7 and err < 1e-9
    # least squares through an exact line, and a scattered set against numpy
    m, c = P.ls_fit(np.array([1.0, 2.0, 3.0]), np.array([5.0, 7.0, 9.0]))
    assert abs(m - 2.0) < 1e-12 and abs(c - 3.0) < 1e-12
    x = np.array([0.3, 1.1, 2.0, 3.7]); y = np.array([1.0, 2.5, 2.9, 5.2])
    mm, cc = np.polyfit(x, y, 1)
    m, c = P.ls_fit(x, y)
    assert abs(m - mm) < 1e-12 and abs(c - cc) < 1e-12


def _ctrl_text():
    txt = open(T007).read()
    return txt.replace("500000         # Steps between writes to config. file",
                       "1000         # Steps between writes to config. file", 1)


def _write_con(path, simno, pressure, N, E):
    lines = _ctrl_text().split("\n")[:-1]
    d = dict(ctrl_filename="test007.ctr", ctrl_lines=lines, simno=simno, siminfo="synthetic",
             flags=dict(tot_nrg=True, pair_nrg=True, intra_nrg=False, posns=True, vels=False, time=False),
             extra=D.extra_info_lines(simno, pressure, 300.0, "GCMC", ["CH4"], [False], 20000, 1000),
             configs=[dict(time=None, pair=[(e, 0.0)], kin_intra=[(0.0, 0.0)], total=e,
                           species=[dict(nmoles=int(n), rp=np.full((int(n), 1, 3), 1.5), v=None)])
                      for n, e in zip(N, E)])
    D.write(path, d)


def test_post_loading_and_energy_averages(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    rng = np.random.default_rng(7)
    press = [100.0, 1000.0, 10000.0]
    data = []
    for q, p in enumerate(press):
        N = rng.integers(0, 30, 20)
        E = -1.7 * N + rng.normal(0, 0.5, 20)
        _write_con("temp.con.%d" % (q + 1), q + 1, p, N, E)
        data.append((N, E))
    open("post.ctr", "w").write(
        "-------- Post Processor Information ---------\n"
        "GCMC                  # simulation type\n"
        "temp.con              # base name of config files\n"
        "1, 3                  # first and last file numbers\n"
        "regen.ctr             # new ctrlfile\n"
        "post.out              # output file\n"
        "10, 10                # start skip, end skip (%)\n"
        "-------- Post : Energy Average Info ---------\n"
        "4                     # blocks\n"
        "-------- Post : Loading Average Info --------\n"
        "4                     # blocks\n")
    run = P.PostRun("post.ctr")
    # 20 configurations, 10 % skipped at each end: configurations 3..18,
    # block size Int(15/4)+1 = 4, four blocks ending at 6, 10, 14, 18
    assert (run.ndats, run.firstconf, run.lastconf, run.nuc) == (20, 3, 18, 1)
    res = run.run()
    # the run's ctrlfile, lines without trailing blanks (datafile_gencontrolfile)
    assert open("regen.ctr").read() == "".join(l.rstrip() + "\n" for l in _ctrl_text().split("\n")[:-1])
    for (N, E), r in zip(data, res):
        for b, end in enumerate((6, 10, 14, 18), 1):
            row = r["load_rows"][b][0]
            assert row[0] == end * 1000 and row[1] == N[end - 1]
            assert abs(row[2] - N[end - 4:end].mean()) < 1e-12
            assert abs(row[3] - N[2:end].mean()) < 1e-12
            erow = r["nrg_rows"][b]
            assert abs(erow[2] - E[end - 4:end].mean()) < 1e-12 and abs(erow[3] - E[2:end].mean()) < 1e-12
        assert abs(r["noncoul"][(0, 0)].cumavg - E[2:18].mean()) < 1e-12
        assert abs(r["loading"][0].std() - N[2:18].std()) < 1e-9
    out = open("post.out").read().split("\n")
    assert out[0] == "## Total Energy averages , for the ensemble" and out[1] == "## configfile : temp.con.1"
    assert out[2] == "  Iter. No       Inst.(KJ)      Block (KJ)      Cumul(KJ) "
    assert out[3] == "%10d%16.6f%16.6f%16.6f" % (6000, data[0][1][5], data[0][1][2:6].mean(),
                                                  data[0][1][2:6].mean())
    iso = open("isotherm.CH4").read().split("\n")
    assert iso[0] == " #------------OVERALL ISOTHERM-------------" and iso[1] == "# Molecule : CH4"
    for q, p in enumerate(press):
        zq = data[q][0][2:18].mean()
        assert iso[4 + q] == "%-15s%10s%12.7f" % (P.real2str(p, 15)[:15], "", zq)
    zm, K, err = P.langmuir_fit(press, [d[0][2:18].mean() for d in data])
    assert iso[7] == "# Results of a Langmuir Fit, fug= K * z / (z_max -z)"
    assert iso[9] == "# K =%-15s ; z_max%16.5f" % (P.real2str(K, 15)[:15], zm)


def test_post_rejects_unbuilt_analyses(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    _write_con("temp.con.1", 1, 10.0, [1] * 20, [0.0] * 20)
    base = ("-- Post Processor Information --\nGCMC\ntemp.con\n1\nregen.ctr\npost.out\n0, 0\n")
    open("a.ctr", "w").write(base + "-- Post : Siting Average Info --\n2\n")
    with pytest.raises(NotImplementedError):
        P.PostRun("a.ctr")
    open("b.ctr", "w").write(base.replace("0, 0", "60, 50"))
    with pytest.raises(ValueError):
        P.PostRun("b.ctr")
