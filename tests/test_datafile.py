"""The reference's binary datafile (datafile.F90), pinned on a file the
reference wrote: tests/test010-hgcmc-butane/test.lib.0 (100 butanes, 10
configurations) is parsed record by record and written back byte for byte,
and the extra-info lines of the current writer (datafile_writeExtraInfo)
reproduce that file's lines up to the labels the current source renamed
(nsorbs -> nspcs, sorb -> spc, NOSORBATE -> NOSPECIES)."""
import os

import numpy as np

from musichip import datafile as D

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "golden", "test010-hgcmc-butane", "test.lib.0")


def test_reference_datafile_parses_and_round_trips(tmp_path):
    d = D.read(LIB)
    assert d["ctrl_filename"].strip() == "test.ctr" and len(d["ctrl_lines"]) == 81 and d["simno"] == 1
    assert d["ctrl_lines"][0].strip() == "# md3001.ctr used wrong inter file"
    assert d["flags"] == dict(tot_nrg=True, pair_nrg=True, intra_nrg=False, posns=True, vels=True, time=True)
    cf = d["configs"]
    assert len(cf) == 10
    for k, c in enumerate(cf):
        # 1000 steps between configurations of 0.1 fs
        assert abs(c["time"] - 0.1 * (k + 1)) < 1e-9
        kin, intra = c["kin_intra"][0]
        assert abs(c["total"] - (sum(c["pair"][0]) + kin + intra)) < 1e-9 * abs(c["total"])
        sp = c["species"][0]
        assert sp["nmoles"] == 100 and sp["rp"].shape == (100, 4, 3) and sp["v"].shape == (100, 4, 3)
        # bonded UA butane: neighbouring sites 1.5 - 1.6 A apart
        bl = np.linalg.norm(np.diff(sp["rp"], axis=1), axis=2)
        assert bl.min() > 1.4 and bl.max() < 1.7
    out = tmp_path / "rt.lib"
    D.write(str(out), d)
    assert out.read_bytes() == open(LIB, "rb").read()


def test_extra_info_lines_match_reference_file():
    d = D.read(LIB)
    new = D.extra_info_lines(1, -1.99673, 0.0, "MD", ["Butane"], [False], 10000, 1000)
    ren = {"# nspcs": "# nsorbs", "# spc": "# sorb", "NOSPECIES": "NOSORBATE"}
    for x, y in zip(new, d["extra"]):
        for k, v in ren.items():
            x = x.replace(k, v)
        assert x.rstrip() == y.rstrip()
