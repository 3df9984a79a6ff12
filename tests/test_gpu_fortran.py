"""The Fortran iso_c_binding layer (music-code_amd/fortran/music_hip_mod.f90)
driven by a Fortran program: test012's LJ map built through bind(C)
interfaces gives the reference's counts and the same map-file bytes as the
C-ABI called from Python."""
import hashlib
import os
import subprocess
import tempfile

import numpy as np
import pytest

from musichip import hip as H
from musichip import inputs

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
