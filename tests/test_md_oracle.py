"""Plain NVE MD (the reference's MD driver with velocity Verlet) on the
CPU: the oracle's trajectory (mo_hmc_md) fed through the host statistics
and display of musichip.md reproduces every number the reference prints
for its test001 (2 CH4, 20 000 steps of 0.1 fs from add.res.1, LJ 9 A):
the instantaneous, cumulative, block and std columns of all 200 MD step
blocks, the COM velocity and the summary."""
import io
import json
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from musichip import md as M

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, GOLD)
from extract_md_reference import parse  # noqa: E402

CTR = os.path.join(GOLD, "test001-md-ch4", "test001.ctr")


@pytest.fixture(scope="module")
def reference():
    return json.load(open(os.path.join(GOLD, "md_reference.json")))


def test_md_setup_reads_test001():
    su = M.MdSetup(CTR)
    assert su.spc == "CH4" and su.nmoles == 2 and su.mv["dt"] == 0.0001
    assert np.allclose(su.rp.reshape(2, 3), [[9.82655, 11.47027, 10.29023], [5.68420, 11.65575, 8.76603]])
    assert np.allclose(su.v.reshape(2, 3), [[1.0, 0.0, 0.1], [0.0, 0.5, 0.0]])
    A, B, h2, l2, on = su.spec["lj"]
    assert h2[0] == pytest.approx(81.0) and l2[0] == pytest.approx(0.09) and on[0] == 1


def test_stat_block_and_std():
    """stats_update: the block average restarts after blocksize updates;
    std from running averages."""
    st = M.Stat(blocksize=3)
    for x in (1.0, 2.0, 3.0, 10.0):
        st.update(x)
    assert st.blockavg == 10.0 and st.cumavg == 4.0
    assert st.std() == pytest.approx(np.std([1.0, 2.0, 3.0, 10.0]))


def test_oracle_md_reproduces_test001_output(reference):
    su = M.MdSetup(CTR)
    o = O.Hmc(dict(su.spec), 1)
    o.set_config(su.rp.reshape(su.nmoles, 1, 3))
    done, tr, rp, r, v = o.md(su.v.reshape(su.nmoles, 1, 3), 20000)
    assert done == 20000
    st = M.MdStats(su)
    out = io.StringIO()
    for k in range(200):
        st.feed(tr[100 * k:100 * (k + 1)])
        st.display(out)
    st.summary(out)
    steps, summ = parse(out.getvalue())
    ref = reference["steps"]
    assert len(steps) == len(ref) == 200
    for a, b in zip(steps, ref):
        assert a == b, (a["step"], a, b)
    assert summ == reference["summary"]
