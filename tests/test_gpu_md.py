"""Plain NVE MD on the device (mh_hmc_md, musichip.md): the reference's
test001 (2 CH4, 20 000 velocity-Verlet steps of 0.1 fs from add.res.1)
run end to end on the device reproduces every number the reference prints
(200 MD step blocks and the summary), and the per-step trace of a flexible
butane system with maps and bonded terms (FP64) equals the oracle's to
1e-9 relative; with FP32 pair/map arithmetic the energies stay within 1e-5
and the total momentum within 1e-4."""
import io
import json
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from musichip import hip as H
from musichip import hmc as HM
from musichip import md as M
from musichip import synth
import hmc_systems as HS

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, GOLD)
from extract_md_reference import parse  # noqa: E402


def test_device_md_reproduces_test001_output(hip):
    print("case: test001 MD on the device, FP64, 20000 steps", flush=True)
    run = M.MdRun(os.path.join(GOLD, "test001-md-ch4", "test001.ctr"), ctx=hip)
    out = io.StringIO()
    run.run(out)
    run.close()
    steps, summ = parse(out.getvalue())
    ref = json.load(open(os.path.join(GOLD, "md_reference.json")))
    assert len(steps) == len(ref["steps"]) == 200
    for a, b in zip(steps, ref["steps"]):
        assert a == b, (a["step"], a, b)
    assert summ == ref["summary"]


def _butane(hip, prec):
    fw = HS.framework()
    model = synth.butane_ua()
    hip.upload_framework(fw["xyz"], fw["type"], np.zeros(len(fw["xyz"])), HS.L16, HS.L16)
    dev, host = {}, []
    for t, name in enumerate(model["types"]):
        e, s = model["lj"][name]
        hip.set_lj(*synth.lb_table(fw["eps"], fw["sig"], e, s, HS.HICUT))
        m = hip.build_map(H.MH_MAP_LJ, H.MH_PREC_FP64, HS.GRID, 100.0)
        ptr, vals = m.get()
        dev[t] = m
        host.append((m.npnts, m.step, ptr, vals))
    sp = HS.spec_with_maps(model, host, dt=0.001)
    rp = HS.place(model, host, 7, seed=41)
    v = np.random.default_rng(42).normal(0.0, 3.0, rp.shape)
    return sp, dev, rp, v


@pytest.mark.parametrize("prec", [H.MH_PREC_FP64, H.MH_PREC_FP32])
def test_device_md_trace_matches_oracle(hip, prec):
    print("case: butane MD trace vs oracle, precision %d" % prec, flush=True)
    sp, dev, rp, v = _butane(hip, prec)
    b = HM.HmcBatch(hip, sp, dev, [1, 2], maxN=8, precision=prec)
    for r in range(2):
        b.set_config(r, rp)
        b.set_velocities(r, v)
    nsteps = 400
    tr = b.md(nsteps)
    o = O.Hmc(sp, 1)
    o.set_config(rp)
    done, ref, orp, orr, ov = o.md(v, nsteps)
    assert done == nsteps
    # FP32 pair forces are not exactly antisymmetric, so the total momentum
    # (columns 7-9) drifts by ~1e-5 of its size over the run
    tol = np.full(10, 1e-9) if prec == H.MH_PREC_FP64 else np.array([1e-5] * 7 + [1e-4] * 3)
    scale = np.maximum(np.abs(ref), 1.0)
    for r in range(2):
        err = (np.abs(tr[r] - ref) / scale).max(axis=0)
        assert np.all(err < tol), (r, err)
    # bonded and map parts are exercised
    assert np.all(np.abs(ref[:, 0]) > 0) and np.all(ref[:, 2:5] > 0)
    if prec == H.MH_PREC_FP64:
        prp, pr = b.config(0)
        assert np.abs(prp - orp).max() < 1e-9
        assert np.abs(b.velocities(0) - ov).max() < 1e-9
    b.close()
