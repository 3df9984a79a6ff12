"""HMC bursts on the device (mh_hmc_*) against the CPU oracle: FP64 energies
and forces within 1e-10, FP64 move sequences identical (acceptance, dH within
1e-8 kcal/mol) for butane and for propane (odd Gaussian counts carry the
second Box-Muller value across moves), FP32 forces against FP64."""
import numpy as np
import pytest

from oracle import oracle as O
from musichip import hip as H
from musichip import hmc as HM
from musichip import synth
import hmc_systems as HS

pytestmark = pytest.mark.gpu


def _device_maps(hip, fw, model):
    hip.upload_framework(fw["xyz"], fw["type"], np.zeros(len(fw["xyz"])), HS.L16, HS.L16)
    dev, host = {}, []
    for t, name in enumerate(model["types"]):
        e, s = model["lj"][name]
        hip.set_lj(*synth.lb_table(fw["eps"], fw["sig"], e, s, HS.HICUT))
        m = hip.build_map(H.MH_MAP_LJ, H.MH_PREC_FP64, HS.GRID, 100.0)
        ptr, vals = m.get()
        dev[t] = m
        host.append((m.npnts, m.step, ptr, vals))
    return dev, host


@pytest.fixture(scope="module")
def sysdata():
    fw = HS.framework()
    return fw


@pytest.mark.parametrize("mname,nmol,dt,nmd", [("butane", 3, 0.004, 10), ("propane", 3, 0.004, 10),
                                               ("butane", 7, 0.003, 40)])
def test_gpu_hmc_fp64_matches_oracle(hip, sysdata, mname, nmol, dt, nmd):
    """4 fs steps: large enough that some moves are rejected (restore path).
    40 steps of 3 fs move sites by more than the neighbour-list half skin
    (0.49 A) within a move, so the lists are rebuilt during the burst; the
    oracle loops over all site pairs."""
    print("case: HMC %s x %d, dt %g x %d, FP64 energy/forces and 30-move trajectory" % (mname, nmol, dt, nmd),
          flush=True)
    model = synth.butane_ua() if mname == "butane" else HS.propane_ua()
    dev, host = _device_maps(hip, sysdata, model)
    sp = HS.spec_with_maps(model, host, dt=dt, nmd=nmd)
    seeds = [101, 202]
    rps = [HS.place(model, host, nmol, seed=s) for s in (11, 12)]
    b = HM.HmcBatch(hip, sp, dev, seeds, maxN=max(8, nmol), precision=H.MH_PREC_FP64)
    ors = []
    for r in range(2):
        b.set_config(r, rps[r])
        o = O.Hmc(sp, seeds[r])
        o.set_config(rps[r])
        ors.append(o)
        eg, Fg, okg = b.energy_forces(r)
        eo, Fo, oko = o.energy_forces()
        assert okg and oko
        assert abs(eg - eo) < 1e-10 * max(1.0, abs(eo))
        assert np.abs(Fg - Fo).max() < 1e-10 * max(1.0, np.abs(Fo).max())
    max_disp = 0.0
    for mv in range(30):
        b.run(1)
        s = b.state()
        for r in range(2):
            before = ors[r].rp[:ors[r].st.N].copy()
            ors[r].move()
            st = ors[r].st
            max_disp = max(max_disp, np.abs(ors[r].rp[:st.N] - before).max())
            assert (s["att"][r], s["acc"][r], s["fail"][r]) == (st.att, st.acc, st.fail), (mv, r)
            if np.isnan(st.last_dH):
                assert np.isnan(s["last_dH"][r])
            else:
                assert abs(s["last_dH"][r] - st.last_dH) < 1e-8, (mv, r, s["last_dH"][r], st.last_dH)
    for r in range(2):
        rp, rr = b.config(r)
        n = ors[r].st.N
        assert np.abs(rp - ors[r].rp[:n]).max() < 1e-8
        assert np.abs(rr - ors[r].r[:n]).max() < 1e-8
    s = b.state()
    print("accepted %s of %s, largest site displacement in a move %.2f A" %
          (list(s["acc"]), list(s["att"]), max_disp), flush=True)
    assert (s["acc"] > 5).all()
    if nmd == 40:
        assert max_disp > 1.0
    b.close()


def test_gpu_hmc_fp32_forces_vs_fp64(hip, sysdata):
    model = synth.butane_ua()
    dev, host = _device_maps(hip, sysdata, model)
    sp = HS.spec_with_maps(model, host)
    rp = HS.place(model, host, 6, seed=21)
    b64 = HM.HmcBatch(hip, sp, dev, [1], maxN=8, precision=H.MH_PREC_FP64)
    b32 = HM.HmcBatch(hip, sp, dev, [1], maxN=8, precision=H.MH_PREC_FP32)
    b64.set_config(0, rp)
    b32.set_config(0, rp)
    e64, F64, _ = b64.energy_forces(0)
    e32, F32, _ = b32.energy_forces(0)
    rel_e = abs(e32 - e64) / max(1.0, abs(e64))
    rel_f = np.abs(F32 - F64).max() / np.abs(F64).max()
    print("FP32 vs FP64: energy rel %.3e, forces rel (to max |F|) %.3e" % (rel_e, rel_f), flush=True)
    assert rel_e < 1e-5 and rel_f < 1e-4
    b64.close()
    b32.close()


@pytest.mark.parametrize("prec", [H.MH_PREC_FP64, H.MH_PREC_FP32])
def test_gpu_hmc_workgroup_sizes_agree(hip, sysdata, prec):
    """The same replicas advanced with 128, 256, 512 and 1024 lanes per
    replica: identical acceptance counts, energies and coordinates (the lane
    count changes only the order of the block sums)."""
    print("case: HMC workgroup sizes, precision %d" % prec, flush=True)
    model = synth.butane_ua()
    dev, host = _device_maps(hip, sysdata, model)
    sp = HS.spec_with_maps(model, host, dt=0.003, nmd=10)
    rps = [HS.place(model, host, 7, seed=s) for s in (11, 12)]
    res = {}
    for thr in (128, 256, 512, 1024):
        b = HM.HmcBatch(hip, sp, dev, [31, 32], maxN=8, precision=prec, threads=thr)
        for r in range(2):
            b.set_config(r, rps[r])
        b.run(12)
        s = b.state()
        res[thr] = (s, [b.config(r) for r in range(2)])
        b.close()
    s0, c0 = res[512]
    assert (s0["acc"] > 3).all()
    for thr, (s, c) in res.items():
        assert list(s["acc"]) == list(s0["acc"]) and list(s["att"]) == list(s0["att"]), thr
        assert np.allclose(s["energy"], s0["energy"], rtol=1e-9, atol=1e-9), thr
        for r in range(2):
            assert np.abs(c[r][0] - c0[r][0]).max() < 1e-8, thr

