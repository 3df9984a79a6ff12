"""Small flexible-chain systems for the HMC tests: a 16 A framework (96 LJ
atoms, seed 61), maps for the chain site types built with the CPU oracle or
on the device, and chains placed where every site sits in a tabulated cell
with negative framework energy."""
import numpy as np

from oracle import oracle as O
from musichip import hmc as HM
from musichip import synth

L16 = np.array([16.0, 16.0, 16.0])
GRID = 0.4
HICUT = 7.9


def framework(seed=61):
    rng = np.random.default_rng(seed)
    xyz = synth.lattice_faces(2, 8.0, rng)
    eps, sig, typ = synth.fw_types(rng, len(xyz), 4)
    return dict(xyz=xyz, type=typ, eps=eps, sig=sig)


def propane_ua():
    """CH3-CH2-CH3 cut from the butane model (3 sites: odd Gaussian counts)."""
    m = synth.butane_ua()
    return dict(name="propane", types=m["types"], stype=[0, 1, 0], mass=[15.035, 14.027, 15.035], lj=m["lj"],
                bonds=m["bonds"][:2], angles=m["angles"][:1], torsions=[],
                chain=m["chain"][:3] - m["chain"][:3].mean(0))


def oracle_maps(fw, model):
    """Per site type: (npnts, step, ptr, vals8) built with the oracle."""
    npn = (L16 / GRID).astype(int) + 1
    step = L16 / (npn - 1)
    out = []
    for t in model["types"]:
        e, s = model["lj"][t]
        A, B, h2, l2 = synth.lb_table(fw["eps"], fw["sig"], e, s, HICUT)
        ref = O.build_nodes(1, dict(xyz=fw["xyz"], type=fw["type"], A=A, B=B, hicut2=h2, locut2=l2,
                                    full_ell=L16), npn, step)
        ptr, src, _ = O.compact(npn, ref[:, 0].astype(np.float32), 100.0 / 4.184)
        out.append((npn, step, ptr, ref.astype(np.float32)[src].T.copy()))
    return out


def spec_with_maps(model, maps, dt=0.001, nmd=10):
    sp = HM.hmc_spec(model, L16, L16, hicut=HICUT, dt=dt, nmd=nmd)
    sp["maps"] = maps
    return sp


def place(model, maps, n, seed):
    rng = np.random.default_rng(seed)

    def ok(xyz):
        good = np.ones(len(xyz), dtype=bool)
        for (npn, step, ptr, vals) in maps:
            g, v = O.pothermite(npn, step, ptr, vals, xyz, fund_ell=L16)
            good &= g & (v < 0.0)
        return good

    return synth.place_chains(model, n, L16, rng, ok, min_dist=3.2)
