"""Host mirror of music_mapmaker (drivers/music_mapmaker.F90:6-153,
mapmaker.F90:69-598): reads the ctrlfile, sets up the framework and
forcefield on the device and builds / writes each requested map through
libmusic_hip."""
import os
import time

import numpy as np

from . import constants as C
from . import hip as H
from . import inputs


def map_filename(mapsorb, probe, model):
    """mapmaker_makeName (mapmaker.F90:604-625)."""
    name = mapsorb
    if probe.upper() != "PROBE":
        name += "." + probe
    return name + "." + model + ".map"


def setup_framework(ctx, system, mapsorb):
    fw = system["mols"][mapsorb]
    ctx.upload_framework(fw["full_xyz"], fw["full_type"], fw["full_q"], system["fund_ell"],
                         system["full_ell"])
    # cell angles and the effective box the grid spans (config_makegrid,
    # config.F90:762-766, on the fundamental cell's eff)
    ctx.set_cell(system["angles"], system["eff"])
    return fw


def header_zatoms(system, mapsorb, probe_type, lj=True):
    """Map header zatom records (mapmaker_hackheader, mapmaker.F90:372-388):
    one per atom type of the map species; charge is that of the i-th atom
    of the molecule (molecules_getcharge(mapSorb, i), a reference quirk)."""
    fw = system["mols"][mapsorb]
    types = []
    for t in fw["type"]:
        if t not in types:
            types.append(int(t))
    types.sort()
    out = []
    for i, t in enumerate(types):
        at = system["atypes"][t]
        d = dict(name=at["name"], symbol=at.get("symbol", ""),
                 charge=float(fw["q"][i]) if i < len(fw["q"]) else 0.0)
        pair = system["aa"].get((at["name"], system["atypes"][probe_type]["name"]))
        if lj and pair is not None and pair.get("model") == "LJ" and not pair["off"]:
            d.update(A12=pair["A"], B6=pair["B"], rcuthi=pair["hicut"], rcutlo=pair["locut"])
        out.append(d)
    return out


def make_maps(ctrl_path, ctx=None, precision=H.MH_PREC_FP64, outdir=None, write=True,
              sign_mode=H.MH_EWALD_PARITY):
    """Build every map of the ctrlfile's Mapmaker section (Ewald maps with the
    reference's reciprocal-space signs unless sign_mode is MH_EWALD_PHYSICS).
    Returns a list of (job, Map, seconds)."""
    system = inputs.load_system(ctrl_path)
    own = ctx is None
    ctx = ctx or H.Context(0)
    results = []
    for job in system["ctrl"]["mapmaker"]:
        probe = system["mols"][job["probe"]]
        if len(probe["type"]) != 1:
            raise ValueError("config_makegrid: probe must be a single atom (config.F90:754-759)")
        ptype = int(probe["type"][0])
        setup_framework(ctx, system, job["mapsorb"])
        t0 = time.time()
        if job["model"].upper() == "LJ":
            A, B, h2, l2, on = inputs.lj_table_for_probe(system, ptype)
            ctx.set_lj(A, B, h2, l2, on)
            m = ctx.build_map(H.MH_MAP_LJ, precision, job["gridspace"], job["hipotcut"])
            zat = header_zatoms(system, job["mapsorb"], ptype)
        elif job["model"].upper() == "EWALD":
            e = system["ss"][(job["probe"], job["mapsorb"], "COUL")]["ewald"]
            ctx.set_ewald(e["kappa"], e["kmax"], e["locut"], sign_mode)
            m = ctx.build_map(H.MH_MAP_EWALD, precision, job["gridspace"], job["hipotcut"],
                              probe_q=float(probe["q"][0]))
            zat = header_zatoms(system, job["mapsorb"], ptype, lj=False)
        else:
            raise ValueError("unsupported map model " + job["model"])
        dt = time.time() - t0
        if write:
            fname = job["filename"]
            if fname.upper() == "AUTO":
                fname = map_filename(job["mapsorb"], job["probe"], job["model"].upper())
            path = os.path.join(outdir or os.getcwd(), fname)
            m.write(path, zat, job["hipotcut"])
            job = dict(job, path=path)
        results.append((job, m, dt))
    if own:
        pass
    return results
