"""Molecular dynamics on the device: the reference's MD runs (music.F90
with an "MD Information" ctrlfile section; md_initParams md.F90:83-202,
md_doSim :592-623, md_simdisplay :881-1123, md_simsummary :1128-1320) for
the NVE velocity-Verlet integrator (velverlet_nveIntegrate,
velverlet.F90:148-288), with velocities taken from a restart file.

The trajectory runs in mh_hmc_md (one workgroup per replica, per-step
energy parts and kinetic sums returned to the host); this module keeps the
reference's running statistics (stats_update, stats.F90:152-195, block size
STATS_BLOCKSIZE = 10000, defaults.F90:138) and prints its MD step blocks.

Offered: species of single-atom molecules moving in the cell (sorbate-only
systems, LJ from the atom_atom_file), NVE velocity Verlet.  Generated
velocities, thermostats, other integrators, fixed (framework) species and
multi-atom molecules are not offered and raise NotImplementedError."""
import math
import os
import sys

import numpy as np

from . import constants as C
from . import hip as H
from . import hmc as HM
from . import inputs as I
from .stats import STATS_BLOCKSIZE, Stat  # noqa: F401

SCALEKE = 0.01             # (A/ps)^2 amu -> kJ/mol, defaults.F90:76


def _time_units(t):
    """md_simdisplay's time label (md.F90:903-913); its bounds are default
    (single precision) reals, so 0.1 ps is still shown in fs."""
    if t < float(np.float32(0.1)):
        return t * 1000.0, "fs"
    if t > 1000.0:
        return t / 1000.0, "ns"
    return t, "ps"


class MdSetup:
    """The system of an MD ctrlfile: species, start configuration and
    velocities, the HMC-form spec of the forcefield (no device work)."""

    def __init__(self, ctrl_path):
        self.ctrl_path = ctrl_path
        self.sysd = I.load_system(ctrl_path)
        c = self.sysd["ctrl"]
        self.general = c["general"]
        self.mdsec = I.read_md_section(ctrl_path)
        if len(self.mdsec["moves"]) != 1:
            raise NotImplementedError("one MD move type is offered")
        mv = self.mdsec["moves"][0]
        if mv["kind"] != "INTEGRATE" or mv["ensemble"] != "NVE" or mv["integrator"] != "VELOCITYVERLET":
            raise NotImplementedError("MD: only INTEGRATE / NVE / VelocityVerlet is offered (got %s %s %s)" %
                                      (mv["kind"], mv["ensemble"], mv["integrator"]))
        if mv["nsteps"] != 1:
            raise NotImplementedError("MD: one integration step per move is offered")
        self.mv = mv
        init = I.read_config_init(ctrl_path)
        moving = [n for n, src in init.items() if src and src[0].upper() != "FIXED"]
        fixed = [n for n, src in init.items() if src and src[0].upper() == "FIXED"]
        if fixed:
            raise NotImplementedError("MD: fixed (framework) species are not offered: %s" % fixed)
        if len(moving) != 1:
            raise NotImplementedError("MD: one moving species is offered")
        self.spc = moving[0]
        src = init[self.spc]
        if src[0].upper() != "RESTARTFILE":
            raise NotImplementedError("MD: configurations from a RESTARTFILE are offered")
        base = c["base"]
        rst = I.read_restartfile(os.path.join(base, src[1]))
        if self.mdsec["velocity_source"].upper() == "GENERATE":
            raise NotImplementedError("MD: generated velocities are not offered; give a restart file")
        vel = I.read_restartfile(os.path.join(base, self.mdsec["velocity_source"]))
        mol = self.sysd["mols"][self.spc]
        if len(mol["type"]) != 1:
            raise NotImplementedError("MD: single-atom molecules are offered")
        sp = rst["species"][self.spc]
        self.rp = sp["rp"].copy()
        self.v = vel["species"][self.spc]["v"]
        if self.v is None:
            raise ValueError("%s holds no _Velocities_ section" % self.mdsec["velocity_source"])
        self.nmoles = len(self.rp)
        atypes = self.sysd["atypes"]
        t0 = int(mol["type"][0])
        aname = atypes[t0]["name"]
        self.mass = float(atypes[t0]["mass"])
        self.mmass = float(mol["mass"])
        lj = self.sysd["aa"].get((aname, aname))
        if lj is None or lj["model"] != "LJ":
            raise NotImplementedError("MD: LJ between the sorbate atoms is required")
        on = 0 if lj["off"] else 1
        A, B = (lj["A"], lj["B"]) if on else (0.0, 0.0)
        full = self.sysd["full_ell"]
        self.volume = float(np.prod(full))
        self.dof = 3  # a single atom's translational degrees of freedom (molecules_getdof)
        self.spec = dict(stype=[0], mass=[self.mass], site_map=[-1], maps=[], full_ell=list(full),
                         fund_ell=list(self.sysd["fund_ell"]), nst=1,
                         lj=([A], [B], [lj["hicut2"]], [lj["locut2"]], [on]), bonds=[], angles=[], torsions=[],
                         T=mv["T"], dt=mv["dt"], nmd=1, max_nrg_md=1000.0, max_nrg0=1.0e6)


class MdStats:
    """The reference's per-species statistics of an MD run (storestats,
    storestats.F90:56-120) fed one trace row per step, and its printed
    blocks."""

    def __init__(self, setup):
        self.setup = setup
        self.spc, self.nmoles, self.mmass, self.dof = setup.spc, setup.nmoles, setup.mmass, setup.dof
        self.dt = setup.mv["dt"]
        self.stats = {k: Stat() for k in ("tmole", "tatom", "intra", "ke", "noncoul", "coul")}
        self.step = 0
        self.time = 0.0
        self.comvel = 0.0

    def _update(self, row):
        """storestats updates of one step from a trace row (energies in
        kcal/mol converted to the reference's kJ/mol)."""
        n = self.nmoles
        ke = 0.5 * row[5] * SCALEKE
        tatom = row[5] * SCALEKE / (C.KJMOLE_KB * (self.dof * n))
        tmole = row[6] * self.mmass * SCALEKE / (C.KJMOLE_KB * (3 * n))
        self.stats["tmole"].update(tmole)
        self.stats["tatom"].update(tatom)
        self.stats["intra"].update((row[2] + row[3] + row[4]) * C.CALTOJ)
        self.stats["ke"].update(ke)
        self.stats["noncoul"].update(row[1] * C.CALTOJ)
        self.stats["coul"].update(0.0)
        self.comvel = math.sqrt(row[7] ** 2 + row[8] ** 2 + row[9] ** 2) / (self.mmass * n)

    def feed(self, trace):
        """Statistics of consecutive steps from trace rows [k][MD_NQ]."""
        for row in trace:
            self._update(row)
        self.step += len(trace)
        self.time = self.step * self.dt

    def _rows(self, key, fmt):
        st = self.stats[key]
        vals = st.row() if key in ("tmole", "tatom") else tuple(x / self.nmoles for x in st.row())
        return "".join(fmt % x for x in vals)

    def display(self, out):
        """md_simdisplay (md.F90:881-1123) for the species: Fortran edit
        descriptors restated (A output right-justified in wider fields)."""
        n = self.nmoles
        t, units = _time_units(self.time)
        w = out.write
        w(" \n")
        w(" MD Step: %s    Time %10.3f %-4s\n" % (self.step, t, units))
        w(" --------------------------------------------------\n")
        w(" %s Information\n" % self.spc)
        w("   %16s%15s%15s%15s%15s\n" % ("Variable", "Current", "CumulAvg", "Block", "Std"))
        w("   %-16s%s\n" % ("Tmole          :", self._rows("tmole", "%15.3f")))
        w("   %-16s%s\n" % ("Tatom          :", self._rows("tatom", "%15.3f")))
        w("   %-16s%s\n" % ("Intramolecular :", self._rows("intra", "%15.6f")))
        w("   %-16s%s\n" % ("Kinetic Energy :", self._rows("ke", "%15.6f")))
        w(" Energy for %s-%s\n" % (self.spc, self.spc))
        w("   %-16s%s\n" % ("Noncoulombic   :", self._rows("noncoul", "%15.6f")))
        w("   %-16s%s\n" % ("Coulombic      :", self._rows("coul", "%15.6f")))
        pe = self.stats["intra"].inst + self.stats["noncoul"].inst + self.stats["coul"].inst
        peavg = self.stats["intra"].cumavg + self.stats["noncoul"].cumavg + self.stats["coul"].cumavg
        ke, keavg = self.stats["ke"].inst, self.stats["ke"].cumavg
        pe, peavg, ke, keavg = pe / n, peavg / n, ke / n, keavg / n
        w(" System Energy Totals(per total moles)\n")
        w("   %-16s%15.6f%15.6f\n" % ("Total Potential:", pe, peavg))
        w("   %-16s%15.6f%15.6f\n" % ("Total Kinetic  :", ke, keavg))
        w("   %-16s%15.6f%15.6f\n" % ("Total Energy   :", pe + ke, peavg + keavg))
        w("   COM Velocity (for each species) : SPC-1: %.10f,\n" % self.comvel)
        w(" Total # Moles %12d\n" % n)
        w(" --------------------------------------------------\n")

    def summary(self, out):
        """md_simsummary (md.F90:1128-1320) without the system-energy line
        of interact_totnrg.  'Total Steps' is the driver's loop index after
        its last step, one past the number of steps, as the reference
        prints it."""
        n = self.nmoles
        t, units = _time_units(self.time)
        st = self.stats
        peavg = st["intra"].cumavg / n + st["noncoul"].cumavg / n
        keavg = st["ke"].cumavg / n
        w = out.write

        def tab(label, val, f="%15.3f"):
            return ("   " + label).ljust(34) + (f % val) + " kJ/mol\n"

        w(" \n \n")
        w(" -----------------------------------------------------\n")
        w("%20s%s\n" % ("", "MD SUMMARY"))
        w("%12s%s%10d\n" % ("", "Total Steps: ", self.step + 1))
        w("%12s%s%10.3f %-4s\n" % ("", "Total Time : ", t, units))
        w(" =====================================================\n")
        w("%15s%s\n" % ("", "System Energy Totals"))
        w(tab("Average Potential Energy: ", peavg))
        w(tab("Average Kinetic Energy: ", keavg))
        w(tab("Average Total Energy: ", peavg + keavg))
        w(" =====================================================\n")
        w(" %s Summary Information\n" % self.spc)
        w("   %20s%15.3f K\n" % ("T(mole)      : ", st["tmole"].cumavg))
        w("   %20s%15.3f K\n" % ("T(atom)      : ", st["tatom"].cumavg))
        w("   %20s%15.3f kJ/mol\n" % ("Total Intramolecular Energy : "[:20], st["intra"].cumavg / n))
        w("   %20s%15.3f kJ/mol\n" % ("Total Kinetic Energy : "[:20], keavg))
        w(" =====================================================\n")
        w("%15s%s\n" % ("", "Pairwise Interaction Energies"))
        w(" Energy for %s-%s\n" % (self.spc, self.spc))
        w("   %20s%15.3f kJ/mol (total of %d mols)\n" % ("Noncoul Inter Energy : "[:20], st["noncoul"].cumavg / n, n))
        w("   %20s%15.3f kJ/mol (total of %d mols)\n" % ("Coul Inter Energy : "[:20], st["coul"].cumavg / n, n))
        w(" -----------------------------------------------------\n")

class MdRun:
    """One MD run described by a reference ctrlfile, on the device."""

    def __init__(self, ctrl_path, ctx=None, precision=H.MH_PREC_FP64, threads=None):
        self.setup = su = MdSetup(ctrl_path)
        self.stats = MdStats(su)
        self.ctx = ctx or H.Context(0)
        full = su.sysd["full_ell"]
        self.ctx.upload_framework(np.zeros((0, 3)), np.zeros(0, dtype=np.int32), None, su.sysd["fund_ell"], full)
        self.batch = HM.HmcBatch(self.ctx, su.spec, {}, [int(su.general["iseed"])], maxN=max(1, su.nmoles),
                                 precision=precision, threads=threads)
        self.batch.set_config(0, su.rp.reshape(su.nmoles, 1, 3))
        self.batch.set_velocities(0, su.v.reshape(su.nmoles, 1, 3))

    def advance(self, nsteps):
        """nsteps MD steps on the device; statistics updated per step."""
        tr = self.batch.md(nsteps)[0]
        self.stats.feed(tr)
        return tr

    def run(self, out=sys.stdout):
        """The whole run: niterations steps, a display block every
        'Steps between writes to stdout' steps, then the summary."""
        total = int(self.setup.general["niterations"])
        every = max(1, int(self.setup.general["iprint"]))
        out.write(" Begining Simulation\n Number of Iterations to Perform : %9d\n" % total)
        st = self.stats
        while st.step < total:
            k = min(every - st.step % every, total - st.step)
            self.advance(k)
            if st.step % every == 0:
                st.display(out)
        st.summary(out)

    def close(self):
        self.batch.close()


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="MD run of a reference ctrlfile on the device")
    ap.add_argument("ctrlfile")
    ap.add_argument("--fp32", action="store_true", help="FP32 pair arithmetic (FP64 integration)")
    a = ap.parse_args(argv)
    run = MdRun(a.ctrlfile, precision=H.MH_PREC_FP32 if a.fp32 else H.MH_PREC_FP64)
    run.run()
    run.close()


if __name__ == "__main__":
    main()
