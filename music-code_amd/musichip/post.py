"""The reference's post-processor (drivers/music_post.F90 over post.F90):
energy and loading averages from the binary configuration files of a series
of simulations, and the isotherm with a Langmuir fit.

Post ctrlfile (post_init post.F90:266-417), after the line holding
"Post Processor Information":
    simulation type (GCMC, CBGCMC, HGCMC, MD)
    base name of the configuration files
    first, last file numbers            (files <base>.<first> .. <base>.<last>)
    name of the regenerated ctrlfile    (the run's ctrlfile from the header)
    name of the output file
    start skip, end skip                (percent of configurations)
Optional sections: "Post : Energy Average Info" and "Post : Loading Average
Info", each followed by the number of blocks (post_initAnalTypes
:903-956); "Age of datafile" NEW_DATAFILE.  Other analyses (siting, pore
maps, structure, radial distribution, MSD, velocity histograms,
interactions) and OLD_DATAFILE raise NotImplementedError.

Per file (music_post.F90:70-88): post_reInitStats (:1739-1835, block
size Int((last - first) / nblocks) + 1), post_readandavg (:1865-2087)
updates the stats for configurations firstconf..lastconf, post_writeoutputs
(:3059-3139) samples inst/block/cumulative values at block ends, and
post_writeAvgstoFile (:3209-3549) writes them; after the last file the
isotherm.<molecule> files with post_langmuirFit (:3156-3200) over
mmath_LSFit (mmath.F90:30-69)."""
import math
import sys

import numpy as np

from . import datafile as D
from . import inputs as I
from .stats import Stat

POST_TAG = "Post Processor Information"
NRG_TAG = "Post : Energy Average Info"
LOADING_TAG = "Post : Loading Average Info"
AGE_TAG = "Age of datafile"
OTHER_TAGS = ("Post : Siting Average Info", "Post : Pore Map Info", "Post : Structure Averages Info",
              "MD Post Code Information", "Post : Velocity Histogram", "Post : Interaction Info",
              "Post : Radial Distribution Info")


def _find(lines, tag):
    """filesrchstr (utils.F90:994-1025): first line containing tag, any case."""
    t = tag.strip().upper()
    for i, l in enumerate(lines):
        if t in l.upper():
            return i + 1
    return None


def _e_format(x, w, d):
    """Fortran Ew.d: 0.ddddE+ee, the leading 0 dropped when the field is full."""
    if x == 0.0:
        digits, e = "0" * d, 0
    else:
        s = "%.*e" % (d - 1, abs(x))
        mant, ex = s.split("e")
        digits, e = mant.replace(".", ""), int(ex) + 1
    body = ".%sE%+03d" % (digits, e)
    sign = "-" if x < 0 else ""
    full = sign + "0" + body
    if len(full) <= w:
        return full.rjust(w)
    if len(sign + body) <= w:
        return (sign + body).rjust(w)
    return "*" * w


def real2str(number, sugnchar=12):
    """utils.F90:552-636: a real in about sugnchar characters, fixed point
    when it fits with a significant decimal, else exponential; left
    adjusted."""
    if not math.isfinite(number):
        return "NaN" if number != number else ("Infinity" if number > 0 else "-Infinity")
    want, mindec, minechar = sugnchar, 1, 7
    exp_form = False
    logn = 1.0 if number == 0.0 else math.log10(abs(number))
    xtra = 1 if number < 0 else 0
    beforept = max(0, int(logn)) + 1
    minchar = beforept + xtra + 1 + mindec
    if minchar > want:
        if want < minechar:
            if minchar <= minechar:
                want = minchar
            else:
                exp_form, want = True, minechar
        else:
            exp_form = True
    if not exp_form:
        dec = want - beforept - 1 - xtra
        if dec + logn < 0:
            exp_form, want = True, minechar
    if exp_form:
        s = _e_format(number, want, max(want - 6, mindec))
    else:
        s = "%*.*f" % (want, dec, number)
        if len(s) > want:
            s = "*" * want
    return s.lstrip()


def ls_fit(x, y):
    """mmath_LSFit (mmath.F90:30-69): least-squares line y = m x + c."""
    n = len(x)
    xs, ys = float(np.sum(x)), float(np.sum(y))
    xys = float(np.sum(np.asarray(x) * np.asarray(y)))
    x2 = float(np.sum(np.asarray(x) ** 2))
    if ys == 0.0:
        return 0.0, 0.0
    den = xs * xs - n * x2
    return (xs * ys - n * xys) / den, (xys * xs - x2 * ys) / den


def langmuir_fit(p, z):
    """post_langmuirFit (post.F90:3156-3200): fit 1/z = (K/z_max) (1/p) +
    1/z_max; returns (z_max, K, percent rms error relative to the largest
    loading).  Loadings or pressures below 1e-8 are raised to 1e-8."""
    p = np.maximum(np.asarray(p, dtype=np.float64), 1.0e-8)
    z = np.maximum(np.asarray(z, dtype=np.float64), 1.0e-8)
    m, c = ls_fit(1.0 / p, 1.0 / z)
    with np.errstate(divide="ignore", invalid="ignore"):
        z_max = np.float64(1.0) / np.float64(c)
        K = np.float64(m) / np.float64(c)
        err, mx = 0.0, 1.0e-10
        for pi, zi in zip(p, z):
            err += (z_max * pi / (pi + K) - zi) ** 2
            mx = max(mx, zi)
        err = 100.0 * math.sqrt(err / len(p)) / mx
    return float(z_max), float(K), float(err)


class PostRun:
    """One music_post run over the configuration files a post ctrlfile names."""

    def __init__(self, postfile):
        L = I._lines(postfile)
        i = _find(L, POST_TAG)
        if i is None:
            raise ValueError("%s: no '%s' section" % (postfile, POST_TAG))
        v = [I._strip(x) for x in L[i:i + 6]]
        self.simtype = v[0].upper()
        if self.simtype not in ("GCMC", "CBGCMC", "HGCMC", "MD"):
            raise ValueError("Wrong simType specified in ctrl file: %s" % v[0])
        self.base = v[1]
        fl = [x.strip() for x in v[2].split(",")]
        self.first = int(fl[0])
        self.last = int(fl[1]) if len(fl) > 1 else self.first
        self.newctrl = v[3]
        self.outname = v[4]
        sk = [x.strip() for x in v[5].split(",")]
        self.startskip, self.endskip = float(sk[0]), float(sk[1])
        if not (0 <= self.startskip <= 100 and 0 <= self.endskip <= 100
                and self.startskip + self.endskip <= 100):
            raise ValueError("Wrong value of endskip/startskip specified in control file")
        a = _find(L, AGE_TAG)
        if a is not None:
            age = I._strip(L[a]).split()[0].upper() if I._strip(L[a]) else ""
            if age == "OLD_DATAFILE":
                raise NotImplementedError("configuration files written before 2002 (OLD_DATAFILE)")
            if age != "NEW_DATAFILE":
                raise ValueError("Not able to understand the option :" + age)
        for t in OTHER_TAGS:
            if _find(L, t) is not None:
                raise NotImplementedError("post-processing section '%s'" % t)
        k = _find(L, NRG_TAG)
        self.nrg_blocks = I._ints(L[k])[0] if k is not None else None
        k = _find(L, LOADING_TAG)
        self.load_blocks = I._ints(L[k])[0] if k is not None else None
        self.files = ["%s.%d" % (self.base, n) for n in range(self.first, self.last + 1)]
        # post_extractControlFile: the first file's header gives the ctrlfile
        d0 = D.read(self._path(self.files[0]))
        with open(self._path(self.newctrl), "w") as f:
            for line in d0["ctrl_lines"]:
                f.write(line.rstrip(" \0") + "\n")
        c = I.read_ctrlfile(self._path(self.newctrl))
        self.iconfig = c["general"]["iconfig"]
        self.ndats = c["general"]["niterations"] // self.iconfig
        self.nuc = int(np.prod(c["simcell"]["nxyz"]))
        self.names = [m["name"] for m in c["molecules"]]
        self.firstconf = int(self.ndats * (self.startskip / 100)) + 1
        self.lastconf = self.ndats - int(self.ndats * (self.endskip / 100))
        self.isotherm = {}
        self._out = None

    @staticmethod
    def _path(name):
        """Files are named relative to the working directory, as the
        reference opens them."""
        return name

    def _blocksize(self, nblocks):
        b = int((1.0 * (self.lastconf - self.firstconf)) / nblocks) + 1
        if b == 0:
            raise ValueError("Too many number of blocks specified")
        return b

    def average(self, simno):
        """post_reInitStats + post_readandavg for file simno (1-based):
        returns the sampled rows and the final statistics."""
        name = self.files[simno - 1]
        d = D.read(self._path(name))
        ns = len(self.names)
        fixed = [d["extra"][5 + D.MAX_SORBS + s].split()[0] == "FIXED" for s in range(ns)]
        pressure = float(d["extra"][1].split()[0])
        flags = d["flags"]
        r = dict(name=name, fixed=fixed, pressure=pressure)
        if self.nrg_blocks:
            nb = self._blocksize(self.nrg_blocks)
            r["nrg_bs"] = nb
            r["total"] = Stat(nb)
            r["coul"] = {(a, b): Stat(nb) for a in range(ns) for b in range(a, ns)}
            r["noncoul"] = {(a, b): Stat(nb) for a in range(ns) for b in range(a, ns)}
            r["ke"] = [Stat(nb) for _ in range(ns)]
            r["intra"] = [Stat(nb) for _ in range(ns)]
            r["nrg_rows"] = {}
        if self.load_blocks:
            lb = self._blocksize(self.load_blocks)
            r["load_bs"] = lb
            r["loading"] = [Stat(lb) for _ in range(ns)]
            r["load_rows"] = {}
        cf = d["configs"]
        for i in range(1, self.ndats + 1):
            if i > len(cf):
                raise ValueError("%s holds %d configurations, %d expected" % (name, len(cf), self.ndats))
            if i < self.firstconf:
                continue
            if i > self.lastconf:
                break
            c = cf[i - 1]
            nmol, k = [0] * ns, 0
            for s in range(ns):
                if not fixed[s]:
                    nmol[s] = c["species"][k]["nmoles"]
                    k += 1
            if self.nrg_blocks:
                r["total"].update(c["total"] if flags["tot_nrg"] else 0.0)
                k = 0
                for a in range(ns):
                    for b in range(a, ns):
                        nc, co = c["pair"][k] if flags["pair_nrg"] else (0.0, 0.0)
                        k += 1
                        r["noncoul"][(a, b)].update(nc)
                        r["coul"][(a, b)].update(co)
                    r["ke"][a].update(c["kin_intra"][a][0])
                    r["intra"][a].update(c["kin_intra"][a][1])
            if self.load_blocks:
                for s in range(ns):
                    r["loading"][s].update(float(nmol[s]))
            itn = i * self.iconfig
            j = i - self.firstconf + 1
            if self.nrg_blocks:
                st = r["total"]
                if j % r["nrg_bs"] == 0:
                    r["nrg_rows"][j // r["nrg_bs"]] = (itn, st.inst, st.blockavg, st.cumavg)
                elif i == self.lastconf:
                    r["nrg_rows"][self.nrg_blocks] = (itn, st.inst, st.blockavg, st.cumavg)
            if self.load_blocks and j % r["load_bs"] == 0:
                r["load_rows"][j // r["load_bs"]] = [(itn, st.inst, st.blockavg, st.cumavg)
                                                     for st in r["loading"]]
        return r

    def write(self, simno, r, out):
        """post_writeAvgstoFile for one file; the isotherm files after the last."""
        w = out.write
        ns = len(self.names)
        n = self.lastconf - self.firstconf + 1
        if self.nrg_blocks:
            w("## Total Energy averages , for the ensemble\n")
            w("## configfile : %s\n" % r["name"])
            w("  Iter. No       Inst.(KJ)      Block (KJ)      Cumul(KJ) \n")
            for b in range(1, n // r["nrg_bs"] + 1):
                row = r["nrg_rows"][b]
                w("%10d%16.6f%16.6f%16.6f\n" % (int(row[0]), row[1], row[2], row[3]))
            w("\n")
            if self.load_blocks:
                w(" ## All the averages below are based on cum avg\n")
                w(" ## Unstored/not used nrgs might appear as zero\n")
                w("Sorbs:                                 Nrg-Total        mol/uc    Nrgs(KJ/mol)\n")
                cav = [0.0 if r["fixed"][s] else r["loading"][s].cumavg for s in range(ns)]
                for a in range(ns):
                    for b in range(a, ns):
                        names = (self.names[a] + "--" + self.names[b])[:20].ljust(20)
                        tot = cav[a] if a == b else cav[a] + cav[b]
                        if tot < 0.01e-1:
                            w(names + " -- No Molecs --\n")
                            continue
                        for lab, key in ((" Coulombic : ", "coul"), (" NonCoulom : ", "noncoul")):
                            v = r[key][(a, b)].cumavg
                            w("%s%s%15.5f%15.5f%15.5f\n" % (names, lab, v, tot / self.nuc, v / tot))
                for s in range(ns):
                    nm = self.names[s][:20].ljust(20)
                    tot = cav[s]
                    if tot < 0.01e-1:
                        w(nm + " -- No Molecs --\n")
                        continue
                    for lab, st in (("  Kintetic : ", r["ke"][s]), (" Intra(Tot): ", r["intra"][s])):
                        w("%s%s%15.5f%15.5f%15.5f\n" % (nm, lab, st.cumavg, tot / self.nuc, st.cumavg / tot))
        if self.load_blocks:
            w("\n")
            nblk = n // r["load_bs"]
            for s in range(ns):
                w("## Per Unitcell Loading averages ,  for the ensemble\n")
                w("## configfile : %s\n" % r["name"])
                w("  Iter. No    Inst.(molec/uc)   Block (molec/uc)   Cumul(molec/uc \n")
                w(" Molecule : %s\n" % self.names[s])
                if r["fixed"][s]:
                    w(" FIXED CONFIG, SO NO LOADING AVG REPORTED FOR THIS SORB\n\n")
                    continue
                for b in range(1, nblk + 1):
                    row = r["load_rows"][b][s]
                    w("%10d%16.6f%16.6f%16.6f\n" % (int(row[0]), row[1] / self.nuc, row[2] / self.nuc,
                                                   row[3] / self.nuc))
                self.isotherm[(simno, s)] = (r["pressure"], r["load_rows"][nblk][s][3] / self.nuc)
            if simno == len(self.files):
                self._write_isotherms(r["fixed"])

    def _write_isotherms(self, fixed):
        nsims = len(self.files)
        for s, name in enumerate(self.names):
            with open(self._path("isotherm." + name), "w") as f:
                w = f.write
                w(" #------------OVERALL ISOTHERM-------------\n")
                w("# Molecule : %s\n" % name)
                if fixed[s]:
                    w(" FIXED CONFIG, SO NO LOADING AVG REPORTED FOR THIS SORB\n\n\n")
                    continue
                w(" #  Pressure                  loading   \n")
                w(" #  kpa(1 st gcmc species)    (molec/uc)\n")
                pts = [self.isotherm[(q, s)] for q in range(1, nsims + 1)]
                for p, z in pts:
                    w("%-15s%10s%12.7f\n" % (real2str(p, 15)[:15], "", z))
                if nsims > 1 and self.simtype in ("GCMC", "HGCMC"):
                    z_max, K, err = langmuir_fit([p for p, _ in pts], [z for _, z in pts])
                    w("# Results of a Langmuir Fit, fug= K * z / (z_max -z)\n")
                    w("# Where K is is kPa, z , and z_max  are zeo-loadings in molec/uc\n")
                    w("# K =%-15s ; z_max%16.5f\n" % (real2str(K, 15)[:15], z_max))
                    w("# %% Error in loading %16.5f\n" % err)
                    if err > 5.0:
                        w("# This high error could be from statistically wrong valuesat low pressure values\n")
                        w("# skip them and try again\n")
                w("\n")

    def run(self):
        """All files in order; returns the per-file averages."""
        res = []
        with open(self._path(self.outname), "w") as out:
            for q in range(1, len(self.files) + 1):
                r = self.average(q)
                self.write(q, r, out)
                res.append(r)
        return res


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print("usage: python -m musichip.post <post ctrlfile>", file=sys.stderr)
        return 2
    PostRun(argv[0]).run()
    return 0


if __name__ == "__main__":
    sys.exit(main())
