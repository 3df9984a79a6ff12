"""Readers for the reference's input files: ctrlfile sections, .atm/.mol
files, the atom-atom and species-species forcefield files.

Formats follow the reference readers:
  ctrlfile sections located by tag string     general.F90:55-113, utils filesrchstr
  atom file  (Atom_Name:, Atom_Symbol:, ...)   atom.F90
  molecule file (Coord_Info: Listed ...)       molecule.F90:830-926 (principal axes)
  simulation cell section                      simcell.F90:77-225
  atom_atom_file LJ lines                      lj.F90:88-262 (lj_init), :900-981 (mix)
  spc_spc_file lines                           ssdriver.F90:81-253
  mapmaker section                             mapmaker.F90:69-312
Only the subset needed by the GCMC energy hot path is read: LJ / EWALD
pair models, rigid molecules; slanted cells for the map path.
"""
import math
import os
import re

import numpy as np

from . import constants as C


def _strip(line):
    return line.split("#", 1)[0].strip()


def _lines(path):
    with open(path, errors="replace") as f:
        return f.read().splitlines()


def find_section(lines, tag):
    """Index of the line after the first line containing tag (filesrchstr)."""
    for i, l in enumerate(lines):
        if tag.lower() in l.lower():
            return i + 1
    raise KeyError("section '%s' not found" % tag)


def _nonblank(lines, i):
    while i < len(lines) and _strip(lines[i]) == "":
        i += 1
    return i


def _ints(s):
    return [int(x) for x in re.split(r"[,\s]+", _strip(s)) if x]


def _floats(s):
    return [float(x) for x in re.split(r"[,\s]+", _strip(s)) if x]


def find_file(name, envvar, base):
    """Resolve a reference input file: $ENVVAR/name, base/name, then the
    shared potentials tree next to base."""
    cands = []
    d = os.environ.get(envvar)
    if d:
        cands.append(os.path.join(d, name))
    cands.append(os.path.join(base, name))
    sub = {"ATOMSDIR": "atoms", "MOLSDIR": "molecules", "PMAPDIR": "pmaps",
           "EMAPDIR": "pmaps"}.get(envvar, "")
    p = base
    for _ in range(3):
        cands.append(os.path.join(p, "potentials", sub, name))
        p = os.path.dirname(p)
    for c in cands:
        if os.path.exists(c):
            return c
    raise FileNotFoundError("%s (looked in %s)" % (name, cands))


# ----------------------------------------------------------------- atoms
def read_atom(path):
    d = {"sscharge": 0.0, "mass": 0.0}
    for l in _lines(path):
        s = _strip(l)
        if ":" not in s:
            continue
        k, v = [x.strip() for x in s.split(":", 1)]
        if k == "Atom_Name":
            d["name"] = v
        elif k == "Atom_Symbol":
            d["symbol"] = v
        elif k == "Atom_Mass":
            d["mass"] = float(v)
        elif k == "Atom_SS_Charge":
            d["sscharge"] = float(v)
    return d


# ------------------------------------------------------------- molecules
def read_molecule(path):
    """Rigid 'Listed Cartesian' molecule file.  Returns dict with names,
    xyz (as listed), charges, rigid / translate / rotate flags and optional
    fundcell."""
    L = _lines(path)
    mol = {"fundcell": None}
    i = 0
    while i < len(L):
        s = _strip(L[i])
        if s.startswith("Molecule_Name:"):
            mol["name"] = s.split(":", 1)[1].split()[0]
        elif s.startswith("Coord_Info:"):
            info = s.split(":", 1)[1].split()
            mol["coordinfo"] = info
            # molecule.F90:1008-1028: info = 'Listed', coordsys, gcmodel, flag
            flag = info[3].upper() if len(info) > 3 else ""
            mol["rigid"] = len(info) > 2 and info[2].upper() == "RIGID"
            mol["translate"] = flag != "NOTRANSFORM"
            mol["rotate"] = flag not in ("NOROTATE", "NOTRANSFORM")
            mol["transform"] = mol["translate"]
            i += 1
            n = _ints(L[i])[0]
            names, xyz, q = [], [], []
            for j in range(n):
                f = _strip(L[i + 1 + j]).split()
                xyz.append([float(f[1]), float(f[2]), float(f[3])])
                names.append(f[4])
                q.append(float(f[5]) if len(f) > 5 else 0.0)
            mol["atoms"] = names
            mol["xyz"] = np.array(xyz, dtype=np.float64)
            mol["q"] = np.array(q, dtype=np.float64)
            i += n
        elif s.startswith("Fundcell_Info:"):
            ell = _floats(L[i + 1])
            ang = _floats(L[i + 2])
            org = _floats(L[i + 3])
            eff = _floats(L[i + 4])
            mol["fundcell"] = dict(ell=np.array(ell), angles=np.array(ang),
                                   origin=np.array(org), eff=np.array(eff))
            i += 4
        i += 1
    return mol


def _pythag(a, b):
    """sqrt(a^2 + b^2) without destructive underflow or overflow: the
    iteration of EISPACK's pythag (Moler and Morrison, IBM J. Res. Dev. 27
    (1983) 577), which tql2 uses."""
    p = max(abs(a), abs(b))
    if p == 0.0:
        return p
    r = (min(abs(a), abs(b)) / p) ** 2
    while True:
        t = 4.0 + r
        if t == 4.0:
            return p
        s = r / t
        u = 1.0 + 2.0 * s
        p = u * p
        r = (s / u) ** 2 * r


def _tred2(a):
    """Householder reduction of a symmetric matrix to tridiagonal form with
    the transformations accumulated (EISPACK tred2; Martin, Reinsch and
    Wilkinson, Numer. Math. 11 (1968) 181).  Reads the lower triangle
    a[j][i], j >= i.  Returns the diagonal d, the subdiagonal e (e[0] = 0)
    and the orthogonal z.  The operations follow the published algorithm
    step for step, so that the eigenvectors tql2 derives from z carry the
    same signs as those of the reference (molecule.F90:896-897)."""
    n = a.shape[0]
    z = np.zeros((n, n))
    d = np.zeros(n)
    e = np.zeros(n)
    for i in range(n):
        for j in range(i, n):
            z[j, i] = a[j, i]
        d[i] = a[n - 1, i]
    if n > 1:
        for i in range(n - 1, 0, -1):
            l = i - 1
            h = 0.0
            scale = 0.0
            if l >= 1:
                for k in range(l + 1):
                    scale = scale + abs(d[k])
            if l < 1 or scale == 0.0:
                e[i] = d[l]
                for j in range(l + 1):
                    d[j] = z[l, j]
                    z[i, j] = 0.0
                    z[j, i] = 0.0
            else:
                for k in range(l + 1):
                    d[k] = d[k] / scale
                    h = h + d[k] * d[k]
                f = d[l]
                g = -math.copysign(math.sqrt(h), f)
                e[i] = scale * g
                h = h - f * g
                d[l] = f - g
                for j in range(l + 1):
                    e[j] = 0.0
                for j in range(l + 1):
                    f = d[j]
                    z[j, i] = f
                    g = e[j] + z[j, j] * f
                    for k in range(j + 1, l + 1):
                        g = g + z[k, j] * d[k]
                        e[k] = e[k] + z[k, j] * f
                    e[j] = g
                f = 0.0
                for j in range(l + 1):
                    e[j] = e[j] / h
                    f = f + e[j] * d[j]
                hh = f / (h + h)
                for j in range(l + 1):
                    e[j] = e[j] - hh * d[j]
                for j in range(l + 1):
                    f = d[j]
                    g = e[j]
                    for k in range(j, l + 1):
                        z[k, j] = z[k, j] - f * e[k] - g * d[k]
                    d[j] = z[l, j]
                    z[i, j] = 0.0
            d[i] = h
        for i in range(1, n):
            l = i - 1
            z[n - 1, l] = z[l, l]
            z[l, l] = 1.0
            h = d[i]
            if h != 0.0:
                for k in range(l + 1):
                    d[k] = z[k, i] / h
                for j in range(l + 1):
                    g = 0.0
                    for k in range(l + 1):
                        g = g + z[k, i] * z[k, j]
                    for k in range(l + 1):
                        z[k, j] = z[k, j] - g * d[k]
            for k in range(l + 1):
                z[k, i] = 0.0
    for i in range(n):
        d[i] = z[n - 1, i]
        z[n - 1, i] = 0.0
    z[n - 1, n - 1] = 1.0
    e[0] = 0.0
    return d, e, z


def _tql2(d, e, z):
    """Eigenvalues and eigenvectors of the tridiagonal (d, e) by the implicit
    QL method, z accumulating the rotations, then sorted ascending (EISPACK
    tql2; Bowdler, Martin, Reinsch and Wilkinson, Numer. Math. 11 (1968)
    293).  In place; raises after 30 iterations on one eigenvalue, where the
    reference stops (molecule.F90:898-902)."""
    n = len(d)
    if n == 1:
        return
    for i in range(1, n):
        e[i - 1] = e[i]
    f = 0.0
    tst1 = 0.0
    e[n - 1] = 0.0
    for l in range(n):
        j = 0
        h = abs(d[l]) + abs(e[l])
        if tst1 < h:
            tst1 = h
        m = l
        while m < n:
            tst2 = tst1 + abs(e[m])
            if tst2 == tst1:
                break
            m += 1
        if m != l:
            while True:
                if j == 30:
                    raise ValueError("tql2: no convergence for eigenvalue %d" % (l + 1))
                j += 1
                l1 = l + 1
                l2 = l1 + 1
                g = d[l]
                p = (d[l1] - g) / (2.0 * e[l])
                r = _pythag(p, 1.0)
                d[l] = e[l] / (p + math.copysign(r, p))
                d[l1] = e[l] * (p + math.copysign(r, p))
                dl1 = d[l1]
                h = g - d[l]
                for i in range(l2, n):
                    d[i] = d[i] - h
                f = f + h
                p = d[m]
                c = 1.0
                c2 = c
                el1 = e[l1]
                s = 0.0
                s2 = 0.0
                c3 = 0.0
                for ii in range(1, m - l + 1):
                    c3 = c2
                    c2 = c
                    s2 = s
                    i = m - ii
                    g = c * e[i]
                    h = c * p
                    r = _pythag(p, e[i])
                    e[i + 1] = s * r
                    s = e[i] / r
                    c = p / r
                    p = c * d[i] - s * g
                    d[i + 1] = h + s * (c * g + s * d[i])
                    for k in range(n):
                        h = z[k, i + 1]
                        z[k, i + 1] = s * z[k, i] + c * h
                        z[k, i] = c * z[k, i] - s * h
                p = -s * s2 * c3 * el1 * e[l] / dl1
                e[l] = s * p
                d[l] = c * p
                tst2 = tst1 + abs(e[l])
                if not tst2 > tst1:
                    break
        d[l] = d[l] + f
    for ii in range(1, n):
        i = ii - 1
        k = i
        p = d[i]
        for j in range(ii, n):
            if d[j] < p:
                k = j
                p = d[j]
        if k != i:
            d[k] = d[i]
            d[i] = p
            for j in range(n):
                p = z[j, i]
                z[j, i] = z[j, k]
                z[j, k] = p


def principal_axes(xyz, masses, translate=True, rotate=True, moments=False):
    """molecule_genprincipleaxes (molecule.F90:830-926), in the reference's
    arithmetic order: shift by the centre of mass of the real masses
    (molecule_getrefcom, molecule.F90:777-798: sum of r m, times the
    reciprocal of the mass sum), form the inertia tensor with the masses
    truncated to integers (molecule.F90:833, 872-892), diagonalise it by
    tred2/tql2 and rotate the atoms by the transposed eigenvector matrix
    (rot_xform * r, matrix_m_mult_v, matrix.F90:137-150).  The principal
    moments come back too when moments is set."""
    xyz = np.asarray(xyz, dtype=np.float64)
    mreal = [float(x) for x in masses]
    n = len(xyz)
    r = np.array(xyz, dtype=np.float64)
    if translate:
        com = np.zeros(3)
        msum = 0.0
        for a in range(n):
            msum = msum + mreal[a]
            for c in range(3):
                com[c] = com[c] + xyz[a][c] * mreal[a]
        inv = 1.0 / msum
        for c in range(3):
            com[c] = com[c] * inv
        for a in range(n):
            for c in range(3):
                r[a, c] = r[a, c] + (-com[c])
    mint = [float(int(x)) for x in mreal]
    I = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            moi = 0.0
            for a in range(n):
                if i == j:
                    rsq = r[a, 0] * r[a, 0] + r[a, 1] * r[a, 1] + r[a, 2] * r[a, 2]
                    moi = moi + mint[a] * (rsq - r[a, i] * r[a, i])
                else:
                    moi = moi - mint[a] * r[a, i] * r[a, j]
            I[i, j] = moi
    d, e, z = _tred2(I)
    _tql2(d, e, z)
    out = r
    if rotate:
        out = np.zeros_like(r)
        for a in range(n):
            for k in range(3):
                v = 0.0
                for c in range(3):
                    v = v + z[c, k] * r[a, c]
                out[a, k] = v
    return (out, d.copy()) if moments else out


# ------------------------------------------------------------- ctrlfile
def read_ctrlfile(path):
    """Parse the ctrlfile sections used by music_mapmaker / music_gcmc."""
    L = _lines(path)
    base = os.path.dirname(os.path.abspath(path))
    c = {"path": path, "base": base}
    i = find_section(L, "General Information")
    g = [l for l in L[i:i + 12]]
    vals = [_strip(x) for x in g]
    c["general"] = dict(desc=vals[0], niterations=int(vals[1].split()[0]),
                        iprint=int(vals[2].split()[0]), icrash=int(vals[3].split()[0]),
                        iconfig=int(vals[4].split()[0]), simstart=int(vals[5].split()[0]),
                        iseed=int(vals[6].split()[0]))
    # atom types
    i = find_section(L, "Atomic Types")
    n = _ints(L[i])[0]
    i += 1
    atoms = []
    for _ in range(n):
        i = _nonblank(L, i)
        name = _strip(L[i]).split()[0]
        fname = _strip(L[i + 1]).split()[0]
        atoms.append(dict(name=name, file=fname))
        i += 2
    c["atoms"] = atoms
    i = find_section(L, "Molecule Types")
    n = _ints(L[i])[0]
    i += 1
    mols = []
    for _ in range(n):
        i = _nonblank(L, i)
        name = _strip(L[i]).split()[0]
        fname = _strip(L[i + 1]).split()[0]
        mols.append(dict(name=name, file=fname))
        i += 2
    c["molecules"] = mols
    # simulation cell
    i = find_section(L, "Simulation Cell Information")
    first = _strip(L[i]).split()[0]
    if first.upper() == "NULL":
        ell = _floats(L[i + 1])
        ang = _floats(L[i + 2])
        org = _floats(L[i + 3])
        eff = _floats(L[i + 4])
        nxyz = _ints(L[i + 5])
        pbc = _ints(L[i + 6])
        c["simcell"] = dict(fundcell_source=None,
                            fund=dict(ell=np.array(ell), angles=np.array(ang),
                                      origin=np.array(org), eff=np.array(eff)),
                            nxyz=nxyz, pbc=pbc)
    else:
        nxyz = _ints(L[i + 1])
        pbc = _ints(L[i + 2])
        c["simcell"] = dict(fundcell_source=first, fund=None, nxyz=nxyz, pbc=pbc)
    # forcefield
    i = find_section(L, "Forcefield Information")
    ff = [_strip(x).split()[0] for x in L[i:i + 5]]
    c["forcefield"] = dict(kind=ff[0], storage=ff[1], atom_atom_file=ff[2],
                           spc_spc_file=ff[3], intra_file=ff[4])
    # optional sections
    try:
        i = find_section(L, "Mapmaker Information")
        n = _ints(L[i])[0]
        i += 1
        maps = []
        for _ in range(n):
            i = _nonblank(L, i)
            mapsorb = _strip(L[i]).split()[0]
            probe = _strip(L[i + 1]).split()[0]
            inter = _strip(L[i + 2]).split()
            grid = float(_strip(L[i + 3]).split()[0])
            hicut = float(_strip(L[i + 4]).split()[0])
            fname = _strip(L[i + 5]).split()[0]
            maps.append(dict(mapsorb=mapsorb, probe=probe, int=inter[0], model=inter[1],
                             gridspace=grid, hipotcut=hicut, filename=fname))
            i += 6
        c["mapmaker"] = maps
    except KeyError:
        c["mapmaker"] = []
    try:
        c["gcmc"] = _read_gcmc_section(L)
    except KeyError:
        c["gcmc"] = None
    return c


def _read_gcmc_section(L):
    i = find_section(L, "GCMC Information")
    v = [_strip(x) for x in L[i:i + 6]]
    g = dict(niterations=int(v[0].split()[0]), tk=float(v[1].split()[0]),
             npressures=int(v[3].split()[0]), blocksize=int(v[4].split()[0]),
             nspc=int(v[5].split()[0]))
    i += 6
    spcs = []
    for _ in range(g["nspc"]):
        while i < len(L) and (_strip(L[i]) == "" or set(_strip(L[i])) <= set("-")):
            i += 1
        name = _strip(L[i]).split()[0]
        fug = _floats(L[i + 1].replace("-", " ") if False else L[i + 1])
        smap = _strip(L[i + 2]).split()[0]
        nmoves = _ints(L[i + 3])[0]
        wts = _floats(L[i + 4])[:nmoves]
        i += 5
        moves = []
        for m in range(nmoves):
            mt = _strip(L[i]).split()[0].upper()
            i += 1
            mv = dict(type=mt, weight=wts[m])
            if mt in ("RTRANSLATE", "RROTATE"):
                p = _floats(L[i])
                mv["delta"] = p[0]
                mv["adjust"] = int(p[1]) if len(p) > 1 else 0
                i += 1
            moves.append(mv)
        spcs.append(dict(name=name, fugacity=fug, smap=smap, moves=moves))
    g["species"] = spcs
    return g


def read_md_section(path):
    """The ctrlfile's "MD Information" section (md_initParams, md.F90:83-202;
    moves_init / integrate_init for one INTEGRATE move type): velocity source
    (a restart filename or Generate), steps per move, dt (ps), temperature,
    I/O intervals, ensemble and integrator."""
    L = _lines(path)
    i = find_section(L, "MD Information")
    nmv = _ints(L[i])[0]
    i += 1
    src = _strip(L[i]).split()[0]
    i += 2  # the line after the velocity source is a separator
    mv = []
    for _ in range(nmv):
        kind = _strip(L[i]).split()[0].upper()
        vals = [_strip(L[i + k]).split()[0] for k in range(1, 8)]
        mv.append(dict(kind=kind, nsteps=int(vals[0]), dt=float(vals[1]), T=float(vals[2]),
                       iprint=int(vals[3]), idata=int(vals[4]), ensemble=vals[5].upper(),
                       integrator=vals[6].upper()))
        i += 8
    return dict(velocity_source=src, moves=mv)


def read_config_init(path):
    """The ctrlfile's "Configuration Initialization" section: per molecule
    type its source (RESTARTFILE <file>, FIXED, ...)."""
    L = _lines(path)
    i = find_section(L, "Configuration Initialization")
    out = {}
    while i + 1 < len(L):
        name = _strip(L[i]).split()
        src = _strip(L[i + 1]).split()
        if not name or _strip(L[i]).startswith("-"):
            break
        out[name[0]] = src
        i += 2
    return out


def read_restartfile(path):
    """A restart file in the layout of config_writerestartfile
    (config.F90:2114-2430): description, iteration, time, then per molecule
    type _MOLECULE_NAME_, nmoles, natoms, the principal coordinates
    [nmoles][natoms][3], optional _Velocities_ (A/ps) and the generalized
    coordinates.  Returns dict(desc, iteration, time, species={name: dict(
    rp, v)})."""
    with open(path) as f:
        L = [l.rstrip("\n") for l in f]
    out = dict(desc=L[0].strip(), iteration=int(L[1].split()[0]), time=float(L[2].split()[0].replace("D", "E")),
               species={})
    i = 3
    while i < len(L):
        if not L[i].strip().startswith("_MOLECULE_NAME_:"):
            i += 1
            continue
        name = L[i].split(":", 1)[1].split()[0]
        nmol = int(L[i + 1].split()[0])
        nat = int(L[i + 2].split()[0])
        i += 3
        if L[i].strip() != "_Principal Coordinates Only_":
            raise ValueError("%s: expected _Principal Coordinates Only_ at line %d" % (path, i + 1))
        n = nmol * nat
        rp = np.array([[float(x) for x in L[i + 1 + k].split()[:3]] for k in range(n)]).reshape(nmol, nat, 3)
        i += 1 + n
        v = None
        if i < len(L) and L[i].strip() == "_Velocities_":
            v = np.array([[float(x) for x in L[i + 1 + k].split()[:3]] for k in range(n)]).reshape(nmol, nat, 3)
            i += 1 + n
        out["species"][name] = dict(rp=rp, v=v, nmoles=nmol, natoms=nat)
    return out


# ------------------------------------------------------------ forcefield
def _kv(tok):
    if "@" in tok:
        k, v = tok.split("@", 1)
        return k.upper(), v
    return tok.upper(), None


def read_atom_atom(path):
    """atom_atom_file -> {(a1,a2): dict(model, sig, eps, hicut2, locut2, off)}
    following lj_init (lj.F90:140-262) defaults: locut = 1e-3, hicut = 13."""
    out = {}
    for l in _lines(path):
        s = _strip(l)
        f = s.split()
        if len(f) < 3:
            continue
        a1, a2, model = f[0], f[1], f[2].upper()
        d = dict(model=model, off=False, line=s)
        if model == "LJ":
            d.update(sig=None, eps=None, locut=C.LOW_DIST_CUTOFF,
                     locut2=C.LOW_DIST_CUTOFF * C.LOW_DIST_CUTOFF,
                     hicut=C.HIGH_DIST_CUTOFF, hicut2=C.HIGH_DIST_CUTOFF * C.HIGH_DIST_CUTOFF,
                     mix=False)
            for tok in f[3:]:
                k, v = _kv(tok)
                if k == "OFF":
                    d.update(off=True, hicut=0.0, hicut2=0.0)
                elif k == "SIG":
                    if "MIX" in v.upper():
                        d["mix"] = v.upper()
                    else:
                        d["sig"] = float(v)
                elif k == "EPS":
                    if "MIX" not in v.upper():
                        d["eps"] = float(v)
                elif k in ("HICUT", "TRUNC"):
                    d["hicut"] = float(v)
                    d["hicut2"] = d["hicut"] ** 2
                elif k in ("LOCUT", "LOWCUT"):
                    d["locut"] = float(v)
                    d["locut2"] = d["locut"] ** 2
            if not d["off"] and d["sig"] is not None and d["eps"] is not None:
                d["A"], d["B"] = lj_AB(d["sig"], d["eps"])
        out[(a1, a2)] = d
        out[(a2, a1)] = d
    # mixed pairs from the pure ones (lj_calc_mix, lj.F90:905-979): sigma
    # arithmetic (LBMIX) or geometric (JORGMIX), epsilon geometric
    for (a1, a2), d in out.items():
        if d["model"] != "LJ" or d["off"] or not d.get("mix") or "A" in d:
            continue
        p1, p2 = out.get((a1, a1)), out.get((a2, a2))
        if not p1 or not p2 or p1.get("sig") is None or p2.get("sig") is None:
            raise ValueError("mixing rule without the pure pairs: " + d["line"])
        if d["mix"] == "JORGMIX":
            d["sig"] = float(np.sqrt(p1["sig"] * p2["sig"]))
        else:
            d["sig"] = (p1["sig"] + p2["sig"]) / 2.0
        d["eps"] = float(np.sqrt(p1["eps"] * p2["eps"]))
        d["A"], d["B"] = lj_AB(d["sig"], d["eps"])
    return out


def lj_AB(sig, eps):
    """lj_calc_AB (lj.F90:860-869): B = 4 eps sig^6 kB, A = B sig^6."""
    s6 = sig ** 6
    B = 4.0 * eps * s6 * C.KCALMOLE_KB
    return B * s6, B


def read_spc_spc(path):
    """spc_spc_file -> {(s1,s2,'NCOUL'|'COUL'): dict(kind, tokens)}."""
    out = {}
    for l in _lines(path):
        s = _strip(l)
        f = s.split()
        if len(f) < 3 or f[2].upper() not in ("NCOUL", "COUL"):
            continue
        s1, s2, ic = f[0], f[1], f[2].upper()
        kind = f[3].upper() if len(f) > 3 else "OFF"
        d = dict(kind=kind, tokens=[t.upper() for t in f[3:]], line=s)
        if kind == "SUM" and "EWALD" in d["tokens"]:
            e = dict(kmax=None, kappa=None, locut=None, sfactor=False)
            for t in f[3:]:
                k, v = _kv(t)
                if k == "KMAX":
                    e["kmax"] = int(v)
                elif k == "KAPPA":
                    e["kappa"] = float(v)
                elif k == "LOCUT":
                    e["locut"] = float(v)
                elif k == "SFACTOR":
                    e["sfactor"] = True
            d["ewald"] = e
        out[(s1, s2, ic)] = d
        out[(s2, s1, ic)] = d
    return out


# ---------------------------------------------------------------- system
def load_system(ctrl_path):
    """Read a ctrlfile and everything it references.  Returns a dict with
    atom types, molecules (with replicated framework coordinates), the
    simulation cell (fund/full edges), forcefield tables and map jobs."""
    c = read_ctrlfile(ctrl_path)
    base = c["base"]
    atypes = []
    for a in c["atoms"]:
        d = read_atom(find_file(a["file"], "ATOMSDIR", base))
        d["name"] = a["name"]
        atypes.append(d)
    aidx = {a["name"]: i for i, a in enumerate(atypes)}
    mols = {}
    for m in c["molecules"]:
        d = read_molecule(find_file(m["file"], "MOLSDIR", base))
        d["name"] = m["name"]
        d["type"] = np.array([aidx[n] for n in d["atoms"]], dtype=np.int32)
        masses = [atypes[t]["mass"] for t in d["type"]]
        d["mass"] = float(np.sum(masses))
        # molecule_coordinit (molecule.F90:714-716): rigid molecules only,
        # and molecule_genprincipleaxes returns for fewer than 3 atoms (:843)
        if d.get("rigid") and len(d["xyz"]) >= 3:
            d["ref"] = principal_axes(d["xyz"], masses, d["translate"], d["rotate"])
        else:
            d["ref"] = d["xyz"].copy()
        mols[m["name"]] = d
    sc = c["simcell"]
    if sc["fundcell_source"]:
        fc = mols[sc["fundcell_source"]]["fundcell"]
    else:
        fc = sc["fund"]
    fund = np.asarray(fc["ell"], dtype=np.float64)
    angles = np.asarray(fc["angles"], dtype=np.float64)
    eff = np.asarray(fc["eff"], dtype=np.float64) if fc.get("eff") is not None else fund.copy()
    nxyz = np.asarray(sc["nxyz"], dtype=np.int64)
    full = fund * nxyz
    lvec = lattice_vectors(fund, angles)
    # replicate the framework (simcell.F90:200-216 -> molecule_expand,
    # molecule.F90:637-700): copies along the fundamental cell's lattice
    # vectors a, then b, then c, copy i shifted by lvec * i
    if sc["fundcell_source"]:
        fw = mols[sc["fundcell_source"]]
        xyz, typ, q = fw["xyz"], fw["type"], fw["q"]
        for ax in range(3):
            n = int(nxyz[ax])
            off = lvec[ax]
            xyz = np.concatenate([xyz + off * float(i) for i in range(n)])
            typ = np.concatenate([typ] * n)
            q = np.concatenate([q] * n)
        fw["full_xyz"], fw["full_type"], fw["full_q"] = xyz, typ, q
    ff = c["forcefield"]
    aa = read_atom_atom(find_file(ff["atom_atom_file"], "NONE", base))
    ss = read_spc_spc(find_file(ff["spc_spc_file"], "NONE", base))
    slanted = float(np.sum(np.abs(angles - 90.0))) > 0.001  # fundcell_init orthoflag
    return dict(ctrl=c, atypes=atypes, aidx=aidx, mols=mols, fund_ell=fund, full_ell=full,
                nxyz=nxyz, aa=aa, ss=ss, seed=c["general"]["iseed"], angles=angles, eff=eff,
                slanted=slanted, lvec=lvec)


def lattice_vectors(ell, angles_deg):
    """fundcell_genlatticevects (fundcell.F90:174-208): a along x, b in the
    xy plane, c completing the edge lengths; the angles in radians with the
    reference's pi = 3.1415927 (defaults.F90:33-35).  Rows a, b, c."""
    degtorad = 3.1415927 / 180.0
    abc, aac, aab = [float(a) * degtorad for a in angles_deg]
    e = [float(x) for x in ell]
    l1 = [e[0], 0.0, 0.0]
    l2 = [e[1] * math.cos(aab), e[1] * math.sin(aab), 0.0]
    c1 = e[2] * math.cos(aac)
    c2 = e[2] * (math.cos(abc) - math.cos(aac) * math.cos(aab)) / math.sin(aab)
    m2 = 0.0
    m2 = m2 + c1 * c1
    m2 = m2 + c2 * c2
    m2 = m2 + 0.0 * 0.0
    return np.array([l1, l2, [c1, c2, math.sqrt(e[2] * e[2] - m2)]])


def lj_table_for_probe(system, probe_type, ntypes=None):
    """Per framework-atom-type LJ parameters against one probe atom type:
    arrays A, B, hicut2, locut2 and an 'on' mask (OFF pairs -> on=0)."""
    at = system["atypes"]
    nt = len(at) if ntypes is None else ntypes
    A = np.zeros(nt)
    B = np.zeros(nt)
    hc2 = np.zeros(nt)
    lc2 = np.zeros(nt)
    on = np.zeros(nt, dtype=np.uint8)
    pn = at[probe_type]["name"]
    for t in range(nt):
        d = system["aa"].get((at[t]["name"], pn))
        if d is None or d["model"] != "LJ" or d["off"]:
            continue
        A[t], B[t], hc2[t], lc2[t], on[t] = d["A"], d["B"], d["hicut2"], d["locut2"], 1
    return A, B, hc2, lc2, on
