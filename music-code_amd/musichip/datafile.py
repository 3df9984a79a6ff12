"""The reference's binary datafile ("config file", datafile.F90): Fortran
unformatted sequential records with 4-byte length markers.

Layout (datafile_initout datafile.F90:148-225, datafile_writeheader
:596-638, datafile_writeExtraInfo :384-454, datafile_writeconfig :909-958,
datafile_writenrg :1034-1073):

* header FileHeader_Info (:102-108): ctrl_filename (48 chars), nlines
  (int32), ctrl_line(200) (192 chars each: the ctrlfile's lines read as
  '(a)' into 96 chars; lines past nlines are whatever the writer's memory
  held, zeros here), simno (int32), siminfo_string (480 chars);
* ContentDetails (:83-87): tot_nrg, pair_nrg, intra_nrg, posns, vels, time
  as 4-byte logicals;
* 100 extra-info lines of 48 chars (simno, pressure, temperature, simType,
  nspcs, 10 species names, 10 fixed flags, niters, iconfig);
* per configuration: [time], then per species i: [(noncoul, coul) for
  j >= i if pair_nrg], (kinetic, intra); [total if tot_nrg]; then per
  moving species with atoms: nmoles, [rp (natoms, nmoles, 3 as Fortran
  (1:natoms, 1:nmoles)%rp: x, y, z per atom)], [v].

read() parses such a file; write() produces it byte for byte from the same
fields.  Energies are kJ/mol as the reference's storestats holds them."""
import struct

import numpy as np

STRLEN = 48
LSTRLEN = 96
MAX_CTRL_LINES = 200
MAX_EXTRA_INFO_LINES = 100
MAX_SORBS = 10
FLAGS = ("tot_nrg", "pair_nrg", "intra_nrg", "posns", "vels", "time")


def _records(data):
    off, out = 0, []
    while off < len(data):
        n = struct.unpack_from("<i", data, off)[0]
        if n < 0 or off + 8 + n > len(data) or struct.unpack_from("<i", data, off + 4 + n)[0] != n:
            raise ValueError("not a Fortran sequential record at byte %d" % off)
        out.append(data[off + 4:off + 4 + n])
        off += 8 + n
    return out


def _rec(b):
    return struct.pack("<i", len(b)) + b + struct.pack("<i", len(b))


def _fix(s, n):
    b = s.encode("latin-1") if isinstance(s, str) else bytes(s)
    return (b + b" " * n)[:n]


def extra_info_lines(simno, pressure, temperature, simtype, species, fixed, niters, iconfig):
    """The 100 lines of datafile_writeExtraInfo (datafile.F90:384-454)."""
    L = [" " * STRLEN] * MAX_EXTRA_INFO_LINES
    L[0] = "%5d%5s%s" % (simno, "", " # simno")
    L[1] = "%12.5f%5s%s" % (pressure, "", " # simPressure")
    L[2] = "%12.5f%5s%s" % (temperature, "", " # simTemperature")
    L[3] = (simtype.strip() + "     " + " # simType")
    L[4] = "%5d%5s%s" % (len(species), "", " # nspcs")
    for i in range(MAX_SORBS):
        if i < len(species):
            L[5 + i] = "%-20s # spc%5d" % (species[i][:20], i + 1)
        else:
            L[5 + i] = "NOSPECIES%5s # spc" % ""
    for i in range(MAX_SORBS):
        if i < len(species):
            L[5 + MAX_SORBS + i] = "%s%5s # spc%5d" % ("FIXED" if fixed[i] else "NOT_FIXED", "", i + 1)
        else:
            L[5 + MAX_SORBS + i] = "NOSPECIES%5s # spc%5d" % ("", i + 1)
    L[5 + 2 * MAX_SORBS] = "%12d%5s%s" % (niters, "", " # niters")
    L[6 + 2 * MAX_SORBS] = "%12d%5s%s" % (iconfig, "", " # iconfig")
    return [_fix(x, STRLEN).decode("latin-1") for x in L]


def read(path, natoms=None):
    """Parse a datafile.  natoms: atoms per molecule of each moving species
    with atoms (taken from the record lengths when None).  Returns dict(
    ctrl_filename, ctrl_lines, ctrl_tail, simno, siminfo, flags, extra,
    configs=[dict(time, pair, kin_intra, total, species=[dict(nmoles, rp,
    v)])])."""
    recs = _records(open(path, "rb").read())
    h = recs[0]
    out = dict(ctrl_filename=h[:STRLEN].decode("latin-1"))
    nl = struct.unpack_from("<i", h, STRLEN)[0]
    o = STRLEN + 4
    lines = [h[o + 2 * LSTRLEN * i:o + 2 * LSTRLEN * (i + 1)] for i in range(MAX_CTRL_LINES)]
    out["ctrl_lines"] = [x.decode("latin-1") for x in lines[:nl]]
    out["ctrl_tail"] = b"".join(lines[nl:])
    o += 2 * LSTRLEN * MAX_CTRL_LINES
    out["simno"] = struct.unpack_from("<i", h, o)[0]
    out["siminfo"] = h[o + 4:o + 4 + 10 * STRLEN].decode("latin-1")
    fl = struct.unpack("<6i", recs[1])
    out["flags"] = {k: bool(v) for k, v in zip(FLAGS, fl)}
    out["flag_words"] = fl
    out["extra"] = [r.decode("latin-1") for r in recs[2:2 + MAX_EXTRA_INFO_LINES]]
    nspc = int(out["extra"][4].split()[0])
    fixed = [out["extra"][5 + MAX_SORBS + i].split()[0] == "FIXED" for i in range(nspc)]
    f = out["flags"]
    configs, i = [], 2 + MAX_EXTRA_INFO_LINES
    while i < len(recs):
        c = dict(time=None, pair=[], kin_intra=[], total=None, species=[])
        if f["time"]:
            c["time"] = struct.unpack("<d", recs[i])[0]
            i += 1
        for a in range(nspc):
            for b in range(a, nspc):
                if f["pair_nrg"]:
                    c["pair"].append(struct.unpack("<2d", recs[i]))
                    i += 1
            c["kin_intra"].append(struct.unpack("<2d", recs[i]))
            i += 1
        if f["tot_nrg"]:
            c["total"] = struct.unpack("<d", recs[i])[0]
            i += 1
        k = 0
        for a in range(nspc):
            if fixed[a]:
                continue
            if natoms is not None and natoms[k] == 0:
                k += 1
                continue
            nm = struct.unpack("<i", recs[i])[0]
            i += 1
            sp = dict(nmoles=nm, rp=None, v=None)
            for key, flag in (("rp", "posns"), ("v", "vels")):
                if nm > 0 and f[flag]:
                    arr = np.frombuffer(recs[i], dtype="<f8")
                    na = natoms[k] if natoms is not None else len(arr) // (3 * nm)
                    sp[key] = arr.reshape(nm, na, 3).copy()
                    i += 1
            c["species"].append(sp)
            k += 1
        configs.append(c)
    out["configs"] = configs
    return out


def write(path, d):
    """Write a datafile from the fields read() returns."""
    h = bytearray()
    h += _fix(d["ctrl_filename"], STRLEN)
    h += struct.pack("<i", len(d["ctrl_lines"]))
    for x in d["ctrl_lines"]:
        h += _fix(x, 2 * LSTRLEN)
    tail = d.get("ctrl_tail")
    ntail = 2 * LSTRLEN * (MAX_CTRL_LINES - len(d["ctrl_lines"]))
    h += tail if tail is not None else b"\0" * ntail
    h += struct.pack("<i", d["simno"])
    h += _fix(d["siminfo"], 10 * STRLEN)
    out = bytearray(_rec(bytes(h)))
    fl = d.get("flag_words") or tuple(1 if d["flags"][k] else 0 for k in FLAGS)
    out += _rec(struct.pack("<6i", *fl))
    for x in d["extra"]:
        out += _rec(_fix(x, STRLEN))
    f = d["flags"]
    for c in d["configs"]:
        if f["time"]:
            out += _rec(struct.pack("<d", c["time"]))
        k = 0
        nspc = len(c["kin_intra"])
        for a in range(nspc):
            for b in range(a, nspc):
                if f["pair_nrg"]:
                    out += _rec(struct.pack("<2d", *c["pair"][k]))
                    k += 1
            out += _rec(struct.pack("<2d", *c["kin_intra"][a]))
        if f["tot_nrg"]:
            out += _rec(struct.pack("<d", c["total"]))
        for sp in c["species"]:
            out += _rec(struct.pack("<i", sp["nmoles"]))
            for key, flag in (("rp", "posns"), ("v", "vels")):
                if sp["nmoles"] > 0 and f[flag]:
                    out += _rec(np.ascontiguousarray(sp[key], dtype="<f8").tobytes())
    with open(path, "wb") as fh:
        fh.write(bytes(out))
