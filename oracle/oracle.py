"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product path (music-code_amd/) never
imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ctypes.c_double
_i = ctypes.c_int
_ll = ctypes.c_longlong
_vp = ctypes.c_void_p


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def _host_tag():
    """Tag of the -march=native build: the CPU model and flags of this host."""
    import hashlib
    info = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith(("model name", "flags")):
                info += line
                if line.startswith("flags"):
                    break
    except OSError:
        pass
    return hashlib.sha1(info.encode()).hexdigest()[:10]


def lib(native=False):
    """Load (building once if absent) liboracle.so."""
    global _LIB
    if _LIB is not None and not native:
        return _LIB
    name = "liboracle_native_%s.so" % _host_tag() if native else "liboracle.so"
    path = os.path.join(HERE, name)
    srcs = [os.path.join(HERE, f) for f in os.listdir(HERE) if f.endswith((".c", ".h"))]
    if not os.path.exists(path) or os.path.getmtime(path) < max(os.path.getmtime(f) for f in srcs):
        cmd = ["make", "-s", "-C", HERE]
        subprocess.check_call(cmd + (["native", "NATIVE_OUT=" + name] if native else ["all"]))
    L = ctypes.CDLL(path)
    L.mo_rranf_n.argtypes = [ctypes.c_int32, _i, _vp, _vp, _vp]
    L.mo_erfc.argtypes = [_d]
    L.mo_erfc.restype = _d
    L.mo_lj_node.argtypes = [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.mo_lj_node.restype = _i
    L.mo_buck_node.argtypes = [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.mo_buck_node.restype = _i
    L.mo_ewald_sfactor.argtypes = [_i, _vp, _vp, _vp, _d, _i, _d, _i, _vp, _vp, _vp, _vp, _vp]
    L.mo_ewald_sfactor.restype = _i
    L.mo_ewald_node.argtypes = [_vp, _i, _vp, _vp, _d, _d, _i, _vp, _vp, _vp, _vp, _vp]
    L.mo_ewald_node.restype = _i
    L.mo_build_nodes.argtypes = [_i, _vp, _vp, _vp, _ll, _ll, _vp]
    L.mo_compact.argtypes = [_vp, _vp, _d, _vp, _vp, _vp]
    L.mo_compact.restype = _ll
    L.mo_map_write.argtypes = [ctypes.c_char_p, _i, _vp, ctypes.c_float, _vp, _ll, _vp, _vp]
    L.mo_map_write.restype = _i
    L.mo_map_read_sizes.argtypes = [ctypes.c_char_p, _vp, _vp, _vp, _vp]
    L.mo_map_read_sizes.restype = _i
    L.mo_map_read.argtypes = [ctypes.c_char_p, _vp, _vp]
    L.mo_map_read.restype = _i
    L.mo_pothermite.argtypes = [_vp, _vp, _vp, _vp, _ll, _vp, _vp]
    L.mo_pothermite.restype = _i
    L.mo_forcehermite.argtypes = [_vp, _vp, _vp, _vp, _ll, _vp, _vp]
    L.mo_forcehermite.restype = _i
    L.mo_maptouc.argtypes = [_vp, _vp, _vp]
    L.mo_pothermite_nowrap.argtypes = [_vp, _vp, _vp, _vp, _ll, _vp, _vp]
    L.mo_pothermite_nowrap.restype = _i
    L.mo_cell_init.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
    L.mo_cell_minimage.argtypes = [_vp, _vp, _vp, _vp]
    L.mo_cell_maptouc.argtypes = [_vp, _vp, _vp]
    L.mo_cell_pbc.argtypes = [_vp, _vp, _vp]
    L.mo_node_at.argtypes = [_i, _vp, _vp, _vp]
    L.mo_node_at.restype = _i
    L.mo_cell_uniform.argtypes = [_vp, _vp, _vp]
    L.mo_cell_volume.argtypes = [_vp]
    L.mo_cell_volume.restype = _d
    if not native:
        _LIB = L
    return L


def rranf(seed, n):
    out = np.zeros(n)
    ix = ctypes.c_int32()
    iy = ctypes.c_int32()
    lib().mo_rranf_n(seed, n, _p(out), ctypes.byref(ix), ctypes.byref(iy))
    return out, ix.value, iy.value


def erfc(x):
    return np.array([lib().mo_erfc(float(v)) for v in np.atleast_1d(x)])


class MoCell(ctypes.Structure):
    _fields_ = [("ortho", _i), ("fell", _d * 3), ("finv", _d * 3), ("ell", _d * 3), ("half", _d * 3),
                ("origin", _d * 3), ("eff", _d * 3), ("U", (_d * 3) * 3), ("S", (_d * 3) * 3),
                ("lvec", (_d * 3) * 3)]


def cell(fund_ell, angles_deg, eff=None, ncell=(1, 1, 1), origin=(0.0, 0.0, 0.0)):
    """A (possibly slanted) simulation cell: fundcell_init + fundcell_construct."""
    c = MoCell()
    fe = np.ascontiguousarray(fund_ell, dtype=np.float64)
    an = np.ascontiguousarray(angles_deg, dtype=np.float64)
    og = np.ascontiguousarray(origin, dtype=np.float64)
    ef = np.ascontiguousarray(fund_ell if eff is None else eff, dtype=np.float64)
    nc = np.ascontiguousarray(ncell, dtype=np.int32)
    lib().mo_cell_init(ctypes.byref(c), _p(fe), _p(an), _p(og), _p(ef), _p(nc))
    return c


def cell_matrices(c):
    return (np.array([[c.U[i][j] for j in range(3)] for i in range(3)]),
            np.array([[c.S[i][j] for j in range(3)] for i in range(3)]),
            np.array([[c.lvec[i][j] for j in range(3)] for i in range(3)]))


def cell_minimage(c, a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.zeros(3)
    lib().mo_cell_minimage(ctypes.byref(c), _p(a), _p(b), _p(out))
    return out


def cell_maptouc(c, r):
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros(3)
    lib().mo_cell_maptouc(ctypes.byref(c), _p(r), _p(out))
    return out


def cell_pbc(c, r):
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros(3)
    lib().mo_cell_pbc(ctypes.byref(c), _p(r), _p(out))
    return out


class _Framework(ctypes.Structure):
    _fields_ = [("natoms", _i), ("xyz", _vp), ("type", _vp), ("q", _vp),
                ("A", _vp), ("B", _vp), ("hicut2", _vp), ("locut2", _vp),
                ("bA", _vp), ("brho", _vp), ("bC", _vp), ("bD", _vp),
                ("charge1", _d), ("kappa", _d), ("nk", _i),
                ("k", _vp), ("s", _vp), ("si", _vp), ("full_ell", _d * 3), ("cell", _vp)]


def ewald_sfactor(xyz, q, ell, kappa_param, kmax, locut, nkmax=50000):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    ell = np.ascontiguousarray(ell, dtype=np.float64)
    ncand = (2 * kmax + 1) ** 3
    k = np.zeros((ncand, 3))
    s = np.zeros(ncand)
    si = np.zeros(ncand)
    kidx = np.zeros((ncand, 3), dtype=np.int32)
    kap = ctypes.c_double()
    nk = lib().mo_ewald_sfactor(len(q), _p(xyz), _p(q), _p(ell), kappa_param, kmax, locut,
                                nkmax, _p(k), _p(s), _p(si), _p(kidx), ctypes.byref(kap))
    if nk < 0:
        raise RuntimeError("too many k vectors")
    return dict(nk=nk, k=k[:nk].copy(), s=s[:nk].copy(), si=si[:nk].copy(),
                kidx=kidx[:nk].copy(), kappa=kap.value)


def node_at(kind, fw, points):
    """Nodes at arbitrary points [n,3] -> (ok[n], vals[n,8]); fw as in
    build_nodes (with an optional slanted 'cell')."""
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    st, keep = _framework_struct(fw)
    ok = np.zeros(len(pts), dtype=bool)
    out = np.zeros((len(pts), 8))
    o = np.zeros(8)
    for i in range(len(pts)):
        ok[i] = bool(lib().mo_node_at(kind, ctypes.byref(st), _p(pts[i]), _p(o)))
        out[i] = o
    return ok, out


def cell_volume(c):
    """Simulation cell volume (simcell_getvolume)."""
    return lib().mo_cell_volume(ctypes.byref(c))


def cell_uniform(c, seed, n):
    """n points of simcell_uniformRandPt after random_init(seed)."""
    _gcmc_lib()
    r = _Rng()
    lib().mo_rng_init(ctypes.byref(r), seed)
    out = np.zeros((n, 3))
    v = np.zeros(3)
    for i in range(n):
        lib().mo_cell_uniform(ctypes.byref(c), ctypes.byref(r), _p(v))
        out[i] = v
    return out


def _framework_struct(fw):
    """mo_framework from a dict (xyz, type, q, A, B, hicut2, locut2, bA,
    brho, bC, bD, charge1, kappa, k, s, si, full_ell, optional cell);
    returns the struct and the arrays it points into."""
    keep = {}

    def arr(name, dt=np.float64):
        v = fw.get(name)
        if v is None:
            return None
        a = np.ascontiguousarray(v, dtype=dt)
        keep[name] = a
        return _p(a)

    st = _Framework()
    st.natoms = len(fw["xyz"])
    st.xyz = arr("xyz")
    st.type = arr("type", np.int32)
    st.q = arr("q")
    st.A = arr("A")
    st.B = arr("B")
    st.hicut2 = arr("hicut2")
    st.locut2 = arr("locut2")
    st.bA = arr("bA")
    st.brho = arr("brho")
    st.bC = arr("bC")
    st.bD = arr("bD")
    st.charge1 = float(fw.get("charge1", 1.0))
    st.kappa = float(fw.get("kappa", 0.0))
    st.nk = int(len(fw["s"])) if fw.get("s") is not None else 0
    st.k = arr("k")
    st.s = arr("s")
    st.si = arr("si")
    for i in range(3):
        st.full_ell[i] = float(fw["full_ell"][i])
    c = fw.get("cell")
    st.cell = ctypes.cast(ctypes.byref(c), _vp) if c is not None else None
    keep["cell"] = c
    return st, keep


def build_nodes(kind, fw, npnts, step, m0=0, m1=None):
    """Evaluate grid nodes m0..m1 (config_makegrid order) -> [n, 8] FP64.

    fw: dict with xyz, type, q, A, B, hicut2, locut2 (per type), charge1,
    kappa, k, s, si, full_ell (and bA, brho, bC, bD for Buckingham; a
    slanted 'cell' from cell())."""
    npnts = np.ascontiguousarray(npnts, dtype=np.int32)
    step = np.ascontiguousarray(step, dtype=np.float64)
    if m1 is None:
        m1 = int(np.prod(npnts.astype(np.int64)))
    st, keep = _framework_struct(fw)
    out = np.zeros((m1 - m0, 8))
    lib().mo_build_nodes(kind, ctypes.byref(st), _p(npnts), _p(step), m0, m1, _p(out))
    return out


def compact(npnts, r, cutoff):
    npnts = np.ascontiguousarray(npnts, dtype=np.int32)
    r = np.ascontiguousarray(r, dtype=np.float32)
    n = int(np.prod(npnts.astype(np.int64)))
    ptr = np.zeros(n, dtype=np.int32)
    src = np.zeros(n, dtype=np.int64)
    counts = np.zeros(4, dtype=np.int64)
    nacc = lib().mo_compact(_p(npnts), _p(r), float(cutoff), _p(ptr), _p(src), _p(counts))
    return ptr, src[:nacc].copy(), counts


class _ZAtom(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 26), ("symbol", ctypes.c_char * 3),
                ("charge", ctypes.c_float), ("A12", ctypes.c_float), ("B6", ctypes.c_float),
                ("C2", ctypes.c_float), ("D0", ctypes.c_float), ("rcuthi", ctypes.c_float),
                ("rcutlo", ctypes.c_float)]


def map_write(path, zatoms, pcut, npnts, ptr, vals8):
    """zatoms: list of dicts(name, symbol, charge, A12, B6, C2, D0, rcuthi, rcutlo)."""
    arr = (_ZAtom * max(1, len(zatoms)))()
    for i, z in enumerate(zatoms):
        arr[i].name = z["name"].encode()[:25]
        arr[i].symbol = z["symbol"].encode()[:2]
        for f in ("charge", "A12", "B6", "C2", "D0", "rcuthi", "rcutlo"):
            setattr(arr[i], f, float(z.get(f, 0.0)))
    npnts = np.ascontiguousarray(npnts, dtype=np.int32)
    ptr = np.ascontiguousarray(ptr, dtype=np.int32)
    vals8 = np.ascontiguousarray(vals8, dtype=np.float32)
    nf = vals8.shape[1]
    rc = lib().mo_map_write(path.encode(), len(zatoms), ctypes.cast(arr, _vp), float(pcut),
                            _p(npnts), nf, _p(ptr), _p(vals8))
    if rc:
        raise IOError("mo_map_write failed")


def map_read(path):
    np3 = np.zeros(3, dtype=np.int32)
    nf = ctypes.c_longlong()
    pc = ctypes.c_float()
    nz = ctypes.c_int()
    if lib().mo_map_read_sizes(path.encode(), _p(np3), ctypes.byref(nf), ctypes.byref(pc),
                               ctypes.byref(nz)):
        raise IOError("bad map header " + path)
    ptr = np.zeros(int(np.prod(np3.astype(np.int64))), dtype=np.int32)
    vals = np.zeros((8, nf.value), dtype=np.float32)
    if lib().mo_map_read(path.encode(), _p(ptr), _p(vals)):
        raise IOError("bad map body " + path)
    return dict(npnts=np3, nfsize=nf.value, pcut=pc.value, ptr=ptr, vals=vals)


def pothermite(npnts, step, ptr, vals8, coords, fund_ell=None, force=False, cell=None):
    """Interpolate at coords [n,3]; if fund_ell (or a slanted cell) is given,
    coords are first folded with simcell_maptouc.  Returns (ok[n], V[n]) or
    (ok, p[n,4]) if force.  A slanted cell interpolates without the
    periodic corner wrap (interpolate.F90:675-679)."""
    npnts = np.ascontiguousarray(npnts, dtype=np.int32)
    step = np.ascontiguousarray(step, dtype=np.float64)
    ptr = np.ascontiguousarray(ptr, dtype=np.int32)
    vals8 = np.ascontiguousarray(vals8, dtype=np.float32)
    coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
    n = len(coords)
    ok = np.zeros(n, dtype=bool)
    out = np.zeros((n, 4)) if force else np.zeros(n)
    L = lib()
    fe = np.ascontiguousarray(fund_ell, dtype=np.float64) if fund_ell is not None else None
    c = np.zeros(3)
    for i in range(n):
        if cell is not None:
            L.mo_cell_maptouc(ctypes.byref(cell), _p(coords[i]), _p(c))
        elif fe is not None:
            L.mo_maptouc(_p(fe), _p(coords[i]), _p(c))
        else:
            c[:] = coords[i]
        if force:
            p = np.zeros(4)
            ok[i] = bool(L.mo_forcehermite(_p(npnts), _p(step), _p(ptr), _p(vals8),
                                           vals8.shape[1], _p(c), _p(p)))
            out[i] = p
        else:
            v = ctypes.c_double()
            f = L.mo_pothermite_nowrap if (cell is not None and not cell.ortho) else L.mo_pothermite
            ok[i] = bool(f(_p(npnts), _p(step), _p(ptr), _p(vals8), vals8.shape[1], _p(c), ctypes.byref(v)))
            out[i] = v.value
    return ok, out


# ------------------------------------------------------------------ GCMC
MV_INSERT, MV_DELETE, MV_TRANSLATE, MV_ROTATE, MV_BINSERT, MV_BDELETE = 0, 1, 2, 3, 4, 5
MOVE_NAMES = {"RINSERT": 0, "RDELETE": 1, "RTRANSLATE": 2, "RROTATE": 3,
              "INSERT": 0, "DELETE": 1, "TRANSLATE": 2, "ROTATE": 3, "BINSERT": 4, "BDELETE": 5}


class MoBmap(ctypes.Structure):
    _fields_ = [("ncub", _i * 3), ("ncubelets", _ll), ("numtab", _ll), ("step", _d * 3), ("eff", _d * 3),
                ("cum_wt", _d), ("clamped", _i), ("maxnrgval", _d), ("cube", _vp), ("wt", _vp),
                ("cum_prob", _vp), ("cube_index", _vp)]


class Bmap:
    """The oracle's bias map (mo_bmap_build, a restatement of bmap_cubeinit,
    bmap.F90:181-403) of a compacted map (ptr [nx][ny][nz] 1-based, vals8
    [8][nfsize]) at a bias temperature."""

    def __init__(self, npnts, step, eff, ptr, vals8, T_bias):
        L = _gcmc_lib()
        self.b = MoBmap()
        npn = np.ascontiguousarray(npnts, dtype=np.int32)
        st = np.ascontiguousarray(step, dtype=np.float64)
        ef = np.ascontiguousarray(eff, dtype=np.float64)
        self._ptr = np.ascontiguousarray(ptr, dtype=np.int32)
        self._V = np.ascontiguousarray(np.asarray(vals8, dtype=np.float32)[0])
        rc = L.mo_bmap_build(ctypes.byref(self.b), _p(npn), _p(st), _p(ef), _p(self._ptr), _p(self._V),
                             float(T_bias))
        if rc:
            raise ValueError("mo_bmap_build failed (no cubelet below 30 kcal/mol)")
        n = self.b.numtab
        self.numtab, self.ncubelets, self.cum_wt = n, self.b.ncubelets, self.b.cum_wt
        self.clamp_energy = self.b.maxnrgval
        self.cube = np.ctypeslib.as_array(ctypes.cast(self.b.cube, ctypes.POINTER(ctypes.c_int32)), (n,)).copy()
        self.wt = np.ctypeslib.as_array(ctypes.cast(self.b.wt, ctypes.POINTER(_d)), (n,)).copy()
        self.cum = np.ctypeslib.as_array(ctypes.cast(self.b.cum_prob, ctypes.POINTER(_d)), (n,)).copy()

    def index(self, key):
        """bmap_getbiasindex (the reference's bisection), 0-based"""
        return _gcmc_lib().mo_bmap_index(ctypes.byref(self.b), float(key))

    def __del__(self):
        try:
            _gcmc_lib().mo_bmap_free(ctypes.byref(self.b))
        except Exception:
            pass


class GcmcSys(ctypes.Structure):
    _fields_ = [("full_ell", _d * 3), ("fund_ell", _d * 3), ("ns", _i),
                ("ref", (_d * 3) * 8), ("stype", _i * 8), ("sq", _d * 8),
                ("site_pmap", _i * 8), ("site_emap", _i * 8), ("nmaps", _i),
                ("map_ptr", _vp * 8), ("map_vals", _vp * 8), ("map_npnts", (_i * 3) * 8),
                ("map_step", (_d * 3) * 8), ("map_nf", _ll * 8),
                ("nst", _i), ("lj_A", _d * 64), ("lj_B", _d * 64), ("lj_h2", _d * 64),
                ("lj_l2", _d * 64), ("lj_on", _i * 64),
                ("ewald", _i), ("kmax", _i), ("kappa", _d), ("rc2_real", _d), ("T", _d),
                ("nmoves", _i), ("move_type", _i * 4), ("cumwt", _d * 4),
                ("insdelratio", _d), ("dtrans", _d), ("drot", _d), ("max_nrg", _d), ("bias", _vp),
                ("coul", _i), ("coul_on", _i * 64), ("coul_hicut", _d), ("coul_locut", _d),
                ("coul_shield", _d), ("adjust_trans", _i), ("adjust_rot", _i), ("cell", _vp),
                ("lj_comcut2", _d * 4), ("coul_comcut2", _d * 4), ("cav", _vp), ("cav_ntrials", _i),
                ("cav_map", _i)]


class MoCavity(ctypes.Structure):
    _fields_ = [("n", _i * 3), ("d", _d * 3), ("tab", _vp), ("spc_mask", _i)]


class GcmcTrace(ctypes.Structure):
    _fields_ = [("move", _i), ("accepted", _i), ("N", _i), ("dE", _d)]


class _Rng(ctypes.Structure):
    _fields_ = [("ix", ctypes.c_int32), ("iy", ctypes.c_int32), ("iseed", ctypes.c_int32),
                ("first", ctypes.c_int32)]


class GcmcState(ctypes.Structure):
    _fields_ = [("N", _i), ("maxN", _i), ("com", _vp), ("ang", _vp), ("rp", _vp), ("r", _vp),
                ("nk", _i), ("kv", _vp), ("wk", _vp), ("Sre", _vp), ("Sim", _vp),
                ("energy", _d), ("B", _d), ("rng", _Rng), ("att", _ll * 4), ("acc", _ll * 4),
                ("sumN", _d), ("nsteps", _ll), ("trace", _vp), ("trace_cap", _i),
                ("delta", _d * 2), ("cube_cnt", _vp), ("cav_avg", _vp), ("cav_n", _vp), ("cav_cap", _i)]


def _gcmc_lib():
    L = lib()
    if not hasattr(L, "_gcmc_sig"):
        L.mo_gcmc_kvecs.argtypes = [_vp, _vp, _vp, _vp]
        L.mo_gcmc_mol_const.argtypes = [_vp]
        L.mo_gcmc_mol_const.restype = _d
        L.mo_gcmc_run.argtypes = [_vp, _vp, _ll]
        L.mo_gcmc_run.restype = _i
        L.mo_gcmc_total_energy.argtypes = [_vp, _vp]
        L.mo_gcmc_total_energy.restype = _d
        L.mo_gcmc_recompute_sk.argtypes = [_vp, _vp]
        L.mo_gcmc_set_mol.argtypes = [_vp, _vp, _i, _vp, _vp]
        L.mo_gcmc_run_mix.argtypes = [_vp, _i, _vp, _ll]
        L.mo_gcmc_run_mix.restype = _i
        L.mo_gcmc_total_energy_mix.argtypes = [_vp, _i, _vp]
        L.mo_gcmc_total_energy_mix.restype = _d
        L.mo_gcmc_recompute_sk_mix.argtypes = [_vp, _i, _vp]
        L.mo_gcmc_set_mol_mix.argtypes = [_vp, _vp, _i, _i, _vp, _vp]
        L.mo_gcmc_cav_rebuild.argtypes = [_vp, _i, _vp]
        L.mo_adams_B.argtypes = [_d, _d, _d, _d]
        L.mo_adams_B.restype = _d
        L.mo_rng_init.argtypes = [_vp, ctypes.c_int32]
        L.mo_bmap_build.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _d]
        L.mo_bmap_build.restype = _i
        L.mo_bmap_free.argtypes = [_vp]
        L.mo_bmap_index.argtypes = [_vp, _d]
        L.mo_bmap_index.restype = _ll
        L.mo_euler_xyz.argtypes = [_i, _vp, _vp, _vp, _vp]
        L.mo_coul_pair.argtypes = [_d, _d, _d, _d, _d, _d, ctypes.POINTER(_d)]
        L.mo_coul_pair.restype = _i
        L.mo_bonded_term.argtypes = [_i, _vp, _vp, _vp]
        L.mo_bonded_term.restype = _d
        L._gcmc_sig = True
    return L


def euler_xyz(ref, com, ang):
    """Site coordinates com + R(phi, theta, psi) ref of one rigid molecule."""
    ref = np.ascontiguousarray(ref, dtype=np.float64)
    com = np.ascontiguousarray(com, dtype=np.float64)
    ang = np.ascontiguousarray(ang, dtype=np.float64)
    out = np.zeros_like(ref)
    _gcmc_lib().mo_euler_xyz(len(ref), _p(ref), _p(com), _p(ang), _p(out))
    return out


def coul_pair(hicut, locut, shield, sep, q1, q2):
    """(ok, energy) of one Coulomb site pair at separation sep."""
    sep = [float(x) for x in sep]
    r2 = 0.0
    for x in sep:
        r2 = r2 + x * x
    e = _d()
    ok = _gcmc_lib().mo_coul_pair(hicut, locut, shield, r2, q1, q2, ctypes.byref(e))
    return bool(ok), e.value


def bonded_term(kind, par, coords):
    """(energy, forces [natoms][3]) of one bond (1), angle (2) or torsion (3)."""
    par = np.ascontiguousarray(par, dtype=np.float64)
    r = np.zeros(12)
    c = np.asarray(coords, dtype=np.float64).ravel()
    r[:len(c)] = c
    f = np.zeros(12)
    u = _gcmc_lib().mo_bonded_term(kind, _p(par), _p(r), _p(f))
    return u, f[:len(c)].reshape(-1, 3)


def adams_B(fug_kPa, volume, T, Zig=1.0):
    return _gcmc_lib().mo_adams_B(fug_kPa, volume, T, Zig)


class Gcmc:
    """One GCMC chain on the CPU oracle.  spec: dict with full_ell, fund_ell,
    ref [ns,3], stype [ns], sq [ns], nst, lj (A, B, h2, l2, on: [nst*nst]),
    maps: list of (npnts, step, ptr, vals8) and site_pmap/site_emap,
    ewald (kappa, kmax, rc) or coul (on [nst*nst], hicut, locut, shield), T,
    moves [(type, weight)], dtrans, drot, and for BINSERT / BDELETE bias (a
    Bmap); a slanted simulation cell as 'cell' (from cell()).  Cavity bias:
    'cavity' = dict(n=(nx, ny, nz), tab=None or int8 [nx, ny, nz] (1 =
    tabulated), mask=species bits that occupy cubes), and per species
    'cav_ntrials' (> 0: its BINSERT / BDELETE are cavity biased) and
    'cav_map' (MAP_CAVITY_BIAS: centres from its bias map)."""

    def __init__(self, spec, fugacity_kPa, seed, maxN=2048, trace=0):
        L = _gcmc_lib()
        s = GcmcSys()
        self.keep = []
        for i in range(3):
            s.full_ell[i] = spec["full_ell"][i]
            s.fund_ell[i] = spec["fund_ell"][i]
        ref = np.asarray(spec["ref"], dtype=np.float64).reshape(-1, 3)
        s.ns = ns = len(ref)
        for a in range(ns):
            for i in range(3):
                s.ref[a][i] = ref[a][i]
            s.stype[a] = int(spec["stype"][a])
            s.sq[a] = float(spec["sq"][a])
            s.site_pmap[a] = int(spec.get("site_pmap", [-1] * ns)[a])
            s.site_emap[a] = int(spec.get("site_emap", [-1] * ns)[a])
        maps = spec.get("maps", [])
        s.nmaps = len(maps)
        for j, (npn, step, ptr, vals) in enumerate(maps):
            ptr = np.ascontiguousarray(ptr, dtype=np.int32)
            vals = np.ascontiguousarray(vals, dtype=np.float32)
            self.keep += [ptr, vals]
            s.map_ptr[j] = ptr.ctypes.data
            s.map_vals[j] = vals.ctypes.data
            s.map_nf[j] = vals.shape[1]
            for i in range(3):
                s.map_npnts[j][i] = int(npn[i])
                s.map_step[j][i] = float(step[i])
        s.nst = spec["nst"]
        A, B, h2, l2, on = spec["lj"]
        for t in range(s.nst * s.nst):
            s.lj_A[t], s.lj_B[t], s.lj_h2[t], s.lj_l2[t], s.lj_on[t] = A[t], B[t], h2[t], l2[t], int(on[t])
        ew = spec.get("ewald")
        if ew:
            s.ewald, s.kappa, s.kmax, s.rc2_real = 1, ew["kappa"], ew["kmax"], ew["rc"] ** 2
        cl = spec.get("coul")  # plain Coulomb: dict(on=[nst*nst], hicut, locut, shield)
        if cl:
            s.coul = 1
            for t in range(s.nst * s.nst):
                s.coul_on[t] = int(cl["on"][t])
            s.coul_hicut, s.coul_locut, s.coul_shield = cl["hicut"], cl["locut"], cl.get("shield", 1.0)
        s.T = spec["T"]
        mv = spec["moves"]
        s.nmoves = len(mv)
        wsum = sum(w for _, w in mv)
        c = 0.0
        wi = wd = None
        for i, (t, w) in enumerate(mv):
            s.move_type[i] = MOVE_NAMES[t] if isinstance(t, str) else int(t)
            c = c + w / wsum
            s.cumwt[i] = c
            if s.move_type[i] in (0, 4):
                wi = w / wsum
            if s.move_type[i] in (1, 5):
                wd = w / wsum
        s.insdelratio = (wi / wd) if (wi and wd) else 1.0
        s.dtrans = spec.get("dtrans", 0.0)
        s.drot = spec.get("drot", 0.0)
        s.adjust_trans = int(spec.get("adjust_trans", 0))
        s.adjust_rot = int(spec.get("adjust_rot", 0))
        s.max_nrg = spec.get("max_nrg", 1.0e10)
        # BINSERT / BDELETE: spec["bias"] is a Bmap
        self.bias = spec.get("bias")
        s.bias = ctypes.cast(ctypes.byref(self.bias.b), _vp) if self.bias is not None else None
        self.cell = spec.get("cell")
        s.cell = ctypes.cast(ctypes.byref(self.cell), _vp) if self.cell is not None else None
        # centre-of-mass cutoffs (ssbasic COMCUT@r): dict(lj=r, coul=r), each
        # a radius for every species pair or a list per partner species
        cc = spec.get("comcut") or {}
        for key, fld in (("lj", s.lj_comcut2), ("coul", s.coul_comcut2)):
            v = cc.get(key, 0.0)
            v = list(v) if np.ndim(v) else [v] * 4
            for i, r in enumerate(v):
                fld[i] = float(r) ** 2 if r else 0.0
        cv = spec.get("cavity")
        self.cav = None
        if cv is not None:
            c = MoCavity()
            for i in range(3):
                c.n[i] = int(cv["n"][i])
                c.d[i] = float(spec["full_ell"][i]) / int(cv["n"][i])
            tab = cv.get("tab")
            if tab is not None:
                tab = np.ascontiguousarray(tab, dtype=np.int8).reshape(-1)
                self.keep.append(tab)
                c.tab = tab.ctypes.data
            c.spc_mask = int(cv.get("mask", 1))
            self.cav = c
            s.cav = ctypes.cast(ctypes.byref(c), _vp)
        s.cav_ntrials = int(spec.get("cav_ntrials", 0))
        s.cav_map = int(spec.get("cav_map", 0))
        self.sys = s
        st = GcmcState()
        st.maxN = maxN
        st.delta[0], st.delta[1] = s.dtrans, s.drot
        self.com = np.zeros((maxN, 3))
        self.ang = np.zeros((maxN, 3))
        self.rp = np.zeros((maxN, ns, 3))
        self.r = np.zeros((maxN, ns, 3))
        st.com, st.ang, st.rp, st.r = [a.ctypes.data for a in (self.com, self.ang, self.rp, self.r)]
        nk = ctypes.c_int()
        L.mo_gcmc_kvecs(ctypes.byref(s), ctypes.byref(nk), None, None)
        n = nk.value if ew else 0
        self.kv = np.zeros((max(n, 1), 3))
        self.wk = np.zeros(max(n, 1))
        self.Sre = np.zeros(max(n, 1))
        self.Sim = np.zeros(max(n, 1))
        if ew:
            L.mo_gcmc_kvecs(ctypes.byref(s), ctypes.byref(nk), self.kv.ctypes.data, self.wk.ctypes.data)
        st.nk = n
        st.kv, st.wk, st.Sre, st.Sim = [a.ctypes.data for a in (self.kv, self.wk, self.Sre, self.Sim)]
        vol = float(np.prod(spec["full_ell"]))
        if self.cell is not None and not self.cell.ortho:
            vol = cell_volume(self.cell)
        st.B = adams_B(fugacity_kPa, vol, spec["T"])
        L.mo_rng_init(ctypes.byref(st.rng), int(seed))
        ncube = int(np.prod(cv["n"])) if cv is not None else 1
        self.cube_cnt = np.zeros(ncube, dtype=np.int32)
        # cav_prob(t) of rigidmoves_equateInsBias: one update of ntrials / 2
        self.cav_cap = 4 * maxN + 2
        self.cav_avg = np.full(self.cav_cap, s.cav_ntrials / 2.0)
        self.cav_n = np.ones(self.cav_cap, dtype=np.int64)
        st.cube_cnt, st.cav_avg, st.cav_n = [a.ctypes.data for a in (self.cube_cnt, self.cav_avg, self.cav_n)]
        st.cav_cap = self.cav_cap
        self.trace = (GcmcTrace * max(trace, 1))()
        st.trace = ctypes.cast(self.trace, _vp) if trace else None
        st.trace_cap = trace
        self.st = st

    def set_config(self, com, ang=None):
        L = _gcmc_lib()
        com = np.asarray(com, dtype=np.float64).reshape(-1, 3)
        ang = np.zeros_like(com) if ang is None else np.asarray(ang, dtype=np.float64).reshape(-1, 3)
        for m in range(len(com)):
            L.mo_gcmc_set_mol(ctypes.byref(self.sys), ctypes.byref(self.st), m,
                              np.ascontiguousarray(com[m]).ctypes.data,
                              np.ascontiguousarray(ang[m]).ctypes.data)
        self.st.N = len(com)
        L.mo_gcmc_recompute_sk(ctypes.byref(self.sys), ctypes.byref(self.st))
        sp = (_vp * 1)(ctypes.cast(ctypes.byref(self.sys), _vp))
        st = (_vp * 1)(ctypes.cast(ctypes.byref(self.st), _vp))
        L.mo_gcmc_cav_rebuild(sp, 1, st)
        self.st.energy = self.total_energy()

    def run(self, nsteps):
        rc = _gcmc_lib().mo_gcmc_run(ctypes.byref(self.sys), ctypes.byref(self.st), nsteps)
        if rc:
            raise RuntimeError("mo_gcmc_run failed %d" % rc)

    def total_energy(self):
        return _gcmc_lib().mo_gcmc_total_energy(ctypes.byref(self.sys), ctypes.byref(self.st))

    def mol_const(self):
        return _gcmc_lib().mo_gcmc_mol_const(ctypes.byref(self.sys))

    def trace_array(self, n=None):
        n = self.st.trace_cap if n is None else n
        return np.array([(t.move, t.accepted, t.N, t.dE) for t in self.trace[:n]])


GLOBAL_KEYS = ("full_ell", "fund_ell", "nst", "lj", "coul", "ewald", "T", "cell", "cavity")


class GcmcMix:
    """A GCMC chain of several species on the CPU oracle (mo_gcmc_run_mix):
    specs[i] as for Gcmc; the cell, pair tables (nst site types shared by
    the species), Coulomb / Ewald and T are specs[0]'s.  fugacities[i] per
    species.  The trace (RNG, energy, S(k)) lives with species 0."""

    def __init__(self, specs, fugacities, seed, maxN=2048, trace=0):
        self.nspc = len(specs)
        glob = {k: specs[0][k] for k in GLOBAL_KEYS if k in specs[0]}
        self.ch = [Gcmc(dict(sp, **glob), fugacities[i], seed if i == 0 else 1, maxN=maxN,
                        trace=trace if i == 0 else 0) for i, sp in enumerate(specs)]
        self.sysp = (_vp * self.nspc)(*[ctypes.cast(ctypes.byref(c.sys), _vp) for c in self.ch])
        self.stp = (_vp * self.nspc)(*[ctypes.cast(ctypes.byref(c.st), _vp) for c in self.ch])

    def set_config(self, spc_coms, spc_angs=None):
        """Molecules per species: lists of com [n,3] and angles [n,3]."""
        L = _gcmc_lib()
        for i in range(self.nspc):
            com = np.asarray(spc_coms[i], dtype=np.float64).reshape(-1, 3)
            ang = np.zeros_like(com) if spc_angs is None else np.asarray(spc_angs[i], dtype=np.float64).reshape(-1, 3)
            for m in range(len(com)):
                L.mo_gcmc_set_mol_mix(self.sysp, self.stp, i, m, np.ascontiguousarray(com[m]).ctypes.data,
                                      np.ascontiguousarray(ang[m]).ctypes.data)
            self.ch[i].st.N = len(com)
        L.mo_gcmc_recompute_sk_mix(self.sysp, self.nspc, self.stp)
        L.mo_gcmc_cav_rebuild(self.sysp, self.nspc, self.stp)
        self.ch[0].st.energy = self.total_energy()

    def run(self, nsteps):
        rc = _gcmc_lib().mo_gcmc_run_mix(self.sysp, self.nspc, self.stp, nsteps)
        if rc:
            raise RuntimeError("mo_gcmc_run_mix failed %d" % rc)

    def total_energy(self):
        return _gcmc_lib().mo_gcmc_total_energy_mix(self.sysp, self.nspc, self.stp)

    @property
    def energy(self):
        return self.ch[0].st.energy

    def N(self):
        return [c.st.N for c in self.ch]

    def trace_array(self, n=None):
        return self.ch[0].trace_array(n)


# ---------------------------------------------------------------- HMC
class HmcSys(ctypes.Structure):
    _fields_ = [("ns", _i), ("stype", _i * 8), ("mass", _d * 8), ("site_map", _i * 8), ("nmaps", _i),
                ("map_ptr", _vp * 8), ("map_vals", _vp * 8), ("map_npnts", (_i * 3) * 8),
                ("map_step", (_d * 3) * 8), ("map_nf", _ll * 8), ("full_ell", _d * 3), ("fund_ell", _d * 3),
                ("nst", _i), ("lj_A", _d * 64), ("lj_B", _d * 64), ("lj_h2", _d * 64), ("lj_l2", _d * 64),
                ("lj_on", _i * 64), ("nb", _i), ("bi", (_i * 2) * 16), ("bk", _d * 16), ("br0", _d * 16),
                ("na", _i), ("ai", (_i * 3) * 16), ("ak", _d * 16), ("at0", _d * 16), ("nt", _i),
                ("ti", (_i * 4) * 16), ("tc", (_d * 6) * 16), ("T", _d), ("dt", _d), ("nmd", _i),
                ("max_nrg_md", _d), ("max_nrg0", _d)]


class MoRng(ctypes.Structure):
    _fields_ = [("ix", ctypes.c_int32), ("iy", ctypes.c_int32), ("iseed", ctypes.c_int32),
                ("first", ctypes.c_int32)]


class HmcState(ctypes.Structure):
    _fields_ = [("N", _i), ("maxN", _i), ("rp", _vp), ("r", _vp), ("v", _vp), ("a", _vp), ("rng", MoRng),
                ("g_have", _i), ("g_z2", _d), ("energy", _d), ("att", _ll), ("acc", _ll), ("fail", _ll),
                ("last_dH", _d)]


def _hmc_lib():
    L = lib()
    if getattr(L, "_hmc_sig", False):
        return L
    L.mo_hmc_energy_forces.argtypes = [_vp, _i, _vp, _vp, _vp, ctypes.POINTER(_i)]
    L.mo_hmc_energy_forces.restype = _d
    L.mo_hmc_gaussian.argtypes = [_vp, _d, _d]
    L.mo_hmc_gaussian.restype = _d
    L.mo_hmc_move.argtypes = [_vp, _vp]
    L.mo_hmc_move.restype = _i
    L.mo_hmc_nve.argtypes = [_vp, _i, _vp, _vp, _vp, _i, _vp]
    L.mo_hmc_nve.restype = _i
    L.mo_hmc_md.argtypes = [_vp, _i, _vp, _vp, _vp, _i, _vp]
    L.mo_hmc_md.restype = _i
    L._hmc_sig = True
    return L


def hmc_sys(spec):
    """mo_hmc_sys from an HMC spec (keys: stype, mass, site_map, maps
    [(npnts, step, ptr, vals8)], full_ell, fund_ell, nst, lj (A, B, h2, l2,
    on), bonds [(i, j, k, r0)], angles [(i, j, l, k, theta0)], torsions
    [(i, j, k, l, c[6])], T, dt, nmd, max_nrg_md, max_nrg0)."""
    s = HmcSys()
    keep = []
    s.ns = ns = len(spec["stype"])
    for a in range(ns):
        s.stype[a] = int(spec["stype"][a])
        s.mass[a] = float(spec["mass"][a])
        s.site_map[a] = int(spec["site_map"][a])
    s.nmaps = len(spec["maps"])
    for j, (npn, step, ptr, vals) in enumerate(spec["maps"]):
        ptr = np.ascontiguousarray(ptr, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float32)
        keep += [ptr, vals]
        s.map_ptr[j] = ptr.ctypes.data
        s.map_vals[j] = vals.ctypes.data
        s.map_nf[j] = vals.shape[1]
        for i in range(3):
            s.map_npnts[j][i] = int(npn[i])
            s.map_step[j][i] = float(step[i])
    for i in range(3):
        s.full_ell[i] = spec["full_ell"][i]
        s.fund_ell[i] = spec["fund_ell"][i]
    s.nst = spec["nst"]
    A, B, h2, l2, on = spec["lj"]
    for t in range(s.nst * s.nst):
        s.lj_A[t], s.lj_B[t], s.lj_h2[t], s.lj_l2[t], s.lj_on[t] = A[t], B[t], h2[t], l2[t], int(on[t])
    s.nb = len(spec["bonds"])
    for q, (i, j, k, r0) in enumerate(spec["bonds"]):
        s.bi[q][0], s.bi[q][1], s.bk[q], s.br0[q] = i, j, k, r0
    s.na = len(spec["angles"])
    for q, (i, j, l, k, t0) in enumerate(spec["angles"]):
        s.ai[q][0], s.ai[q][1], s.ai[q][2], s.ak[q], s.at0[q] = i, j, l, k, t0
    s.nt = len(spec["torsions"])
    for q, (i, j, k, l, c) in enumerate(spec["torsions"]):
        s.ti[q][0], s.ti[q][1], s.ti[q][2], s.ti[q][3] = i, j, k, l
        for n in range(6):
            s.tc[q][n] = c[n]
    s.T, s.dt, s.nmd = spec["T"], spec["dt"], spec["nmd"]
    s.max_nrg_md = spec.get("max_nrg_md", 1000.0)
    s.max_nrg0 = spec.get("max_nrg0", 1.0e6)
    return s, keep


class Hmc:
    """One HMC chain of flexible molecules on the CPU oracle."""

    def __init__(self, spec, seed, maxN=256):
        L = _hmc_lib()
        self.sys, self.keep = hmc_sys(spec)
        ns = self.sys.ns
        self.ns = ns
        self.rp = np.zeros((maxN, ns, 3))
        self.r = np.zeros((maxN, ns, 3))
        self.v = np.zeros((maxN, ns, 3))
        self.a = np.zeros((maxN, ns, 3))
        st = HmcState()
        st.maxN = maxN
        st.rp, st.r, st.v, st.a = [x.ctypes.data for x in (self.rp, self.r, self.v, self.a)]
        L.mo_rng_init(ctypes.byref(st.rng), int(seed))
        self.st = st
        self.full = np.asarray(spec["full_ell"], dtype=np.float64)

    def set_config(self, rp):
        rp = np.asarray(rp, dtype=np.float64).reshape(-1, self.ns, 3)
        n = len(rp)
        self.rp[:n] = rp
        self.r[:n] = rp - np.floor(rp / self.full) * self.full
        self.st.N = n
        e, _, ok = self.energy_forces()
        self.st.energy = e
        return e, ok

    def energy_forces(self, rp=None, r=None):
        L = _hmc_lib()
        n = self.st.N if rp is None else len(rp)
        rp = self.rp[:n] if rp is None else np.ascontiguousarray(rp, dtype=np.float64)
        r = self.r[:n] if r is None else np.ascontiguousarray(r, dtype=np.float64)
        rp = np.ascontiguousarray(rp)
        r = np.ascontiguousarray(r)
        F = np.zeros((max(n, 1), self.ns, 3))
        ok = _i()
        e = L.mo_hmc_energy_forces(ctypes.byref(self.sys), n, rp.ctypes.data, r.ctypes.data, F.ctypes.data,
                                   ctypes.byref(ok))
        return e, F[:n], bool(ok.value)

    def move(self):
        return _hmc_lib().mo_hmc_move(ctypes.byref(self.sys), ctypes.byref(self.st))

    def gaussian(self, sigma=1.0):
        return _hmc_lib().mo_hmc_gaussian(ctypes.byref(self.st), 0.0, sigma)

    def md(self, v, nsteps):
        """Plain NVE MD from the current configuration and velocities v
        [N][ns][3] (mo_hmc_md); returns (steps done, trace [nsteps][10],
        final rp, r, v)."""
        n = self.st.N
        rp = np.ascontiguousarray(self.rp[:n].copy())
        r = np.ascontiguousarray(self.r[:n].copy())
        v = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(n, -1, 3).copy())
        tr = np.zeros((nsteps, 10))
        done = _hmc_lib().mo_hmc_md(ctypes.byref(self.sys), n, rp.ctypes.data, r.ctypes.data, v.ctypes.data, nsteps,
                                    tr.ctypes.data)
        return done, tr, rp, r, v

    def nve(self, v, nsteps):
        """Plain NVE trajectory from velocities v [N][ns][3]; total energy
        (kcal/mol) after each step."""
        n = self.st.N
        rp = np.ascontiguousarray(self.rp[:n].copy())
        r = np.ascontiguousarray(self.r[:n].copy())
        v = np.ascontiguousarray(np.asarray(v, dtype=np.float64).copy())
        et = np.zeros(nsteps + 1)
        ok = _hmc_lib().mo_hmc_nve(ctypes.byref(self.sys), n, rp.ctypes.data, r.ctypes.data, v.ctypes.data, nsteps,
                                   et.ctypes.data)
        return bool(ok), et
