"""ctypes binding of libmusic_hip.so (C-ABI declared in include/music_hip.h).

The library must be built (music-code_amd/Makefile, __graft_entry__.build()).
There is no CPU fallback: every compute call goes to the HIP device and a
missing library or device raises.
"""
import atexit
import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(os.path.dirname(HERE), "lib", "libmusic_hip.so")

MH_MAP_LJ, MH_MAP_EWALD, MH_MAP_BUCK = 1, 2, 3
MH_PREC_FP64, MH_PREC_FP32 = 0, 1
MH_EWALD_PARITY, MH_EWALD_PHYSICS = 0, 1

_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_i64 = ctypes.c_int64
_LIB = None
# live handles, released children first at interpreter exit (before the HIP
# runtime itself is torn down)
_LIVE = {"gcmc": weakref.WeakSet(), "bmap": weakref.WeakSet(), "map": weakref.WeakSet(),
         "ctx": weakref.WeakSet()}


@atexit.register
def _release_all():
    for kind in ("gcmc", "bmap", "map", "ctx"):
        for obj in list(_LIVE[kind]):
            try:
                obj.close()
            except Exception:
                pass


def real_space_cutoff_default(kappa):
    """FP32 Ewald real-space cutoff used by the library: erfc(kappa r) <= 1e-9
    at kappa r = 4.32 (mirrors real_space_cutoff in csrc/mh_mapbuild.hip)."""
    return 4.32 / kappa


class MusicHipError(RuntimeError):
    pass


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise MusicHipError("libmusic_hip.so not built (%s); run __graft_entry__.build()" % LIBPATH)
    L = ctypes.CDLL(LIBPATH)
    L.mh_last_error.restype = ctypes.c_char_p
    L.mh_version.restype = ctypes.c_char_p
    sig = {
        "mh_ctx_create": [_i, ctypes.POINTER(_vp)],
        "mh_ctx_destroy": [_vp],
        "mh_ctx_set_timing": [_vp, _i],
        "mh_ctx_get_timing": [_vp, ctypes.c_char_p, _i],
        "mh_framework_upload": [_vp, _i, _vp, _vp, _vp, _vp, _vp],
        "mh_ff_set_lj": [_vp, _i, _vp, _vp, _vp, _vp, _vp],
        "mh_ff_set_buck": [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp],
        "mh_ewald_set": [_vp, _d, _i, _d, _i, ctypes.POINTER(_i)],
        "mh_framework_set_cell": [_vp, _vp, _vp],
        "mh_ewald_get": [_vp, _vp, _vp, _vp, _vp],
        "mh_map_build": [_vp, _i, _i, _d, _d, _d, _i, ctypes.POINTER(_vp)],
        "mh_map_destroy": [_vp],
        "mh_map_info": [_vp, _vp, _vp, _vp, _vp],
        "mh_map_get": [_vp, _vp, _vp],
        "mh_map_get_full": [_vp, _vp, _vp],
        "mh_map_write": [_vp, ctypes.c_char_p, _i, ctypes.c_char_p, ctypes.c_char_p, _vp,
                         ctypes.c_float],
        "mh_map_load": [_vp, ctypes.c_char_p, _vp, ctypes.POINTER(_vp)],
        "mh_map_from_host": [_vp, _vp, _i64, _vp, _vp, _vp, ctypes.POINTER(_vp)],
        "mh_map_interp": [_vp, _vp, _i, _vp, _d, _vp, _vp],
        "mh_bmap_create": [_vp, _vp, _d, ctypes.POINTER(_vp)],
        "mh_bmap_destroy": [_vp],
        "mh_bmap_info": [_vp, _vp, _vp, _vp, _vp],
        "mh_bmap_get": [_vp, _vp, _vp, _vp],
        "mh_bmap_search": [_vp, _i, _vp, _vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = _i
    _LIB = L
    return L


def check(rc, what=""):
    if rc != 0:
        msg = lib().mh_last_error().decode(errors="replace")
        raise MusicHipError("%s failed (code %d): %s" % (what, rc, msg))


class Map:
    """A potential map held on the device (mh_map)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle
        _LIVE["map"].add(self)
        npn = np.zeros(3, dtype=np.int32)
        nf = ctypes.c_int64()
        counts = np.zeros(4, dtype=np.int64)
        step = np.zeros(3)
        check(lib().mh_map_info(self.h, _p(npn), ctypes.byref(nf), _p(counts), _p(step)), "mh_map_info")
        self.npnts, self.nfsize, self.counts, self.step = npn, nf.value, counts, step

    def get(self):
        ptr = np.zeros(int(np.prod(self.npnts.astype(np.int64))), dtype=np.int32)
        vals = np.zeros((8, self.nfsize), dtype=np.float32)
        check(lib().mh_map_get(self.h, _p(ptr), _p(vals)), "mh_map_get")
        return ptr, vals

    def get_full(self, want64=True):
        n = int(np.prod(self.npnts.astype(np.int64)))
        f32 = np.zeros((n, 8), dtype=np.float32)
        f64 = np.zeros((n, 8)) if want64 else None
        check(lib().mh_map_get_full(self.h, _p(f32), _p(f64)), "mh_map_get_full")
        return f32, f64

    def write(self, path, zatoms=(), pcut_kJ=0.0):
        names = b"".join(z["name"].encode()[:25].ljust(25) for z in zatoms)
        syms = b"".join(z["symbol"].encode()[:2].ljust(2) for z in zatoms)
        zp = np.array([[z.get(k, 0.0) for k in ("charge", "A12", "B6", "C2", "D0", "rcuthi", "rcutlo")]
                       for z in zatoms], dtype=np.float32).reshape(-1, 7)
        check(lib().mh_map_write(self.h, path.encode(), len(zatoms), names or None, syms or None,
                                 _p(zp) if len(zatoms) else None, float(pcut_kJ)), "mh_map_write")

    def interp(self, xyz, factor=1.0):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        n = len(xyz)
        e = np.zeros(n)
        ok = np.zeros(n, dtype=np.uint8)
        check(lib().mh_map_interp(self.ctx.h, self.h, n, _p(xyz), float(factor), _p(e), _p(ok)),
              "mh_map_interp")
        return ok.astype(bool), e

    def close(self):
        if self.h:
            lib().mh_map_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BiasMap:
    """Bias map of biased insertion / deletion on the device (mh_bmap), from
    a pmap at a bias temperature (bmap_init + bmap_cubeinit, bmap.F90:69-403)."""

    def __init__(self, pmap, temperature):
        self.ctx = pmap.ctx
        h = _vp()
        check(lib().mh_bmap_create(self.ctx.h, pmap.h, float(temperature), ctypes.byref(h)), "mh_bmap_create")
        self.h = h
        _LIVE["bmap"].add(self)
        self.pmap = pmap
        nt, nc, cw, ce = _i64(), _i64(), _d(), _d()
        check(lib().mh_bmap_info(self.h, ctypes.byref(nt), ctypes.byref(nc), ctypes.byref(cw), ctypes.byref(ce)),
              "mh_bmap_info")
        self.numtab, self.ncubelets, self.cum_wt, self.clamp_energy = nt.value, nc.value, cw.value, ce.value

    def get(self):
        """-> (cubelet ids, weights, cumulative probabilities) in list order"""
        cube = np.zeros(self.numtab, dtype=np.int32)
        wt = np.zeros(self.numtab)
        cum = np.zeros(self.numtab)
        check(lib().mh_bmap_get(self.h, _p(cube), _p(wt), _p(cum)), "mh_bmap_get")
        return cube, wt, cum

    def search(self, keys):
        keys = np.ascontiguousarray(keys, dtype=np.float64)
        out = np.zeros(len(keys), dtype=np.int64)
        check(lib().mh_bmap_search(self.h, len(keys), _p(keys), _p(out)), "mh_bmap_search")
        return out

    def close(self):
        if self.h:
            lib().mh_bmap_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One HIP device context (mh_ctx)."""

    def __init__(self, device=0):
        h = _vp()
        check(lib().mh_ctx_create(device, ctypes.byref(h)), "mh_ctx_create")
        self.h = h
        _LIVE["ctx"].add(self)

    def set_timing(self, on=True):
        check(lib().mh_ctx_set_timing(self.h, 1 if on else 0), "mh_ctx_set_timing")

    def timings(self):
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        check(lib().mh_ctx_get_timing(self.h, buf, len(buf)), "mh_ctx_get_timing")
        return json.loads(buf.value.decode())

    def upload_framework(self, xyz, types, q, fund_ell, full_ell):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        types = np.ascontiguousarray(types, dtype=np.int32)
        q = np.ascontiguousarray(q if q is not None else np.zeros(len(types)), dtype=np.float64)
        fe = np.ascontiguousarray(fund_ell, dtype=np.float64)
        fu = np.ascontiguousarray(full_ell, dtype=np.float64)
        check(lib().mh_framework_upload(self.h, len(types), _p(xyz), _p(types), _p(q), _p(fe), _p(fu)),
              "mh_framework_upload")

    def set_cell(self, angles_deg, eff):
        """Cell angles (degrees) and the effective map box of the framework
        just uploaded (mh_framework_set_cell)."""
        an = np.ascontiguousarray(angles_deg, dtype=np.float64)
        ef = np.ascontiguousarray(eff, dtype=np.float64)
        check(lib().mh_framework_set_cell(self.h, _p(an), _p(ef)), "mh_framework_set_cell")

    def set_lj(self, A, B, hicut2, locut2, on=None):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (A, B, hicut2, locut2)]
        o = np.ascontiguousarray(on if on is not None else np.ones(len(a[0])), dtype=np.uint8)
        check(lib().mh_ff_set_lj(self.h, len(a[0]), *[_p(v) for v in a], _p(o)), "mh_ff_set_lj")

    def set_buck(self, A, rho, C, D, hicut2, locut2):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (A, rho, C, D, hicut2, locut2)]
        check(lib().mh_ff_set_buck(self.h, len(a[0]), *[_p(v) for v in a]), "mh_ff_set_buck")

    def set_ewald(self, kappa_param, kmax, locut, sign_mode=MH_EWALD_PARITY):
        """Ewald parameters and structure factor; sign_mode MH_EWALD_PARITY
        (the reference's skwaldHOT signs) or MH_EWALD_PHYSICS."""
        nk = _i()
        check(lib().mh_ewald_set(self.h, float(kappa_param), int(kmax), float(locut), int(sign_mode),
                                 ctypes.byref(nk)), "mh_ewald_set")
        return nk.value

    def ewald_get(self, nk):
        k = np.zeros((nk, 3))
        s = np.zeros(nk)
        si = np.zeros(nk)
        kap = ctypes.c_double()
        check(lib().mh_ewald_get(self.h, _p(k), _p(s), _p(si), ctypes.byref(kap)), "mh_ewald_get")
        return dict(k=k, s=s, si=si, kappa=kap.value)

    def build_map(self, kind, precision, gridspace, hipotcut_kJ, probe_q=1.0, keep_full=False):
        h = _vp()
        check(lib().mh_map_build(self.h, int(kind), int(precision), float(probe_q), float(gridspace),
                                 float(hipotcut_kJ), 1 if keep_full else 0, ctypes.byref(h)),
              "mh_map_build")
        return Map(self, h)

    def load_map(self, path, fund_ell):
        h = _vp()
        fe = np.ascontiguousarray(fund_ell, dtype=np.float64)
        check(lib().mh_map_load(self.h, path.encode(), _p(fe), ctypes.byref(h)), "mh_map_load")
        return Map(self, h)

    def map_from_host(self, npnts, ptr, vals8, fund_ell):
        npn = np.ascontiguousarray(npnts, dtype=np.int32)
        ptr = np.ascontiguousarray(ptr, dtype=np.int32)
        vals8 = np.ascontiguousarray(vals8, dtype=np.float32)
        fe = np.ascontiguousarray(fund_ell, dtype=np.float64)
        h = _vp()
        check(lib().mh_map_from_host(self.h, _p(npn), vals8.shape[1], _p(ptr), _p(vals8), _p(fe),
                                     ctypes.byref(h)), "mh_map_from_host")
        return Map(self, h)

    def close(self):
        if self.h:
            lib().mh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
