"""Host side of the device-resident GCMC (mh_gcmc_* in include/music_hip.h):
builds mh_gcmc_params, runs batches of replicas and reads statistics.

Mirrors the reference's GCMC driver pieces: move set and weights
(mcmoveset.F90:280-335), Adams B from fugacity (gcmc.F90:266-347), block /
cumulative loading statistics (stats.F90:152-195, gcmcmoves.F90:371-382).
"""
import ctypes

import numpy as np

from . import constants as C
from . import hip as H

MH_MV_INSERT, MH_MV_DELETE, MH_MV_TRANSLATE, MH_MV_ROTATE, MH_MV_BINSERT, MH_MV_BDELETE = 0, 1, 2, 3, 4, 5
MOVES = {"RINSERT": 0, "RDELETE": 1, "RTRANSLATE": 2, "RROTATE": 3,
         "INSERT": 0, "DELETE": 1, "TRANSLATE": 2, "ROTATE": 3, "BINSERT": 4, "BDELETE": 5}

_d, _i, _vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p


class GcmcParams(ctypes.Structure):
    _fields_ = [("nsites", _i), ("ref", (_d * 3) * 8), ("site_type", _i * 8), ("site_q", _d * 8),
                ("site_pmap", _vp * 8), ("site_emap", _vp * 8), ("ntypes", _i),
                ("lj_A", _d * 64), ("lj_B", _d * 64), ("lj_hicut2", _d * 64), ("lj_locut2", _d * 64),
                ("lj_on", ctypes.c_uint8 * 64), ("ewald", _i), ("ewald_kappa", _d), ("ewald_kmax", _i),
                ("ewald_rcut", _d), ("temperature", _d), ("nmoves", _i), ("move_type", _i * 4),
                ("move_weight", _d * 4), ("delta_trans", _d), ("delta_rot", _d), ("max_nrg", _d),
                ("precision", _i), ("bias", _vp), ("coul", _i), ("coul_on", ctypes.c_uint8 * 64),
                ("coul_hicut", _d), ("coul_locut", _d), ("coul_shield", _d),
                ("adjust_trans", _i), ("adjust_rot", _i),
                ("lj_comcut", _d * 4), ("coul_comcut", _d * 4)]


class Cavity(ctypes.Structure):
    """mh_cavity: cube grid, tabulation, listed species, trials per species."""
    _fields_ = [("n", _i * 3), ("tab", _vp), ("species_mask", _i), ("ntrials", _i * 4), ("map_bias", _i * 4)]


def _sig():
    L = H.lib()
    if getattr(L, "_gcmc_sig", False):
        return L
    P = ctypes.POINTER
    sig = {
        "mh_gcmc_create": [_vp, _vp, _i, _vp, _vp, _i, P(_vp)],
        "mh_gcmc_create_mix": [_vp, _vp, _i, _i, _vp, _vp, _i, P(_vp)],
        "mh_gcmc_species": [_vp, _i],
        "mh_gcmc_destroy": [_vp],
        "mh_gcmc_set_threads": [_vp, _i],
        "mh_gcmc_set_config": [_vp, _i, _i, _vp, _vp],
        "mh_gcmc_run": [_vp, _i],
        "mh_gcmc_run_cycles": [_vp, _i, _i],
        "mh_gcmc_get_state": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
        "mh_gcmc_set_rng": [_vp, _i, ctypes.c_int32, ctypes.c_int32],
        "mh_gcmc_get_deltas": [_vp, _vp],
        "mh_gcmc_set_deltas": [_vp, _i, _d, _d],
        "mh_gcmc_set_counts": [_vp, _i, _vp, _vp],
        "mh_gcmc_get_moments": [_vp, _vp, _vp],
        "mh_gcmc_get_config": [_vp, _i, P(_i), _vp, _vp, _vp],
        "mh_gcmc_total_energy": [_vp, _vp, _i],
        "mh_gcmc_set_trace": [_vp, _i],
        "mh_gcmc_get_trace": [_vp, _i, _vp, _vp, _vp, _vp, P(_i)],
        "mh_trial_energy": [_vp, _i, _i, _vp, _vp, _vp, _vp],
        "mh_gcmc_autotune": [_vp, _i, P(_i), P(_i)],
        "mh_pool_create": [_vp, _i, _i, _i, _i, _i, _i, P(_vp)],
        "mh_pool_admit": [_vp, _vp, _vp, _vp, _i, _vp, _vp, _i, P(_i)],
        "mh_pool_launch": [_vp, _i],
        "mh_pool_wait": [_vp],
        "mh_pool_status": [_vp, _i, P(_i), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
        "mh_pool_release": [_vp, _i],
        "mh_gcmc_set_cavity": [_vp, _vp],
        "mh_gcmc_get_cavity": [_vp, _i, _i, _vp, _vp, _vp],
        "mh_cavity_tab": [_vp, _d, _vp, _vp, _vp, _vp],
    }
    for n, a in sig.items():
        f = getattr(L, n)
        f.argtypes = a
        f.restype = _i
    L._gcmc_sig = True
    return L


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def make_params(spec, maps=None, precision=H.MH_PREC_FP64):
    """spec as for the oracle's Gcmc (ref, stype, sq, nst, lj, ewald or
    coul, T, moves, dtrans, drot); maps: dict site index -> (pmap Map or
    None, emap Map or None)."""
    p = GcmcParams()
    ref = np.asarray(spec["ref"], dtype=np.float64).reshape(-1, 3)
    p.nsites = len(ref)
    for a in range(len(ref)):
        for i in range(3):
            p.ref[a][i] = ref[a][i]
        p.site_type[a] = int(spec["stype"][a])
        p.site_q[a] = float(spec["sq"][a])
        if maps and a in maps:
            pm, em = maps[a]
            p.site_pmap[a] = pm.h.value if pm is not None else None
            p.site_emap[a] = em.h.value if em is not None else None
    p.ntypes = spec["nst"]
    A, B, h2, l2, on = spec["lj"]
    for t in range(p.ntypes * p.ntypes):
        p.lj_A[t], p.lj_B[t], p.lj_hicut2[t], p.lj_locut2[t] = A[t], B[t], h2[t], l2[t]
        p.lj_on[t] = 1 if on[t] else 0
    ew = spec.get("ewald")
    if ew:
        p.ewald, p.ewald_kappa, p.ewald_kmax, p.ewald_rcut = 1, ew["kappa"], ew["kmax"], ew["rc"]
    cl = spec.get("coul")  # plain Coulomb: dict(on=[nst*nst], hicut, locut, shield)
    if cl:
        p.coul = 1
        for t in range(p.ntypes * p.ntypes):
            p.coul_on[t] = 1 if cl["on"][t] else 0
        p.coul_hicut, p.coul_locut, p.coul_shield = cl["hicut"], cl["locut"], cl.get("shield", 1.0)
    p.temperature = spec["T"]
    p.nmoves = len(spec["moves"])
    for i, (t, w) in enumerate(spec["moves"]):
        p.move_type[i] = MOVES[t] if isinstance(t, str) else int(t)
        p.move_weight[i] = w
    p.delta_trans = spec.get("dtrans", 0.0)
    p.delta_rot = spec.get("drot", 0.0)
    p.adjust_trans = int(spec.get("adjust_trans", 0))
    p.adjust_rot = int(spec.get("adjust_rot", 0))
    # centre-of-mass cutoffs (ssbasic COMCUT@r): dict(lj=r, coul=r), a radius
    # for every partner species or a list per partner species
    cc = spec.get("comcut") or {}
    for key, fld in (("lj", p.lj_comcut), ("coul", p.coul_comcut)):
        v = cc.get(key, 0.0)
        v = list(v) if np.ndim(v) else [v] * 4
        for i, r in enumerate(v):
            fld[i] = float(r or 0.0)
    p.max_nrg = spec.get("max_nrg", C.VERY_LRG_NRG)
    p.precision = precision
    bias = spec.get("bias_map")  # a hip.BiasMap for BINSERT / BDELETE
    p.bias = bias.h.value if bias is not None else None
    return p


class GcmcBatch:
    """nrep replicas advanced together on the device (one workgroup each)."""

    def __init__(self, ctx, spec, fugacities, seeds, maxN=1024, maps=None,
                 precision=H.MH_PREC_FP64, threads=None):
        L = _sig()
        self.ctx = ctx
        self.params = make_params(spec, maps, precision)
        self.nrep = len(fugacities)
        self.ns = self.params.nsites
        self.maxN = maxN
        f = np.ascontiguousarray(fugacities, dtype=np.float64)
        s = np.ascontiguousarray(seeds, dtype=np.int32)
        h = _vp()
        H.check(L.mh_gcmc_create(ctx.h, ctypes.byref(self.params), self.nrep, _p(f), _p(s), maxN,
                                 ctypes.byref(h)), "mh_gcmc_create")
        self.h = h
        H._LIVE["gcmc"].add(self)
        self.maps = maps  # keep the map objects alive
        self.bias_map = spec.get("bias_map")
        if threads:
            self.set_threads(threads)

    def set_threads(self, n):
        H.check(_sig().mh_gcmc_set_threads(self.h, int(n)), "mh_gcmc_set_threads")

    def autotune(self, probe_steps=200):
        """Fastest (threads, table mode) for this batch; the state is unchanged."""
        t, x = ctypes.c_int(), ctypes.c_int()
        H.check(_sig().mh_gcmc_autotune(self.h, int(probe_steps), ctypes.byref(t), ctypes.byref(x)),
                "mh_gcmc_autotune")
        return t.value, x.value

    def set_config(self, rep, com, ang=None):
        com = np.ascontiguousarray(com, dtype=np.float64).reshape(-1, 3)
        ang = None if ang is None else np.ascontiguousarray(ang, dtype=np.float64).reshape(-1, 3)
        H.check(_sig().mh_gcmc_set_config(self.h, rep, len(com), _p(com), _p(ang)), "mh_gcmc_set_config")

    def run(self, nsteps):
        H.check(_sig().mh_gcmc_run(self.h, int(nsteps)), "mh_gcmc_run")

    def run_cycles(self, nsteps, ncycles):
        """ncycles cycles of nsteps steps replayed from one captured hipGraph."""
        H.check(_sig().mh_gcmc_run_cycles(self.h, int(nsteps), int(ncycles)), "mh_gcmc_run_cycles")

    def state(self):
        R = self.nrep
        N = np.zeros(R, dtype=np.int32)
        e = np.zeros(R)
        sN = np.zeros(R)
        ns = np.zeros(R, dtype=np.int64)
        att = np.zeros((R, 4), dtype=np.int64)
        acc = np.zeros((R, 4), dtype=np.int64)
        ix = np.zeros(R, dtype=np.int32)
        iy = np.zeros(R, dtype=np.int32)
        H.check(_sig().mh_gcmc_get_state(self.h, _p(N), _p(e), _p(sN), _p(ns), _p(att), _p(acc), _p(ix),
                                         _p(iy)), "mh_gcmc_get_state")
        sN2 = np.zeros(R)
        H.check(_sig().mh_gcmc_get_moments(self.h, None, _p(sN2)), "mh_gcmc_get_moments")
        return dict(N=N, energy=e, sumN=sN, sumN2=sN2, nsteps=ns, att=att, acc=acc, ix=ix, iy=iy)

    def deltas(self):
        """Displacements in use, [nrep][2] = (translate A, rotate rad)."""
        d = np.zeros((self.nrep, 2))
        H.check(_sig().mh_gcmc_get_deltas(self.h, _p(d)), "mh_gcmc_get_deltas")
        return d

    def set_deltas(self, rep, dtrans, drot):
        H.check(_sig().mh_gcmc_set_deltas(self.h, rep, float(dtrans), float(drot)), "mh_gcmc_set_deltas")

    def set_counts(self, rep, att, acc):
        """Restore a replica's attempt / acceptance counts by move type."""
        a = np.ascontiguousarray(att, dtype=np.int64)
        c = np.ascontiguousarray(acc, dtype=np.int64)
        H.check(_sig().mh_gcmc_set_counts(self.h, rep, _p(a), _p(c)), "mh_gcmc_set_counts")

    def set_rng(self, rep, ix, iy):
        H.check(_sig().mh_gcmc_set_rng(self.h, rep, int(ix), int(iy)), "mh_gcmc_set_rng")

    def config(self, rep):
        n = ctypes.c_int()
        com = np.zeros((self.maxN, 3))
        ang = np.zeros((self.maxN, 3))
        r = np.zeros((self.maxN, self.ns, 3))
        H.check(_sig().mh_gcmc_get_config(self.h, rep, ctypes.byref(n), _p(com), _p(ang), _p(r)),
                "mh_gcmc_get_config")
        k = n.value
        return com[:k].copy(), ang[:k].copy(), r[:k].copy()

    def total_energy(self, refresh=False):
        e = np.zeros(self.nrep)
        H.check(_sig().mh_gcmc_total_energy(self.h, _p(e), 1 if refresh else 0), "mh_gcmc_total_energy")
        return e

    def set_trace(self, cap):
        H.check(_sig().mh_gcmc_set_trace(self.h, int(cap)), "mh_gcmc_set_trace")

    def trace(self, rep):
        n = ctypes.c_int()
        _sig().mh_gcmc_get_trace(self.h, rep, None, None, None, None, ctypes.byref(n))
        c = n.value
        mv = np.zeros(c, dtype=np.int32)
        ac = np.zeros(c, dtype=np.int32)
        N = np.zeros(c, dtype=np.int32)
        dE = np.zeros(c)
        H.check(_sig().mh_gcmc_get_trace(self.h, rep, _p(mv), _p(ac), _p(N), _p(dE), ctypes.byref(n)),
                "mh_gcmc_get_trace")
        return np.stack([mv, ac, N, dE], 1)

    def trial_energy(self, rep, com, ang=None):
        com = np.ascontiguousarray(com, dtype=np.float64).reshape(-1, 3)
        ang = np.zeros_like(com) if ang is None else np.ascontiguousarray(ang, dtype=np.float64).reshape(-1, 3)
        n = len(com)
        e = np.zeros(n)
        ok = np.zeros(n, dtype=np.uint8)
        H.check(_sig().mh_trial_energy(self.h, rep, n, _p(com), _p(ang), _p(e), _p(ok)), "mh_trial_energy")
        return ok.astype(bool), e

    def set_cavity(self, n, ntrials, map_bias=None, tab=None, species_mask=1):
        """Cavity-biased BINSERT / BDELETE (mh_gcmc_set_cavity): n cubes per
        axis, tab (uint8 [n0, n1, n2], None = all tabulated), ntrials and
        map_bias per species, species_mask of the species occupying cubes."""
        c = Cavity()
        for i in range(3):
            c.n[i] = int(n[i])
        nt = list(ntrials) if np.ndim(ntrials) else [int(ntrials)]
        mb = list(map_bias) if map_bias is not None and np.ndim(map_bias) else [int(map_bias or 0)] * len(nt)
        for s_, (t, m) in enumerate(zip(nt, mb)):
            c.ntrials[s_], c.map_bias[s_] = int(t), int(m)
        self.cav_tab = None if tab is None else np.ascontiguousarray(tab, dtype=np.uint8).reshape(-1)
        c.tab = _p(self.cav_tab)
        c.species_mask = int(species_mask)
        self.cav_n = tuple(int(v) for v in n)
        H.check(_sig().mh_gcmc_set_cavity(self.h, ctypes.byref(c)), "mh_gcmc_set_cavity")

    def cavity(self, rep, cap):
        """(cav_prob averages [cap], update counts [cap], cube site counts) of
        the selected species in replica rep."""
        avg = np.zeros(cap)
        n = np.zeros(cap, dtype=np.int64)
        cubes = np.zeros(int(np.prod(self.cav_n)), dtype=np.int32)
        H.check(_sig().mh_gcmc_get_cavity(self.h, rep, cap, _p(avg), _p(n), _p(cubes)), "mh_gcmc_get_cavity")
        return avg, n, cubes

    def close(self):
        if self.h:
            _sig().mh_gcmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SHARED_KEYS = ("nst", "lj", "coul", "ewald", "T", "full_ell", "fund_ell")


class GcmcMix(GcmcBatch):
    """Replicas of a mixture (gcmc_dosim's species draw every step,
    gcmc.F90:240-242): specs[i] and maps[i] per species, the pair tables,
    Coulomb / Ewald settings and temperature from specs[0] (site types
    index its tables); fugacities [nrep][nspc].  The per-species calls
    (state, config, set_config, deltas, counts, trial_energy) address the
    species chosen with select()."""

    def __init__(self, ctx, specs, fugacities, seeds, maxN=1024, maps=None,
                 precision=H.MH_PREC_FP64, threads=None):
        L = _sig()
        self.ctx = ctx
        self.nspc = len(specs)
        shared = {k: specs[0][k] for k in SHARED_KEYS if k in specs[0]}
        maps = maps or [None] * self.nspc
        plist = [make_params(dict(sp, **shared), maps[i], precision) for i, sp in enumerate(specs)]
        self.params = (GcmcParams * self.nspc)(*plist)
        self.ns_list = [p.nsites for p in plist]
        f = np.ascontiguousarray(fugacities, dtype=np.float64).reshape(-1, self.nspc)
        self.nrep = len(f)
        self.ns = self.ns_list[0]
        self.maxN = maxN
        s = np.ascontiguousarray(seeds, dtype=np.int32)
        h = _vp()
        H.check(L.mh_gcmc_create_mix(ctx.h, self.params, self.nspc, self.nrep, _p(f), _p(s), maxN,
                                     ctypes.byref(h)), "mh_gcmc_create_mix")
        self.h = h
        H._LIVE["gcmc"].add(self)
        self.maps = maps
        self.bias_map = None
        if threads:
            self.set_threads(threads)

    def select(self, spc):
        H.check(_sig().mh_gcmc_species(self.h, int(spc)), "mh_gcmc_species")
        self.ns = self.ns_list[spc]


class Pool(GcmcBatch):
    """Slot pool for continuous batching (mh_pool_*): requests of different
    systems share one persistent launch; admitted and released between
    launches."""

    def __init__(self, ctx, nslots, nsys, maxN=256, kmax_cap=0, max_sites=1, precision=H.MH_PREC_FP32,
                 threads=None):
        L = _sig()
        self.ctx = ctx
        self.nrep = nslots
        self.ns = max_sites
        self.maxN = maxN
        self.keep = {}  # ticket -> (params, maps) kept alive while admitted
        h = _vp()
        H.check(L.mh_pool_create(ctx.h, nslots, nsys, maxN, kmax_cap, max_sites, precision, ctypes.byref(h)),
                "mh_pool_create")
        self.h = h
        H._LIVE["gcmc"].add(self)
        if threads:
            self.set_threads(threads)

    def admit(self, spec, maps, fund_ell, full_ell, fugacities, seeds, nsteps):
        """Ticket of the admitted request, or None when the pool is full."""
        params = make_params(spec, maps, H.MH_PREC_FP32)
        f = np.ascontiguousarray(fugacities, dtype=np.float64)
        sd = np.ascontiguousarray(seeds, dtype=np.int32)
        fe = np.ascontiguousarray(fund_ell, dtype=np.float64)
        fu = np.ascontiguousarray(full_ell, dtype=np.float64)
        t = ctypes.c_int()
        rc = _sig().mh_pool_admit(self.h, ctypes.byref(params), _p(fe), _p(fu), len(f), _p(f), _p(sd),
                                  int(nsteps), ctypes.byref(t))
        if rc == 6:  # MH_ECAPACITY
            return None
        H.check(rc, "mh_pool_admit")
        self.keep[t.value] = (params, maps)
        return t.value

    def launch(self, nsteps):
        H.check(_sig().mh_pool_launch(self.h, int(nsteps)), "mh_pool_launch")

    def wait(self):
        H.check(_sig().mh_pool_wait(self.h), "mh_pool_wait")

    def status(self, ticket):
        n = ctypes.c_int()
        cap = self.nrep
        sl, rem, err, N = (np.zeros(cap, dtype=np.int32) for _ in range(4))
        e, sN = np.zeros(cap), np.zeros(cap)
        ns = np.zeros(cap, dtype=np.int64)
        att = np.zeros((cap, 4), dtype=np.int64)
        acc = np.zeros((cap, 4), dtype=np.int64)
        H.check(_sig().mh_pool_status(self.h, int(ticket), ctypes.byref(n), _p(sl), _p(rem), _p(err), _p(N), _p(e),
                                      _p(sN), _p(ns), _p(att), _p(acc)), "mh_pool_status")
        k = n.value
        return dict(slots=sl[:k], remaining=rem[:k], err=err[:k], N=N[:k], energy=e[:k], sumN=sN[:k],
                    nsteps=ns[:k], att=att[:k], acc=acc[:k])

    def release(self, ticket):
        H.check(_sig().mh_pool_release(self.h, int(ticket)), "mh_pool_release")
        self.keep.pop(ticket, None)


def cavity_tab(pmap, ucutoff, n, ncell, full_ell):
    """mh_cavity_tab: the cavity list's tabulation from a pmap (cubes holding
    a map cubelet centre below ucutoff kcal/mol), uint8 [n0, n1, n2]."""
    tab = np.zeros(tuple(int(v) for v in n), dtype=np.uint8)
    nn = np.ascontiguousarray(n, dtype=np.int32)
    nc = np.ascontiguousarray(ncell, dtype=np.int32)
    fe = np.ascontiguousarray(full_ell, dtype=np.float64)
    H.check(_sig().mh_cavity_tab(pmap.h, float(ucutoff), _p(nn), _p(nc), _p(fe), _p(tab)), "mh_cavity_tab")
    return tab


def adams_B(fug_kPa, volume, T, Zig=1.0):
    """gcmc_genfuglist (gcmc.F90:303-319)."""
    sivolume = volume * 1.0e-30 * C.NAV
    return np.log(fug_kPa * 1.0e3 * sivolume / (C.RGAS * T)) - np.log(Zig)
