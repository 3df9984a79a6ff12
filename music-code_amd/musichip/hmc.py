"""Host side of the HMC bursts of flexible sorbates (mh_hmc_* in
include/music_hip.h): parameters from a model, batches of replicas.

Mirrors hybridmoves_integrate_boltz (hybridmoves.F90:186-241) driven by the
reference's hybrid MC loop (hybridmc.F90:530-594) for the NVT burst."""
import ctypes

import numpy as np

from . import constants as C
from . import hip as H
from .inputs import lj_AB

_d, _i, _vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
MD_NQ = 10  # MH_MD_NQ


class HmcParams(ctypes.Structure):
    _fields_ = [("nsites", _i), ("site_type", _i * 8), ("site_mass", _d * 8), ("site_map", _vp * 8),
                ("ntypes", _i), ("lj_A", _d * 64), ("lj_B", _d * 64), ("lj_hicut2", _d * 64),
                ("lj_locut2", _d * 64), ("lj_on", ctypes.c_uint8 * 64), ("nbonds", _i), ("bond", (_i * 2) * 16),
                ("bond_k", _d * 16), ("bond_r0", _d * 16), ("nangles", _i), ("angle", (_i * 3) * 16),
                ("angle_k", _d * 16), ("angle_theta0", _d * 16), ("ntorsions", _i), ("torsion", (_i * 4) * 16),
                ("torsion_c", (_d * 6) * 16), ("temperature", _d), ("dt_ps", _d), ("md_steps", _i),
                ("max_nrg_md", _d), ("max_nrg_start", _d), ("precision", _i)]


def hmc_spec(model, full_ell, fund_ell, hicut=12.8, T=300.0, dt=0.001, nmd=10):
    """HMC spec (shared with the oracle's Hmc) for a model of synth.butane_ua
    form; site_map indexes the maps by site type."""
    types = model["types"]
    nt = len(types)
    A, B = [], []
    for ti in types:
        for tj in types:
            e1, s1 = model["lj"][ti]
            e2, s2 = model["lj"][tj]
            a, b = lj_AB((s1 + s2) / 2.0, np.sqrt(e1 * e2))
            A.append(float(a))
            B.append(float(b))
    return dict(stype=list(model["stype"]), mass=list(model["mass"]), site_map=list(model["stype"]), maps=[],
                full_ell=list(full_ell), fund_ell=list(fund_ell), nst=nt,
                lj=(A, B, [hicut * hicut] * nt * nt, [C.LOW_DIST_CUTOFF ** 2] * nt * nt, [1] * nt * nt),
                bonds=list(model["bonds"]), angles=list(model["angles"]), torsions=list(model["torsions"]),
                T=T, dt=dt, nmd=nmd, max_nrg_md=1000.0, max_nrg0=1.0e6)


def _sig():
    L = H.lib()
    if getattr(L, "_hmc_sig", False):
        return L
    P = ctypes.POINTER
    sig = {
        "mh_hmc_create": [_vp, _vp, _i, _vp, _i, P(_vp)],
        "mh_hmc_destroy": [_vp],
        "mh_hmc_set_threads": [_vp, _i],
        "mh_hmc_set_config": [_vp, _i, _i, _vp],
        "mh_hmc_run": [_vp, _i],
        "mh_hmc_get_state": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
        "mh_hmc_get_config": [_vp, _i, P(_i), _vp, _vp],
        "mh_hmc_energy_forces": [_vp, _i, P(_d), _vp, P(_i)],
        "mh_hmc_set_velocities": [_vp, _i, _i, _vp],
        "mh_hmc_get_velocities": [_vp, _i, P(_i), _vp],
        "mh_hmc_md": [_vp, _i, _vp],
    }
    for n, a in sig.items():
        f = getattr(L, n)
        f.argtypes = a
        f.restype = _i
    L._hmc_sig = True
    return L


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def make_params(spec, maps_by_type, precision):
    p = HmcParams()
    ns = len(spec["stype"])
    p.nsites = ns
    for a in range(ns):
        p.site_type[a] = int(spec["stype"][a])
        p.site_mass[a] = float(spec["mass"][a])
        m = maps_by_type.get(int(spec["site_map"][a])) if maps_by_type else None
        p.site_map[a] = m.h.value if m is not None else None
    p.ntypes = spec["nst"]
    A, B, h2, l2, on = spec["lj"]
    for t in range(p.ntypes * p.ntypes):
        p.lj_A[t], p.lj_B[t], p.lj_hicut2[t], p.lj_locut2[t] = A[t], B[t], h2[t], l2[t]
        p.lj_on[t] = 1 if on[t] else 0
    p.nbonds = len(spec["bonds"])
    for q, (i, j, k, r0) in enumerate(spec["bonds"]):
        p.bond[q][0], p.bond[q][1], p.bond_k[q], p.bond_r0[q] = i, j, k, r0
    p.nangles = len(spec["angles"])
    for q, (i, j, l, k, t0) in enumerate(spec["angles"]):
        p.angle[q][0], p.angle[q][1], p.angle[q][2], p.angle_k[q], p.angle_theta0[q] = i, j, l, k, t0
    p.ntorsions = len(spec["torsions"])
    for q, (i, j, k, l, c) in enumerate(spec["torsions"]):
        p.torsion[q][0], p.torsion[q][1], p.torsion[q][2], p.torsion[q][3] = i, j, k, l
        for n in range(6):
            p.torsion_c[q][n] = c[n]
    p.temperature, p.dt_ps, p.md_steps = spec["T"], spec["dt"], spec["nmd"]
    p.max_nrg_md = spec.get("max_nrg_md", 1000.0)
    p.max_nrg_start = spec.get("max_nrg0", 1.0e6)
    p.precision = precision
    return p


class HmcBatch:
    """nrep HMC replicas advanced together (one workgroup each)."""

    def __init__(self, ctx, spec, maps_by_type, seeds, maxN, precision=H.MH_PREC_FP64, threads=None):
        L = _sig()
        self.ctx = ctx
        self.params = make_params(spec, maps_by_type, precision)
        self.maps = maps_by_type
        self.nrep = len(seeds)
        self.ns = self.params.nsites
        self.maxN = maxN
        sd = np.ascontiguousarray(seeds, dtype=np.int32)
        h = _vp()
        H.check(L.mh_hmc_create(ctx.h, ctypes.byref(self.params), self.nrep, _p(sd), maxN, ctypes.byref(h)),
                "mh_hmc_create")
        self.h = h
        H._LIVE["gcmc"].add(self)
        if threads:
            H.check(L.mh_hmc_set_threads(self.h, int(threads)), "mh_hmc_set_threads")

    def set_config(self, rep, rp):
        rp = np.ascontiguousarray(rp, dtype=np.float64).reshape(-1, self.ns, 3)
        H.check(_sig().mh_hmc_set_config(self.h, rep, len(rp), _p(rp)), "mh_hmc_set_config")

    def set_threads(self, n):
        """Workgroup size per replica (128, 256, 512 or 1024)."""
        H.check(_sig().mh_hmc_set_threads(self.h, int(n)), "mh_hmc_set_threads")

    def set_velocities(self, rep, v):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, self.ns, 3)
        H.check(_sig().mh_hmc_set_velocities(self.h, rep, len(v), _p(v)), "mh_hmc_set_velocities")

    def velocities(self, rep):
        n = ctypes.c_int()
        v = np.zeros((self.maxN, self.ns, 3))
        H.check(_sig().mh_hmc_get_velocities(self.h, rep, ctypes.byref(n), _p(v)), "mh_hmc_get_velocities")
        return v[:n.value].copy()

    def md(self, nsteps):
        """nsteps NVE velocity-Verlet steps; trace [nrep][nsteps][MD_NQ]."""
        tr = np.zeros((self.nrep, int(nsteps), MD_NQ))
        H.check(_sig().mh_hmc_md(self.h, int(nsteps), _p(tr)), "mh_hmc_md")
        return tr

    def run(self, nmoves):
        H.check(_sig().mh_hmc_run(self.h, int(nmoves)), "mh_hmc_run")

    def state(self):
        R = self.nrep
        N = np.zeros(R, dtype=np.int32)
        e, dH = np.zeros(R), np.zeros(R)
        att, acc, fail = (np.zeros(R, dtype=np.int64) for _ in range(3))
        ix, iy = np.zeros(R, dtype=np.int32), np.zeros(R, dtype=np.int32)
        H.check(_sig().mh_hmc_get_state(self.h, _p(N), _p(e), _p(att), _p(acc), _p(fail), _p(dH), _p(ix), _p(iy)),
                "mh_hmc_get_state")
        return dict(N=N, energy=e, att=att, acc=acc, fail=fail, last_dH=dH, ix=ix, iy=iy)

    def config(self, rep):
        n = ctypes.c_int()
        rp = np.zeros((self.maxN, self.ns, 3))
        r = np.zeros((self.maxN, self.ns, 3))
        H.check(_sig().mh_hmc_get_config(self.h, rep, ctypes.byref(n), _p(rp), _p(r)), "mh_hmc_get_config")
        return rp[:n.value].copy(), r[:n.value].copy()

    def energy_forces(self, rep):
        e, ok = ctypes.c_double(), ctypes.c_int()
        F = np.zeros((self.maxN, self.ns, 3))
        H.check(_sig().mh_hmc_energy_forces(self.h, rep, ctypes.byref(e), _p(F), ctypes.byref(ok)),
                "mh_hmc_energy_forces")
        n = len(self.config(rep)[0])
        return e.value, F[:n], bool(ok.value)

    def close(self):
        if self.h:
            _sig().mh_hmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
