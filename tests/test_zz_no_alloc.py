"""mh_map_build_static, the map build that allocates nothing (scratch in a
file-scope device array of at most 2^21 grid nodes, outputs in the caller's
device buffers), against mh_map_build: the same ptr array and records bit
for bit, two calls in a row, the second on a smaller grid (the scratch keeps
the first call's contents), plus the size query and the capacity errors."""
import ctypes

import numpy as np
import pytest

from musichip import hip as H
from musichip import synth

pytestmark = pytest.mark.gpu


def _fn():
    L = H.lib()
    f = L.mh_map_build_static
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                  ctypes.c_void_p]
    f.restype = ctypes.c_int
    return f


def _static(ctx, kind, prec, gs, cut, ptr=None, vals=None, cap=0):
    npn = np.zeros(3, dtype=np.int32)
    nf = ctypes.c_int64(0)
    counts = np.zeros(4, dtype=np.int64)
    rc = _fn()(ctx.h, kind, prec, gs, cut, ptr, vals, cap, npn.ctypes.data, ctypes.byref(nf), counts.ctypes.data)
    return rc, npn, nf.value, counts


class _Dev:
    """Caller-owned device buffer from the HIP runtime the library links."""

    def __init__(self, nbytes):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(nbytes)) == 0
        self.n = nbytes

    def fill(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert self.rt.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.n
        assert self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        if self.p:
            assert self.rt.hipFree(self.p) == 0
            self.p = ctypes.c_void_p()


def _framework(hip, L, n, seed):
    rng = np.random.default_rng(seed)
    Lv = np.array([L, L * 1.1, L * 0.9])
    xyz = rng.uniform(0, 1, (n, 3)) * Lv
    typ = rng.integers(0, synth.NTYPES_FW, n).astype(np.int32)
    eps, sig = rng.uniform(20, 80, synth.NTYPES_FW), rng.uniform(2.5, 3.8, synth.NTYPES_FW)
    hip.upload_framework(xyz, typ, None, Lv, Lv)
    hip.set_lj(*synth.lb_table(eps, sig, 148.0, 3.73, 0.45 * min(Lv)))


def test_static_map_build_equals_allocating_build(hip):
    # two calls in a row on one device, the second on a smaller grid
    for L, n, gs, prec in ((24.0, 400, 0.25, H.MH_PREC_FP32), (14.0, 150, 0.4, H.MH_PREC_FP64)):
        print("case: static map build, L %.1f, %d atoms, grid %.2f, precision %d" % (L, n, gs, prec), flush=True)
        _framework(hip, L, n, int(L * 10))
        m = hip.build_map(H.MH_MAP_LJ, prec, gs, 100.0)
        ptr_ref, vals_ref = m.get()
        rc, npn, _, _ = _static(hip, H.MH_MAP_LJ, prec, gs, 100.0)
        assert rc == 0 and list(npn) == list(m.npnts)
        nodes = int(np.prod(npn.astype(np.int64)))
        assert nodes <= 1 << 21
        ptr_d, vals_d = _Dev(4 * nodes), _Dev(32 * nodes)
        try:
            ptr_d.fill(np.full(nodes, -7, dtype=np.int32))
            vals_d.fill(np.full(8 * nodes, np.nan, dtype=np.float32))
            rc, npn, nf, counts = _static(hip, H.MH_MAP_LJ, prec, gs, 100.0, ptr_d.p.value, vals_d.p.value, nodes)
            assert rc == 0, H.lib().mh_last_error()
            assert nf == m.nfsize and list(counts) == list(m.counts)
            assert np.array_equal(ptr_d.get(np.int32, nodes), ptr_ref.ravel())
            v = vals_d.get(np.float32, 8 * nf).reshape(nf, 8).T
            assert np.array_equal(v, vals_ref)
            # a capacity below the tabulated count is refused before any record is written
            small = _Dev(32 * 4)
            try:
                rc, _, nf2, _ = _static(hip, H.MH_MAP_LJ, prec, gs, 100.0, ptr_d.p.value, small.p.value, 4)
                assert rc != 0 and nf2 == nf
            finally:
                small.free()
        finally:
            ptr_d.free()
            vals_d.free()
        m.close()


def test_static_map_build_refuses_oversized_grid(hip):
    _framework(hip, 40.0, 300, 3)
    rc, npn, _, _ = _static(hip, H.MH_MAP_LJ, H.MH_PREC_FP32, 0.2, 100.0)
    assert rc == 0 and int(np.prod(npn.astype(np.int64))) > 1 << 21
    dummy = ctypes.c_int32(0)
    rc, _, _, _ = _static(hip, H.MH_MAP_LJ, H.MH_PREC_FP32, 0.2, 100.0, ctypes.addressof(dummy),
                          ctypes.addressof(dummy), 1)
    assert rc != 0 and "static scratch" in H.lib().mh_last_error().decode()
