"""The oracle's Buckingham node (mo_buck_node; the reference's buck_snglint
stops before evaluating, buck.F90:303-307, so no reference output exists).

Pinned by what the mathematics fixes: with A = 0 and D = 0 the potential
-C r^-6 is the LJ node with A = 0, B = C, whose values and derivatives are
bit-exact against the reference build (test_oracle_ref.py), so every
column must agree; the exponential and r^-8 terms are pinned by central
finite differences of the node's own outputs (V -> first derivatives,
first -> mixed second, mixed second -> third), each on its own so that a
dropped term or a wrong sign fails."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

L = np.array([14.0, 15.0, 16.0])


def _fw(seed=3, n=40):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0, 1, (n, 3)) * L
    typ = rng.integers(0, 2, n).astype(np.int32)
    return np.ascontiguousarray(xyz), typ


def _buck(p, xyz, typ, A, rho, C, D, hicut=7.0, locut=0.5):
    out = np.zeros(8)
    f = lambda v: np.ascontiguousarray(np.asarray(v, dtype=np.float64))
    ok = O.lib().mo_buck_node(O._p(f(p)), len(xyz), O._p(xyz), O._p(typ), O._p(f(A)), O._p(f(rho)), O._p(f(C)),
                              O._p(f(D)), O._p(f([hicut ** 2] * 2)), O._p(f([locut ** 2] * 2)), O._p(f(L)),
                              O._p(out))
    return ok, out


def _lj(p, xyz, typ, A, B, hicut=7.0, locut=0.5):
    out = np.zeros(8)
    f = lambda v: np.ascontiguousarray(np.asarray(v, dtype=np.float64))
    ok = O.lib().mo_lj_node(O._p(f(p)), len(xyz), O._p(xyz), O._p(typ), O._p(f(A)), O._p(f(B)),
                            O._p(f([hicut ** 2] * 2)), O._p(f([locut ** 2] * 2)), O._p(f(L)), O._p(out))
    return ok, out


def _points(xyz, n, seed):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform(0, 1, 3) * L
        d = xyz - p
        d -= L * np.round(d / L)
        if np.sqrt((d ** 2).sum(1)).min() > 1.6:
            pts.append(p)
    return pts


def test_dispersion_term_equals_lj_node():
    xyz, typ = _fw()
    C = [950.0, 1400.0]
    for p in _points(xyz, 30, 1):
        okb, b = _buck(p, xyz, typ, [0.0, 0.0], [0.3, 0.3], C, [0.0, 0.0])
        okl, l = _lj(p, xyz, typ, [0.0, 0.0], C)
        assert okb == okl == 1
        assert np.allclose(b, l, rtol=1e-12, atol=1e-12 * np.abs(l).max()), (b, l)


@pytest.mark.parametrize("A,rho,C,D", [([9.0e4, 6.0e4], [0.28, 0.33], [0.0, 0.0], [0.0, 0.0]),
                                       ([0.0, 0.0], [0.3, 0.3], [0.0, 0.0], [4000.0, 2500.0]),
                                       ([9.0e4, 6.0e4], [0.28, 0.33], [950.0, 1400.0], [4000.0, 2500.0])])
def test_derivatives_are_finite_differences(A, rho, C, D):
    xyz, typ = _fw(5)
    h = 1e-4
    e = np.eye(3)
    for p in _points(xyz, 12, 2):
        ok, v = _buck(p, xyz, typ, A, rho, C, D)
        assert ok
        side = lambda d: _buck(p + d, xyz, typ, A, rho, C, D)[1]
        scale = max(np.abs(v).max(), 1e-3)
        # first derivatives from V
        for i in range(3):
            fd = (side(h * e[i])[0] - side(-h * e[i])[0]) / (2 * h)
            assert abs(fd - v[1 + i]) <= 1e-6 * scale, (i, fd, v[1 + i])
        # mixed second derivatives: d/dy of dV/dx, d/dz of dV/dx, d/dz of dV/dy
        for col, (src, ax) in zip((4, 5, 6), ((1, 1), (1, 2), (2, 2))):
            fd = (side(h * e[ax])[src] - side(-h * e[ax])[src]) / (2 * h)
            assert abs(fd - v[col]) <= 1e-5 * scale, (col, fd, v[col])
        # third: d/dz of d2V/dxdy
        fd = (side(h * e[2])[4] - side(-h * e[2])[4]) / (2 * h)
        assert abs(fd - v[7]) <= 1e-4 * scale, (fd, v[7])


def test_cutoffs_and_failure():
    xyz, typ = _fw(7, 10)
    p = xyz[0] + np.array([0.3, 0.0, 0.0])
    ok, v = _buck(p, xyz, typ, [9.0e4, 6.0e4], [0.28, 0.33], [950.0, 1400.0], [0.0, 0.0])
    assert ok == 0 and v[0] == 1e10
    # beyond hicut nothing counts: a hicut below every distance gives zero
    q = _points(xyz, 1, 9)[0]
    ok, v = _buck(q, xyz, typ, [9.0e4, 6.0e4], [0.28, 0.33], [950.0, 1400.0], [0.0, 0.0], hicut=1.0, locut=0.1)
    assert ok and np.all(v == 0.0)
