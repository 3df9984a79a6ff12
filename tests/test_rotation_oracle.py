"""Orientation draws of the oracle's rigid moves (rigidmoves_rinsert's
random orientation, rigidmoves.F90:1313-1367, and rigidmoves_rotperturb,
:1675-1720) pinned by what the mathematics fixes, with no interaction so
that every trial is accepted: inserted orientations are uniform on the
rotation group (cos theta, phi and psi uniform), and rotation steps keep a
uniform ensemble uniform (a symmetric proposal in cos theta, phi, psi; a
step in theta instead of cos theta, or a missing wrap, drifts it).  The
Euler matrix itself is bit-exact against the reference build
(test_oracle_ref_moves.py)."""
import numpy as np
from scipy import stats

from oracle import oracle as O

L = np.array([30.0, 30.0, 30.0])
BENT = [[0.0, 0.065, 0.0], [0.8165, -0.5128, 0.0], [-0.8165, -0.5128, 0.0]]


def _free(moves, drot=0.0):
    # no pair table entry on, no maps: every energy is zero
    return dict(full_ell=L, fund_ell=L, nst=1, lj=([0.0], [0.0], [0.0], [0.0], [0]), T=300.0, ref=BENT,
                stype=[0, 0, 0], sq=[0.0, 0.0, 0.0], moves=moves, drot=drot)


def _uniform_checks(ang, p_min=1e-3):
    phi, theta, psi = ang[:, 0], ang[:, 1], ang[:, 2]
    assert stats.kstest(np.cos(theta), "uniform", args=(-1, 2)).pvalue > p_min
    assert stats.kstest(phi, "uniform", args=(-np.pi, 2 * np.pi)).pvalue > p_min
    assert stats.kstest(psi, "uniform", args=(-np.pi, 2 * np.pi)).pvalue > p_min


def test_inserted_orientations_are_uniform():
    g = O.Gcmc(_free([("RINSERT", 1.0)]), 1.0e7, 17, maxN=3000)
    g.run(2500)
    n = g.st.N
    assert n == 2500  # nothing interacts: every insertion is accepted
    ang = g.ang[:n]
    assert np.all((ang[:, 1] >= 0) & (ang[:, 1] <= np.pi))
    _uniform_checks(ang)
    # the body axes then average to zero and their squares to 1/3
    z = np.stack([np.sin(ang[:, 1]) * np.sin(ang[:, 0]), -np.sin(ang[:, 1]) * np.cos(ang[:, 0]),
                  np.cos(ang[:, 1])], 1)
    assert np.all(np.abs(z.mean(0)) < 0.05) and np.all(np.abs((z ** 2).mean(0) - 1 / 3) < 0.03)


def test_rotation_steps_keep_a_uniform_ensemble_uniform():
    rng = np.random.default_rng(5)
    n = 300
    com = rng.uniform(0, 1, (n, 3)) * L
    ang = np.stack([rng.uniform(-np.pi, np.pi, n), np.arccos(rng.uniform(-1, 1, n)),
                    rng.uniform(-np.pi, np.pi, n)], 1)
    g = O.Gcmc(_free([("RROTATE", 1.0)], drot=0.9), 1.0, 23, maxN=n)
    g.set_config(com, ang)
    g.run(120000)
    assert int(g.st.acc[3]) == int(g.st.att[3]) == 120000
    out = g.ang[:n]
    assert np.all((out[:, 1] >= 0) & (out[:, 1] <= np.pi))
    assert np.all(np.abs(out[:, 0]) <= np.pi + 1e-12) and np.all(np.abs(out[:, 2]) <= np.pi + 1e-12)
    _uniform_checks(out)
    # the sites keep their distances under every rotation
    for m in range(0, n, 40):
        rp = O.euler_xyz(np.asarray(BENT), g.com[m], out[m])
        d = np.linalg.norm(rp[:, None] - rp[None], axis=-1)
        d0 = np.linalg.norm(np.asarray(BENT)[:, None] - np.asarray(BENT)[None], axis=-1)
        assert np.allclose(d, d0, atol=1e-12)
