"""Plain Coulomb between sorbate sites (coulomb_snglint, coulomb.F90:258-327)
in the oracle's GCMC energy, pinned without the reference (no fixture of
the reference uses it for a GCMC sorbate):

* the total energy of random CO2-like configurations equals a numpy sum of
  331.5 shield q_a q_b (1/r - 1/hicut) over the site pairs of different
  molecules within hicut (minimum image), with the LJ part off;
* with hicut = 0 the sum is the unshifted 1/r over every minimum-image pair;
* the shifted pair energy is continuous at hicut (a pair just inside adds
  ~0) and a pair closer than locut fails the evaluation;
* switched-off type pairs add nothing.
"""
import numpy as np
import pytest

from oracle import oracle as O
from musichip import synth

L = 24.0


def _spec(hicut, on=(1, 1, 1, 1), shield=1.0, locut=0.5):
    sorb = synth.co2_like()
    return dict(full_ell=[L] * 3, fund_ell=[L] * 3, ref=sorb["ref"], stype=[0, 1, 0], sq=sorb["q"], nst=2,
                lj=([0.0] * 4, [0.0] * 4, [0.0] * 4, [0.0] * 4, [0] * 4), T=300.0,
                coul=dict(on=list(on), hicut=hicut, locut=locut, shield=shield),
                moves=[("RINSERT", 1.0), ("RDELETE", 1.0), ("RTRANSLATE", 1.0), ("RROTATE", 1.0)],
                dtrans=0.3, drot=0.3)


def _numpy_energy(g, n, hicut, on, shield):
    """Site positions from the oracle state; the pair sum restated."""
    sq = np.array(synth.co2_like()["q"])
    st = [0, 1, 0]
    pos = np.array(g.r[:n]).reshape(n, 3, 3)
    e = 0.0
    for m in range(n):
        for k in range(m + 1, n):
            for a in range(3):
                for b in range(3):
                    if not on[st[a] * 2 + st[b]]:
                        continue
                    d = pos[m, a] - pos[k, b]
                    d = d - L * np.round(d / L)
                    rr = np.sqrt(d @ d)
                    if hicut > 0 and rr > hicut:
                        continue
                    e += 331.5 * shield * sq[a] * sq[b] * (1.0 / rr - (1.0 / hicut if hicut > 0 else 0.0))
    return e


def _random_config(n, seed):
    rng = np.random.default_rng(seed)
    # molecules on a jittered lattice so that no pair is closer than locut
    g = np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    com = (g + 0.5) * (L / 4) + rng.uniform(-1.0, 1.0, (n, 3))
    ang = np.column_stack([rng.uniform(-np.pi, np.pi, n), np.arccos(rng.uniform(-1, 1, n)),
                           rng.uniform(-np.pi, np.pi, n)])
    return com, ang


@pytest.mark.parametrize("hicut,on,shield", [(9.5, (1, 1, 1, 1), 1.0), (0.0, (1, 1, 1, 1), 1.0),
                                             (9.5, (1, 0, 0, 1), 0.7)])
def test_oracle_coulomb_energy_matches_numpy(hicut, on, shield):
    g = O.Gcmc(_spec(hicut, on, shield), 1000.0, 5)
    com, ang = _random_config(40, 3)
    g.set_config(com, ang)
    want = _numpy_energy(g, 40, hicut, on, shield)
    assert abs(want) > 1.0
    assert g.total_energy() == pytest.approx(want, rel=1e-12, abs=1e-12)


def test_oracle_coulomb_shift_and_locut():
    # two single-charge molecules along x: the shifted energy vanishes at hicut
    spec = _spec(9.5)
    spec.update(ref=[[0.0, 0.0, 0.0]], stype=[0], sq=[1.0], nst=1,
                lj=([0.0], [0.0], [0.0], [0.0], [0]), coul=dict(on=[1], hicut=9.5, locut=0.5, shield=1.0))
    g = O.Gcmc(spec, 1000.0, 5)
    for d, want in ((9.5 - 1e-9, 0.0), (5.0, 331.5 * (1 / 5.0 - 1 / 9.5)), (9.6, 0.0)):
        g.set_config([[1.0, 1.0, 1.0], [1.0 + d, 1.0, 1.0]])
        assert g.total_energy() == pytest.approx(want, abs=1e-7)
    g.set_config([[1.0, 1.0, 1.0], [1.4, 1.0, 1.0]])  # closer than locut: failed evaluation
    assert np.isnan(g.total_energy())


def test_oracle_coulomb_chain_runs_and_conserves_energy():
    """A chain with Coulomb only: the running energy equals a full recompute."""
    g = O.Gcmc(_spec(9.5), 20000.0, 11)
    g.run(3000)
    assert g.st.N > 5
    assert g.st.energy == pytest.approx(g.total_energy(), rel=1e-9, abs=1e-9)
