"""Centre-of-mass cutoffs (ssbasic COMCUT@r, ssbasic.F90:745-756) on the
device: GCMC of the bent 3-site sorbate (geometric centre 0.54 A off its
centre of mass) with LJ and plain Coulomb cut at the molecules' centres,
FP64 trajectories identical move for move to the oracle's for molecules in
LDS and in HBM and from a configuration placed by the host; a mixture with a
cutoff only between its two species; FP32 trial energies within 1e-5."""
import numpy as np
import pytest

from oracle import oracle as O
from musichip import hip as H
from musichip import gcmc as G

import test_gcmc_comcut_oracle as C

pytestmark = pytest.mark.gpu


def _cell(ctx, L):
    L = np.asarray(L, dtype=np.float64)
    ctx.upload_framework(np.zeros((1, 3)), np.zeros(1, dtype=np.int32), None, L, L)


def _compare(b, ref, nsteps):
    s = b.state()
    for r, (to, No, eo) in enumerate(ref):
        tg = b.trace(r)
        assert np.array_equal(tg[:, :3].astype(int), to[:, :3].astype(int)), r
        scale = np.maximum(np.abs(to[:, 3]), 1.0)
        assert np.max(np.abs(tg[:, 3] - to[:, 3]) / scale) < 1e-10
        assert No == s["N"][r]
        assert abs(eo - s["energy"][r]) < 1e-9 * max(1.0, abs(eo))
    assert np.allclose(b.total_energy(), s["energy"], rtol=1e-10, atol=1e-9)


@pytest.mark.parametrize("threads,maxN", [(256, 512), (512, 512), (256, 20000)])
def test_comcut_trajectory_matches_oracle(hip, threads, maxN):
    print("case: bent sorbate, COMCUT lj 6 / coul 5, FP64 parity, threads %d maxN %d" % (threads, maxN),
          flush=True)
    _cell(hip, C.L)
    spec = C._bent(dict(lj=6.0, coul=5.0))
    fugs = [3.0e3, 2.0e4]
    seeds = [811, 812]
    nsteps = 3000
    ref = []
    for r in range(2):
        g = O.Gcmc(spec, fugs[r], seeds[r], maxN=512, trace=nsteps)
        g.run(nsteps)
        ref.append((g.trace_array(nsteps), g.st.N, g.st.energy))
    assert ref[1][1] > 10
    b = G.GcmcBatch(hip, spec, fugs, seeds, maxN=maxN, threads=threads)
    b.set_trace(nsteps)
    b.run(nsteps)
    _compare(b, ref, nsteps)
    b.close()


def test_comcut_from_placed_configuration(hip):
    """Centres computed on the host for placed molecules (mh_gcmc_set_config)
    equal the oracle's: the chain continued from a placed configuration."""
    print("case: bent sorbate, COMCUT from a placed configuration", flush=True)
    _cell(hip, C.L)
    spec = C._bent(dict(lj=5.5, coul=4.5))
    rng = np.random.default_rng(5)
    com, ang = C._config(rng, 16)
    nsteps = 2000
    g = O.Gcmc(spec, 1.0e4, 901, maxN=512, trace=nsteps)
    g.set_config(com, ang)
    e0 = g.total_energy()
    g.run(nsteps)
    b = G.GcmcBatch(hip, spec, [1.0e4], [901], maxN=512)
    b.set_config(0, com, ang)
    assert abs(b.total_energy()[0] - e0) < 1e-9 * max(1.0, abs(e0))
    b.set_trace(nsteps)
    b.run(nsteps)
    _compare(b, [(g.trace_array(nsteps), g.st.N, g.st.energy)], nsteps)
    # FP32 kernel: trial energies on the final configuration
    c, a, _ = b.config(0)
    b32 = G.GcmcBatch(hip, spec, [1.0e4], [3], maxN=512, precision=H.MH_PREC_FP32)
    b32.set_config(0, c, a)
    tc = rng.uniform(0, 1, (1500, 3)) * C.L
    ta = np.stack([rng.uniform(-np.pi, np.pi, 1500), np.arccos(rng.uniform(-1, 1, 1500)),
                   rng.uniform(-np.pi, np.pi, 1500)], 1)
    ok64, e64 = b.trial_energy(0, tc, ta)
    ok32, e32 = b32.trial_energy(0, tc, ta)
    both = ok64 & ok32
    assert both.sum() > 100 and np.count_nonzero(ok64 != ok32) <= 4
    rel = np.abs(e32[both] - e64[both]) / np.maximum(np.abs(e64[both]), 1.0)
    print("COMCUT FP32 trial energy max rel err %.3e" % rel.max(), flush=True)
    assert rel.max() <= 1e-5
    b32.close()
    b.close()


def test_comcut_mixture_per_species_pair(hip):
    print("case: bent + CH4 mixture, COMCUT 5 between the species only, FP64 parity", flush=True)
    _cell(hip, C.L)
    specs = [C._bent(dict(lj=[0.0, 5.0]), coul=False), C._ch4(dict(lj=[5.0, 0.0]))]
    fugs = [[3.0e3, 2.0e4], [1.0e4, 5.0e4]]
    seeds = [1201, 1202]
    nsteps = 3000
    b = G.GcmcMix(hip, specs, fugs, seeds, maxN=256)
    b.set_trace(nsteps)
    b.run(nsteps)
    for r in range(2):
        g = O.GcmcMix(specs, fugs[r], seeds[r], maxN=256, trace=nsteps)
        g.run(nsteps)
        to, tg = g.trace_array(nsteps), b.trace(r)
        assert np.array_equal(tg[:, :3].astype(int), to[:, :3].astype(int)), r
        scale = np.maximum(np.abs(to[:, 3]), 1.0)
        assert np.max(np.abs(tg[:, 3] - to[:, 3]) / scale) < 1e-10
        for s in range(2):
            b.select(s)
            assert b.state()["N"][r] == g.N()[s]
        assert min(g.N()) >= 2
    b.select(0)
    assert np.allclose(b.total_energy(), b.state()["energy"], rtol=1e-10, atol=1e-9)
    b.close()


def test_comcut_asymmetric_radii_are_refused(hip):
    _cell(hip, C.L)
    specs = [C._bent(dict(lj=[0.0, 5.0]), coul=False), C._ch4(dict(lj=[4.0, 0.0]))]
    with pytest.raises(H.MusicHipError, match="centre-of-mass cutoffs"):
        G.GcmcMix(hip, specs, [[1.0e3, 1.0e3]], [1], maxN=16)
