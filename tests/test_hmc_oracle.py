"""The CPU oracle of the HMC burst pinned by closed forms: forces are the
negative gradient of the energy (central differences), the NVE integrator's
energy error shrinks as dt^2, the Gaussian velocities are the Box-Muller
values of the rranf stream, and the torsion is the reference's polymer
convention of the textbook dihedral."""
import numpy as np
import pytest

from oracle import oracle as O
from musichip import synth
import hmc_systems as HS


@pytest.fixture(scope="module")
def butane_sys():
    fw = HS.framework()
    model = synth.butane_ua()
    maps = HS.oracle_maps(fw, model)
    return model, maps


def test_forces_are_energy_gradient(butane_sys):
    model, maps = butane_sys
    sp = HS.spec_with_maps(model, maps)
    h = O.Hmc(sp, 5)
    rp = HS.place(model, maps, 3, seed=1)
    # strain the chains so every bonded term has a force
    rp = rp + np.random.default_rng(2).normal(0.0, 0.05, rp.shape)
    e0, ok = h.set_config(rp)
    assert ok
    _, F, ok = h.energy_forces()
    assert ok
    r = rp - np.floor(rp / HS.L16) * HS.L16
    d = 1e-6
    num = np.zeros_like(F)
    for m in range(rp.shape[0]):
        for a in range(rp.shape[1]):
            for i in range(3):
                ep, em = [], []
                for sgn, out in ((1, ep), (-1, em)):
                    rp2, r2 = rp.copy(), r.copy()
                    rp2[m, a, i] += sgn * d
                    r2[m, a, i] += sgn * d
                    out.append(h.energy_forces(rp2, r2)[0])
                num[m, a, i] = -(ep[0] - em[0]) / (2 * d)
    scale = np.abs(F).max()
    err = np.abs(num - F).max() / scale
    print("max |F - (-dE/dx)| / max|F| = %.3e (max |F| %.3e kcal/mol/A)" % (err, scale))
    assert err < 1e-6


def test_nve_energy_error_is_second_order(butane_sys):
    model, maps = butane_sys
    rp = HS.place(model, maps, 3, seed=4)
    rng = np.random.default_rng(5)
    mass = np.array(model["mass"])
    sig = np.sqrt(8.314e-3 * 300.0 / (0.01 * mass))
    v = rng.normal(size=(3, 4, 3)) * sig[None, :, None]
    errs = []
    for dt in (0.0005, 0.00025):
        sp = HS.spec_with_maps(model, maps, dt=dt)
        # the sorbate-sorbate LJ is truncated without a shift: a pair crossing
        # the cutoff changes the energy by a step no integrator removes, so the
        # order test uses the smooth part (maps + bonded terms)
        sp["lj"] = sp["lj"][:4] + ([0] * len(sp["lj"][4]),)
        h = O.Hmc(sp, 5)
        h.set_config(rp)
        ok, et = h.nve(v, int(round(0.01 / dt)))
        assert ok
        errs.append(np.abs(et - et[0]).max())
    print("max |dE| at dt 0.5 / 0.25 fs: %.3e / %.3e kcal/mol" % tuple(errs))
    assert 2.5 < errs[0] / errs[1] < 6.0


def test_gaussians_are_box_muller_of_rranf():
    model = HS.propane_ua()
    h = O.Hmc(HS.spec_with_maps(model, []), 22111971)
    g = [h.gaussian() for _ in range(4)]
    u, _, _ = O.rranf(22111971, 4)
    twopi = 2.0 * 3.1415927
    exp = [np.sqrt(-2 * np.log(u[0])) * np.cos(twopi * u[1]), np.sqrt(-2 * np.log(u[0])) * np.sin(twopi * u[1]),
           np.sqrt(-2 * np.log(u[2])) * np.cos(twopi * u[3]), np.sqrt(-2 * np.log(u[2])) * np.sin(twopi * u[3])]
    np.testing.assert_allclose(g, exp, rtol=1e-15, atol=0)


def test_torsion_is_polymer_convention_of_iupac_dihedral():
    model = synth.butane_ua()
    model = dict(model, bonds=[], angles=[])
    sp = HS.spec_with_maps(model, [])
    sp["site_map"] = [-1] * 4
    h = O.Hmc(sp, 1)
    rng = np.random.default_rng(8)
    c = np.array(model["torsions"][0][4])
    for _ in range(20):
        x = rng.normal(size=(1, 4, 3)) * 1.5 + 8.0
        e, _, ok = h.energy_forces(x, x)
        b1, b2, b3 = x[0, 1] - x[0, 0], x[0, 2] - x[0, 1], x[0, 3] - x[0, 2]
        n1, n2 = np.cross(b1, b2), np.cross(b2, b3)
        cphi = n1 @ n2 / np.linalg.norm(n1) / np.linalg.norm(n2)
        exp = sum(c[k] * (-cphi) ** k for k in range(6))
        assert ok and abs(e - exp) < 1e-10 * max(1.0, abs(exp))
    # the trans chain is the minimum (U = 0)
    e, _, _ = h.energy_forces(model["chain"][None] + 8.0, model["chain"][None] + 8.0)
    assert abs(e) < 1e-9


def test_hmc_acceptance_tends_to_one_as_dt_vanishes(butane_sys):
    model, maps = butane_sys
    rp = HS.place(model, maps, 3, seed=9)
    rates = []
    for dt in (0.002, 0.0002):
        h = O.Hmc(HS.spec_with_maps(model, maps, dt=dt, nmd=10), 77)
        h.set_config(rp)
        dHs = []
        for _ in range(30):
            h.move()
            dHs.append(abs(h.st.last_dH))
        rates.append(np.mean(dHs))
    print("mean |dH| at dt 2 / 0.2 fs: %.3e / %.3e kcal/mol" % tuple(rates))
    assert rates[1] < rates[0] / 20
