"""GPU pressure coupling (Berendsen and stochastic cell rescaling, dynamics.cpp:701-723) through the C ABI: the cell
responds to the target pressure as the reference's test asks (tests/test_dynamics.cpp:152-168), one update is the
reference's formula applied to the step's own kinetic energy and virial, the same seed gives the same trajectory
(test_dynamics.cpp:170-181), and the reported density follows the cell."""
import math

import numpy as np
import pytest

import caps_amd as ca
import helpers as H

pytestmark = pytest.mark.gpu

OPTS = ca.EnergyOptions(cutoff=8.0, coulomb=True, electrostatics="pme", pme_order=4, pme_spacing=1.0, tail=False, parts=5)
KB, KATM = 0.0019872041, 68568.415


def _melt():
    return H.melt_case(chains=8, seed=21)


def test_barostat_responds_to_target_pressure():
    rho, vol0 = [], None
    for p0 in (1.0, 20000.0):
        print("case Berendsen at %g atm" % p0, flush=True)
        s = _melt()
        vol0 = abs(np.linalg.det(s.cell[3:].reshape(3, 3)))
        r = ca.run_dynamics(s, ca.DynamicsOptions(dt=0.5, steps=2000, temperature=300, thermostat="bussi", tau_t=100, seed=4,
                                                  thermo_every=100, new_velocities=True, energy=OPTS, barostat="berendsen", pressure=p0,
                                                  tau_p=200))
        vol = abs(np.linalg.det(s.cell[3:].reshape(3, 3)))
        assert vol != vol0
        assert abs(r.thermo[-1]["volume"] - vol) <= 1e-9 * vol
        mass = s.ff.mass.sum()
        assert abs(r.thermo[-1]["density"] - mass / 6.02214076e23 / (vol * 1e-24)) <= 1e-12 * r.thermo[-1]["density"]
        rho.append(r.thermo[-1]["density"])
    print("densities at 1 and 20000 atm: %.4f, %.4f g/cm3" % tuple(rho), flush=True)
    assert rho[1] > rho[0] + 0.02


def test_one_update_is_the_reference_formula():
    # NVE, 10 steps, one Berendsen update after step 10: the same 10 steps without the barostat give the kinetic energy and
    # the virial the update reads; mu = exp(deps / 3), deps = -beta / tau_p (P0 - P) dt * every, clamped to +-0.01
    print("case one Berendsen update", flush=True)
    beta, tau_p, p0, every, dt = 4.5e-5, 200.0, 500.0, 10, 0.5
    a, b = _melt(), _melt()
    d = dict(dt=dt, steps=every, temperature=300, thermostat="none", seed=6, thermo_every=every, new_velocities=True, energy=OPTS)
    ra = ca.run_dynamics(a, ca.DynamicsOptions(**d))
    rb = ca.run_dynamics(b, ca.DynamicsOptions(barostat="berendsen", pressure=p0, tau_p=tau_p, compressibility=beta, barostat_every=every,
                                               **d))
    row = ra.thermo[-1]
    V = abs(np.linalg.det(a.cell[3:].reshape(3, 3)))
    p = (2 * row["kinetic"] + row["terms"].virial) / (3 * V) * KATM
    assert abs(p - row["pressure"]) <= 1e-9 * abs(p)
    deps = min(max(-beta / tau_p * (p0 - p) * dt * every, -0.01), 0.01)
    mu = math.exp(deps / 3)
    print("pressure %.2f atm, mu %.8f" % (p, mu), flush=True)
    assert mu != 1.0
    assert np.allclose(b.cell[3:], mu * a.cell[3:], rtol=1e-14, atol=0)
    o = a.cell[:3]
    assert np.allclose(b.x, o + mu * (a.x - o), rtol=1e-12, atol=1e-12)
    assert np.array_equal(b.v, a.v)   # Berendsen leaves the velocities alone
    assert abs(rb.thermo[-1]["volume"] - mu ** 3 * V) <= 1e-9 * V


def test_same_seed_same_trajectory():
    # test_dynamics.cpp:170-181: Langevin + stochastic cell rescaling, 300 steps, twice
    print("case c-rescale reproducible", flush=True)
    out = []
    for _ in range(2):
        s = _melt()
        ca.run_dynamics(s, ca.DynamicsOptions(dt=0.5, steps=300, temperature=300, thermostat="langevin", seed=9, thermo_every=100,
                                              new_velocities=True, energy=OPTS, barostat="c-rescale", pressure=1.0, tau_p=200))
        out.append((s.x.copy(), s.cell.copy()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert not np.array_equal(out[0][1], _melt().cell)


def test_barostat_options_are_checked():
    s = _melt()
    with pytest.raises(ValueError):
        ca.run_dynamics(s, ca.DynamicsOptions(dt=0.5, steps=2, energy=OPTS, barostat="mtk"))
    with pytest.raises(ValueError):
        ca.run_dynamics(s, ca.DynamicsOptions(dt=0.5, steps=2, energy=OPTS, barostat="berendsen", tau_p=0.0))
