"""GPU dynamics (velocity Verlet + Bussi, Langevin BAOAB, NVE; dynamics.cpp:606-774) on the bonded forces alone, through
the C ABI: the reference's invariants (tests/test_dynamics.cpp:97-151) and the exact initial velocities of the run's own
mt19937_64 stream."""
import numpy as np
import pytest

import caps_amd as ca
import helpers as H
import orcpy as O

pytestmark = pytest.mark.gpu
BONDED = ca.EnergyOptions(cutoff=10.0, coulomb=False, parts=1, tail=False)


def test_initial_velocities_are_the_runs_own_stream():
    s = H.melt_case(chains=4)
    print("case initial velocities", flush=True)
    ca.run_dynamics(s, ca.DynamicsOptions(dt=0.5, steps=0, temperature=300, thermostat="none", seed=17, new_velocities=True,
                                          energy=BONDED))
    v = np.zeros(3 * s.ff.n)
    O.orc().orc_initial_velocities(O.Rng(17).ptr(), np.ascontiguousarray(s.ff.mass), s.ff.n, 300.0, v)
    assert np.abs(s.v.reshape(-1) - v).max() <= 1e-12 * np.abs(v).max()


def test_nve_conserves_the_energy():
    s = ca.kremer_grest(chains=20, beads=30, density=0.85, seed=3)
    print("case nve kg bonded", flush=True)
    r = ca.run_dynamics(s, ca.DynamicsOptions(dt=0.25, steps=400, temperature=200, thermostat="none", seed=5, thermo_every=20,
                                              new_velocities=True, energy=BONDED))
    # test_dynamics.cpp:97-118: after the first rows the total stays in a band of 1 % of ⟨K⟩ and does not drift
    E = np.array([row["total"] for row in r.thermo])[5:]
    K = np.array([row["kinetic"] for row in r.thermo])[5:].mean()
    assert E.max() - E.min() < 0.01 * K
    q = len(E) // 5
    assert abs(E[-q:].mean() - E[:q].mean()) < 0.005 * K


def test_bussi_conserved_quantity():
    s = H.melt_case(chains=4)
    print("case bussi melt bonded", flush=True)
    r = ca.run_dynamics(s, ca.DynamicsOptions(dt=0.25, steps=400, temperature=300, thermostat="bussi", tau_t=50, seed=9,
                                              thermo_every=20, new_velocities=True, energy=BONDED))
    # test_dynamics.cpp:138-151
    C = np.array([row["conserved"] for row in r.thermo])[5:]
    K = np.array([row["kinetic"] for row in r.thermo])[5:].mean()
    assert C.max() - C.min() < 0.01 * K


def test_langevin_is_reproducible_and_holds_the_temperature():
    print("case langevin melt bonded", flush=True)
    out = []
    for _ in range(2):
        s = H.melt_case(chains=4)
        r = ca.run_dynamics(s, ca.DynamicsOptions(dt=0.5, steps=2000, temperature=300, thermostat="langevin", tau_t=100, seed=4,
                                                  thermo_every=20, new_velocities=True, energy=BONDED))
        out.append((s.x.copy(), s.v.copy(), r))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    T = np.array([row["temperature"] for row in out[0][2].thermo])
    assert abs(T[len(T) // 2:].mean() - 300) < 30


@pytest.mark.parametrize("thermostat", ["bussi", "langevin"])
def test_thermostats_hold_the_temperature(thermostat):
    # tests/test_dynamics.cpp:120-135: 400 K, tau_t 50 fs, 4000 steps, thermo every 10; the mean temperature past row 50
    # within 16 K of the target. Bonded terms + PME's reciprocal part on a 1,024-atom melt.
    print("case thermostat %s 400 K" % thermostat, flush=True)
    s = H.melt_case(chains=8, seed=21)
    op = ca.EnergyOptions(cutoff=8.0, coulomb=True, electrostatics="pme", pme_order=4, pme_spacing=1.0, tail=False, parts=5)
    r = ca.run_dynamics(s, ca.DynamicsOptions(dt=0.5, steps=4000, temperature=400, thermostat=thermostat, tau_t=50, seed=11,
                                              thermo_every=10, new_velocities=True, energy=op))
    T = np.array([row["temperature"] for row in r.thermo])
    print("%s: mean T past row 50 = %.2f K over %d rows" % (thermostat, T[50:].mean(), len(T) - 50), flush=True)
    assert abs(T[50:].mean() - 400.0) < 16.0


def test_state_round_trip():
    s = H.melt_case(chains=2)
    print("case state round trip", flush=True)
    ev = ca.Evaluator(s.ff, BONDED)
    v = np.random.default_rng(1).normal(0, 0.01, s.x.shape)
    ev.set_state(s.x, v, s.cell)
    x2, v2, c2 = ev.get_state()
    assert np.array_equal(x2, s.x) and np.array_equal(v2, v) and np.array_equal(c2, s.cell)


def _band(respa, dt, steps):
    s = H.melt_case(chains=4)
    op = ca.EnergyOptions(cutoff=8.0, coulomb=True, electrostatics="pme", pme_order=4, pme_spacing=1.0, tail=False, parts=5)
    r = ca.run_dynamics(s, ca.DynamicsOptions(dt=dt, steps=steps, respa=respa, temperature=300, thermostat="none", seed=7,
                                              thermo_every=5, new_velocities=True, energy=op))
    E = np.array([row["total"] for row in r.thermo])[4:]
    K = np.array([row["kinetic"] for row in r.thermo])[4:].mean()
    return E.max() - E.min(), K


def test_respa_conserves_energy_with_long_outer_steps():
    # test_dynamics.cpp:248-273, with PME's reciprocal part as the slow force and the bonded terms as the fast one: bonded
    # forces at 0.5 fs inside 2 fs steps conserve energy as plain 0.5 fs steps do, far better than plain 2 fs steps
    print("case r-RESPA bands", flush=True)
    plain_fine, kin = _band(1, 0.5, 1200)
    respa, kin2 = _band(4, 2.0, 300)
    plain_coarse, _ = _band(1, 2.0, 300)
    print("bands: plain 0.5 fs %.4g, r-RESPA 4 x 0.5 fs %.4g, plain 2 fs %.4g (kinetic %.4g)" % (plain_fine, respa, plain_coarse, kin2), flush=True)
    assert respa < 0.02 * kin2
    assert respa < 0.5 * plain_coarse
    assert plain_fine < 0.01 * kin


def test_respa_needs_bussi_or_nve_and_both_parts():
    s = H.melt_case(chains=2)
    op = ca.EnergyOptions(cutoff=8.0, coulomb=True, electrostatics="pme", pme_order=4, tail=False, parts=5)
    with pytest.raises(ValueError):
        ca.run_dynamics(s, ca.DynamicsOptions(dt=1.0, steps=2, respa=2, thermostat="langevin", energy=op))
    with pytest.raises(ValueError):
        ca.run_dynamics(s, ca.DynamicsOptions(dt=1.0, steps=2, respa=2, thermostat="none", energy=BONDED))
