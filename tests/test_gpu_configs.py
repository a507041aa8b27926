"""GPU checks at the sizes of the replica-ladder and learned r-RESPA workloads, through the C ABI:

* 64 independent melts of 3,000-40,000 atoms (seeded), each its own cubic box and PME grid, Langevin on a 500 -> 200 K
  ladder, stepped together in grouped launches: every replica ends bit-identical to run_dynamics of that replica alone,
  and the smallest and largest replicas' forces (PME reciprocal part + bonded terms) match the FP64 oracle;
* learned r-RESPA on the 262,144-atom melt with a short training budget: the judging schedule, the fallback count and
  the force errors it reports are consistent.
"""
import numpy as np
import pytest

import caps_amd as ca
import helpers as H
import orcpy as O

pytestmark = pytest.mark.gpu

LADDER = ca.EnergyOptions(cutoff=12.0, coulomb=True, electrostatics="pme", ewald_rtol=1e-5, pme_order=4, pme_spacing=1.0, tail=False,
                          parts=5)
STEPS, SEED = 9, 5


def _ladder():
    rng = np.random.default_rng(1)
    sizes = rng.integers(3000, 40001, 64)
    chains = [max(1, int(round(int(m) / 128))) for m in sizes]
    return [ca.build_melt(c, seed=200 + k) for k, c in enumerate(chains)], np.linspace(500.0, 200.0, 64)


@pytest.fixture(scope="module")
def ladder():
    print("case replica ladder: building 64 melts", flush=True)
    return _ladder()


def test_replica_ladder_matches_each_replica_alone(ladder):
    systems, temps = ladder
    start = [(s.x.copy(), s.cell.copy()) for s in systems]
    atoms = sum(s.ff.n for s in systems)
    assert 64 * 3000 <= atoms <= 64 * 40100 and min(s.ff.n for s in systems) >= 2944 and max(s.ff.n for s in systems) <= 40064
    print("case replica ladder: %d atoms, grouped launches" % atoms, flush=True)
    b = ca.Batch([ca.Evaluator(s.ff, LADDER) for s in systems])
    rep = b.run(systems, temps, ca.BatchOptions(dt=1.0, steps=STEPS, tau_t=100, seed=SEED, new_velocities=True, mode="grouped",
                                                graph_steps=4))
    b.close()
    assert rep.mode == "grouped" and np.all(np.isfinite(rep.temperatures))
    print("case replica ladder: each replica alone", flush=True)
    for r, s in enumerate(systems):
        x0, cell = start[r]
        alone = ca.System(x=x0.copy(), cell=cell, ff=s.ff)
        ev = ca.Evaluator(s.ff, LADDER)
        ca.run_dynamics(alone, ca.DynamicsOptions(dt=1.0, steps=STEPS, temperature=float(temps[r]), thermostat="langevin", tau_t=100,
                                                  seed=SEED + r, thermo_every=STEPS, new_velocities=True, energy=LADDER), evaluator=ev)
        ev.close()
        assert np.array_equal(alone.x, s.x), (r, np.abs(alone.x - s.x).max())
        assert np.array_equal(alone.v, s.v), (r, np.abs(alone.v - s.v).max())


@pytest.mark.parametrize("which", ["smallest", "largest"])
def test_replica_ladder_forces_match_the_oracle(ladder, which):
    systems, _ = ladder
    sizes = [s.ff.n for s in systems]
    s = systems[int(np.argmin(sizes) if which == "smallest" else np.argmax(sizes))]
    s = ca.System(x=s.x.copy(), cell=s.cell.copy(), ff=s.ff)
    print("case replica ladder forces: %s replica, %d atoms" % (which, s.ff.n), flush=True)
    ev = ca.Evaluator(s.ff, LADDER)
    t, f = ev.compute(s.x, s.cell)
    ev.close()
    beta = O.orc().orc_ewald_beta(12.0, 1e-5)
    grid = np.zeros(3, np.int32)
    O.orc().orc_pme_grid(np.ascontiguousarray(s.cell, np.float64), 1.0, 4, grid)
    er, fr, _ = H.oracle_pme(s, beta, grid, 4)
    eb, fb, _ = H.oracle_bonded(s)
    fo = fr + fb
    q = s.ff.charge
    ecoul = er - 332.06371 * beta / np.sqrt(np.pi) * float((q * q).sum())
    rel = np.sqrt(((f - fo) ** 2).sum() / (fo ** 2).sum())
    print("grid %s: coulomb %.6f (oracle %.6f), force rel RMS %.2e" % (list(grid), t.coulomb, ecoul, rel), flush=True)
    assert abs(t.coulomb - ecoul) <= 1e-5 * abs(ecoul)
    assert abs(t.bond - eb[0]) <= 1e-10 * abs(eb[0]) and abs(t.angle - eb[1]) <= 1e-10 * abs(eb[1])
    assert rel < 1e-5


def test_learned_respa_on_the_262k_melt():
    print("case learned r-RESPA 262,144 atoms", flush=True)
    s = ca.build_melt(2048, seed=1)
    op = ca.EnergyOptions(cutoff=12.0, coulomb=True, electrostatics="pme", ewald_rtol=1e-5, pme_order=5, pme_spacing=1.0, tail=False,
                          parts=5)
    ev = ca.Evaluator(s.ff, op)
    train, steps = 40, 80
    r = ca.run_dynamics(s, ca.DynamicsOptions(dt=1.0, steps=train + steps, temperature=300, thermostat="langevin", tau_t=100, seed=3,
                                              thermo_every=4, new_velocities=True, energy=op, learned_k=4, learned_tol=1e-3,
                                              learned_train=train), evaluator=ev)
    judged = [row for row in r.thermo if row["learned_rmse"] > 0]
    print("learned %d, full %d, judgments %d, fallbacks %d" % (r.learned_steps, r.full_steps, r.judgments, r.fallbacks), flush=True)
    assert r.learned_steps + r.full_steps == train + steps
    assert r.judgments >= 1 and 0 <= r.fallbacks <= r.judgments and judged
    lr = np.array([row["learned_rmse"] for row in judged])
    st = np.array([row["learned_rmse_stale"] for row in judged])
    print("judged relative force RMSE: learned %.3e, stale %.3e" % (lr.mean(), st.mean()), flush=True)
    assert np.all(np.isfinite(lr)) and np.all((lr > 0) & (lr < 1)) and np.all((st > 0) & (st < 1))
    # a learned step is taken only after a judgment that met the tolerance
    if r.fallbacks == r.judgments:
        assert r.learned_steps == 0
    assert np.isfinite(s.x).all()
