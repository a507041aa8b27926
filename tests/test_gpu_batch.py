"""GPU checks of replicas stepped together (no reference counterpart: the reference runs a cooling scan's replicas one
after another through run_dynamics, core/src/mechanics.cpp:1381-1470): every launch shape — graphs replayed on several
streams, direct launches, grouped launches (each kernel once for every replica), the tuned choice — leaves each replica
bit-identical to run_dynamics of that replica alone."""
import copy

import numpy as np
import pytest

import caps_amd as ca
import helpers as H

pytestmark = pytest.mark.gpu

OPTS = ca.EnergyOptions(cutoff=9.0, coulomb=True, electrostatics="pme", pme_order=4, pme_spacing=1.0, tail=False, parts=5)
CHAINS = (8, 13, 6)            # 1024, 1664 and 768 atoms: three boxes, three PME grids
TEMPS = (450.0, 350.0, 250.0)
STEPS, SEED = 21, 9            # 21 = 2 graphs of 8 steps + 5 single steps


def _systems():
    return [H.melt_case(chains=c, seed=30 + k) for k, c in enumerate(CHAINS)]


def _alone():
    out = []
    for r, s in enumerate(_systems()):
        ev = ca.Evaluator(s.ff, OPTS)
        ca.run_dynamics(s, ca.DynamicsOptions(dt=1.0, steps=STEPS, temperature=TEMPS[r], thermostat="langevin", tau_t=100, seed=SEED + r,
                                              thermo_every=STEPS, new_velocities=True, energy=OPTS), evaluator=ev)
        out.append((s.x.copy(), s.v.copy()))
    return out


@pytest.fixture(scope="module")
def alone():
    print("case batch: each replica alone through run_dynamics", flush=True)
    return _alone()


@pytest.mark.parametrize("mode,branches,graph_steps", [("graph", 2, 8), ("graph", 1, 4), ("direct", 1, 1), ("grouped", 1, 8), ("grouped", 1, 5)])
def test_batch_matches_each_replica_alone(alone, mode, branches, graph_steps):
    print("case batch %s branches %d graph_steps %d" % (mode, branches, graph_steps), flush=True)
    sys_ = _systems()
    b = ca.Batch([ca.Evaluator(s.ff, OPTS) for s in sys_])
    rep = b.run(sys_, TEMPS, ca.BatchOptions(dt=1.0, steps=STEPS, tau_t=100, seed=SEED, new_velocities=True, mode=mode, branches=branches,
                                             graph_steps=graph_steps))
    assert rep.mode == mode and rep.steps == STEPS and rep.atom_steps_per_s > 0
    for r, s in enumerate(sys_):
        x, v = alone[r]
        assert np.array_equal(s.x, x), (r, np.abs(s.x - x).max())
        assert np.array_equal(s.v, v), (r, np.abs(s.v - v).max())
    assert np.all(np.isfinite(rep.temperatures)) and np.all(rep.temperatures > 0)


def test_tuned_batch_matches_and_keeps_its_choice(alone):
    print("case batch tuned", flush=True)
    sys_ = _systems()
    evs = [ca.Evaluator(s.ff, OPTS) for s in sys_]
    b = ca.Batch(evs)
    o = ca.BatchOptions(dt=1.0, steps=STEPS, tau_t=100, seed=SEED, new_velocities=True)
    rep = b.run(sys_, TEMPS, o)
    print("tuned shape: %s, %d branches, %d steps per graph (tuning %.3f s)" % (rep.mode, rep.branches, rep.graph_steps, rep.tune_seconds),
          flush=True)
    assert rep.mode in ("graph", "direct", "grouped") and rep.tune_seconds > 0
    for r, s in enumerate(sys_):
        assert np.array_equal(s.x, alone[r][0]) and np.array_equal(s.v, alone[r][1]), r
    # the same run again: the tuned shape is reused (no timing) and, from the same start, the same end
    sys2 = _systems()
    rep2 = b.run(sys2, TEMPS, o)
    assert rep2.tune_seconds == 0 and (rep2.mode, rep2.branches, rep2.graph_steps) == (rep.mode, rep.branches, rep.graph_steps)
    for r, s in enumerate(sys2):
        assert np.array_equal(s.x, alone[r][0]) and np.array_equal(s.v, alone[r][1]), r


def test_batch_continues_from_its_state():
    # two runs of 10 steps each continue the trajectory: velocities carried over, the noise counter restarts at each run
    # as run_dynamics's step numbers do
    print("case batch continuation", flush=True)
    s0 = H.melt_case(chains=6, seed=41)
    s_alone = copy.deepcopy(s0)
    ev = ca.Evaluator(s_alone.ff, OPTS)
    d = ca.DynamicsOptions(dt=1.0, steps=10, temperature=300, thermostat="langevin", tau_t=100, seed=5, thermo_every=10, new_velocities=True,
                           energy=OPTS)
    ca.run_dynamics(s_alone, d, evaluator=ev)
    d.new_velocities = False
    ca.run_dynamics(s_alone, d, evaluator=ev)
    s_b = copy.deepcopy(s0)
    b = ca.Batch([ca.Evaluator(s_b.ff, OPTS)])
    b.run([s_b], [300.0], ca.BatchOptions(dt=1.0, steps=10, tau_t=100, seed=5, new_velocities=True, mode="graph", graph_steps=3))
    b.run([s_b], [300.0], ca.BatchOptions(dt=1.0, steps=10, tau_t=100, seed=5, new_velocities=False, mode="graph", graph_steps=3))
    assert np.array_equal(s_b.x, s_alone.x) and np.array_equal(s_b.v, s_alone.v)


def test_batch_refuses_the_pair_path_and_bad_shapes():
    print("case batch errors", flush=True)
    s = H.melt_case(chains=4, seed=2)
    with pytest.raises(ValueError, match="pair path"):
        b = ca.Batch([ca.Evaluator(s.ff, ca.EnergyOptions(cutoff=9.0, coulomb=False, parts=3, tail=False))])
        b.run([s], [300.0], ca.BatchOptions(steps=2))
    b = ca.Batch([ca.Evaluator(s.ff, OPTS)])
    with pytest.raises(ValueError, match="branches"):
        b.run([s], [300.0], ca.BatchOptions(steps=2, mode="graph", branches=2, graph_steps=4))
    with pytest.raises(ValueError, match="temperature"):
        b.run([s], [-1.0], ca.BatchOptions(steps=2))


def test_grouped_launches_need_one_spline_order_and_keep_parts_per_replica():
    # a replica with spline order 5 among order-4 ones: grouped launches are refused, the tuned shape falls back to
    # per-replica launches and stays exact; a replica without bonded terms (parts = 4) takes part in grouped launches
    print("case batch mixed orders and parts", flush=True)
    o5 = ca.EnergyOptions(cutoff=9.0, coulomb=True, electrostatics="pme", pme_order=5, pme_spacing=1.0, tail=False, parts=5)
    o4r = ca.EnergyOptions(cutoff=9.0, coulomb=True, electrostatics="pme", pme_order=4, pme_spacing=1.0, tail=False, parts=4)

    def alone(systems, opts):
        out = []
        for r, (s, op) in enumerate(zip(systems, opts)):
            ca.run_dynamics(s, ca.DynamicsOptions(dt=1.0, steps=9, temperature=TEMPS[r], thermostat="langevin", tau_t=100, seed=SEED + r,
                                                  thermo_every=9, new_velocities=True, energy=op), evaluator=ca.Evaluator(s.ff, op))
            out.append((s.x.copy(), s.v.copy()))
        return out

    for opts, mode in (((OPTS, o5, OPTS), "tuned"), ((OPTS, o4r, OPTS), "grouped")):
        ref = alone(_systems(), opts)
        sys_ = _systems()
        b = ca.Batch([ca.Evaluator(s.ff, op) for s, op in zip(sys_, opts)])
        if opts[1] is o5:
            with pytest.raises(ValueError, match="one spline order"):
                b.run(sys_, TEMPS, ca.BatchOptions(dt=1.0, steps=9, tau_t=100, seed=SEED, mode="grouped", graph_steps=4))
            sys_ = _systems()
        rep = b.run(sys_, TEMPS, ca.BatchOptions(dt=1.0, steps=9, tau_t=100, seed=SEED, mode=mode, graph_steps=4))
        if opts[1] is o5:
            assert rep.mode in ("graph", "direct")
        for r, s in enumerate(sys_):
            assert np.array_equal(s.x, ref[r][0]) and np.array_equal(s.v, ref[r][1]), (mode, r)
