"""Checkpoints of the GPU dynamics through the C ABI (DynamicsOptions::checkpoint_every / checkpoint,
core/include/caps/dynamics.hpp:143-146, called as core/src/dynamics.cpp:758-763 calls it) written as caps md writes them
(a LAMMPS data file with velocities, cli/main.cpp:7820-7846), and a run resumed from one (cli/main.cpp:7809-7820). With
the counter-based Langevin noise drawn for the absolute step, and fixed-order sums, a resumed run is the uninterrupted
run to the last bit; NVE too."""
import numpy as np
import pytest

import caps_amd as ca
import helpers as H

pytestmark = pytest.mark.gpu

OPTS = ca.EnergyOptions(cutoff=8.0, coulomb=True, electrostatics="pme", pme_order=4, pme_spacing=1.0, tail=False, parts=5)


def test_energies_survive_the_round_trip(tmp_path):
    print("case LAMMPS round trip, melt 6 chains: same energies and forces", flush=True)
    s = H.melt_case(chains=6, seed=31)
    p = str(tmp_path / "m.data")
    ca.write_lammps(p, s)
    r = ca.read_lammps(p)
    t0, f0 = ca.Evaluator(s.ff, OPTS).compute(s.x, s.cell)
    t1, f1 = ca.Evaluator(r.ff, OPTS).compute(r.x, r.cell)
    assert t0.total() == t1.total() and np.array_equal(f0, f1)


@pytest.mark.parametrize("thermostat", ["langevin", "none"])
def test_a_resumed_run_continues_the_trajectory(tmp_path, thermostat):
    print("case checkpoint and resume, %s, 300 steps, checkpoints every 100" % thermostat, flush=True)
    s = H.melt_case(chains=6, seed=32)
    path = str(tmp_path / "run.restart.data")
    seen = []

    def keep(x, v, cell, step):
        seen.append((step, x.copy(), v.copy()))
        ca.write_checkpoint(path if step == 200 else path + ".%d" % step, ca.System(x=x, cell=cell, ff=s.ff, v=v), step, 300)

    o = ca.DynamicsOptions(dt=1.0, steps=300, temperature=300.0, thermostat=thermostat, tau_t=100.0, seed=9, thermo_every=50,
                           new_velocities=True, energy=OPTS, checkpoint_every=100, checkpoint=keep)
    full = ca.run_dynamics(s, o)
    # before the last step only, every 100 steps
    assert [q[0] for q in seen] == [100, 200]
    r = ca.read_lammps(path)
    assert ca.checkpoint_step(r) == 200 and np.array_equal(r.x, seen[1][1]) and np.array_equal(r.v, seen[1][2])
    o2 = ca.resume_options(r, o, 300)
    o2.checkpoint_every, o2.checkpoint = 0, None
    rest = ca.run_dynamics(r, o2)
    assert np.array_equal(r.x, s.x) and np.array_equal(r.v, s.v)
    a, b = full.thermo[-1], rest.thermo[-1]
    assert a["step"] == b["step"] == 300 and a["potential"] == b["potential"] and a["temperature"] == b["temperature"]
    assert [q["step"] for q in rest.thermo] == [200, 250, 300]


def test_the_checkpoint_path_form_and_a_failing_callback(tmp_path):
    print("case checkpoint to a path, and a callback that raises", flush=True)
    s = H.melt_case(chains=4, seed=33)
    path = str(tmp_path / "p.restart.data")
    o = ca.DynamicsOptions(dt=1.0, steps=60, temperature=300.0, thermostat="bussi", tau_t=100.0, seed=3, thermo_every=20,
                           new_velocities=True, energy=OPTS, checkpoint_every=25, checkpoint=path)
    ca.run_dynamics(s, o)
    r = ca.read_lammps(path)
    assert ca.checkpoint_step(r) == 50 and r.title.endswith("checkpoint step 50 of 60 · atom_style full · units real")

    def bad(x, v, cell, step):
        raise OSError("disk full")

    o.checkpoint = bad
    with pytest.raises(OSError, match="disk full"):
        ca.run_dynamics(s, o)
