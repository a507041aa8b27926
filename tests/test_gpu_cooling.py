"""The glass-transition scan on the GPU through the C ABI: caps::run_cooling's stages (core/src/mechanics.cpp:1381-1487)
on one context — an unsampled hold at the first temperature, then one hold per temperature of the ladder, each point the
mean of its hold's thermo rows from average_from on — and many replicas scanned together through the replica batch."""
import numpy as np
import pytest

import caps_amd as ca
import helpers as H

pytestmark = pytest.mark.gpu

OPTS = ca.EnergyOptions(cutoff=8.0, coulomb=True, electrostatics="pme", pme_order=4, pme_spacing=1.0, tail=False, parts=5)


def test_run_cooling_stages_and_points():
    print("case run_cooling NVT 600 -> 300 K", flush=True)
    s = H.melt_case(chains=8, seed=23)
    x0 = s.x.copy()
    rows = {}

    def progress(row, index, total):
        assert total == 7
        rows.setdefault(index, []).append(row)
        return True

    o = ca.CoolingOptions(t_start=600, t_end=300, t_step=50, ps_per_step=0.5, equilibrate_ps=0.25, dt=0.5, tau_t=50, barostat="none",
                          seed=5, new_velocities=True, property="potential", energy=OPTS, progress=progress)
    r = ca.run_cooling(s, o)
    T = [p["temperature"] for p in r.points]
    assert T == [600.0, 550.0, 500.0, 450.0, 400.0, 350.0, 300.0]
    assert sorted(rows) == [-1, 0, 1, 2, 3, 4, 5, 6]
    # the equilibration hold: 500 steps, a row every 500 / 50 = 10 steps (and the first); the holds: 1000 steps, every 5
    assert len(rows[-1]) == 51 and all(len(rows[k]) == 201 for k in range(7))
    # steps run on from one hold to the next (step_offset)
    assert rows[0][0]["step"] == 500 and rows[6][-1]["step"] == 500 + 7 * 1000
    for k, p in enumerate(r.points):
        rk = rows[k]
        kept = rk[int(0.5 * len(rk)):]
        assert abs(p["potential"] - np.mean([q["potential"] for q in kept])) <= 1e-9 * abs(p["potential"])
        assert abs(np.mean([q["temperature"] for q in kept]) - T[k]) < 0.1 * T[k]
    pot = np.array([p["potential"] for p in r.points])
    print("potential per hold:", np.round(pot, 1), "fit:", r.fit, flush=True)
    assert pot[0] > pot[-1]
    assert not np.array_equal(s.x, x0)


def test_run_cooling_at_constant_pressure_reports_the_density():
    print("case run_cooling NPT", flush=True)
    s = H.melt_case(chains=8, seed=24)
    v0 = abs(np.linalg.det(s.cell[3:].reshape(3, 3)))
    o = ca.CoolingOptions(t_start=500, t_end=400, t_step=50, ps_per_step=0.5, equilibrate_ps=0.25, dt=0.5, tau_t=50, tau_p=200,
                          pressure=10000.0, barostat="c-rescale", seed=6, new_velocities=True, energy=OPTS)
    r = ca.run_cooling(s, o)
    d = np.array([p["density"] for p in r.points])
    assert len(d) == 3 and np.all(d > 0) and np.all(np.isfinite([p["density_err"] for p in r.points]))
    assert abs(np.linalg.det(s.cell[3:].reshape(3, 3))) != v0
    assert all(abs(p["specific_volume"] * p["density"] - 1) < 1e-12 for p in r.points)


def test_batch_cooling_of_several_replicas():
    print("case batch cooling, 3 replicas", flush=True)
    systems = [H.melt_case(chains=c, seed=40 + c) for c in (6, 8, 10)]
    evs = [ca.Evaluator(s.ff, OPTS) for s in systems]
    b = ca.Batch(evs)
    o = ca.CoolingOptions(t_start=600, t_end=300, t_step=50, ps_per_step=0.5, equilibrate_ps=0.25, dt=0.5, tau_t=50, barostat="none",
                          seed=7, new_velocities=True, property="potential")
    res = ca.batch_cooling(b, systems, o, samples=4)
    assert len(res) == 3
    for r, s in zip(res, systems):
        T = [p["temperature"] for p in r.points]
        pot = np.array([p["potential"] for p in r.points])
        print("replica of %d atoms: potential per hold %s" % (s.ff.n, np.round(pot, 1)), flush=True)
        assert T == [600.0, 550.0, 500.0, 450.0, 400.0, 350.0, 300.0]
        assert np.all(np.isfinite(pot)) and pot[0] > pot[-1]
        assert np.isfinite(s.x).all()
    b.close()
    with pytest.raises(ValueError):
        ca.batch_cooling(ca.Batch(evs), systems, ca.CoolingOptions(barostat="c-rescale"))
