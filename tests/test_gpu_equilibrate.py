"""Chained equilibration stages on the GPU (caps::equilibrate, core/src/equilibrate.cpp:188-262) through the C ABI: a
temperature ramp, a compression at constant pressure and an NVE stage run one after another from the state each leaves,
their thermo rows continuous in steps, each summarised by the means over its second half."""
import numpy as np
import pytest

import caps_amd as ca
import helpers as H

pytestmark = pytest.mark.gpu

OPTS = ca.EnergyOptions(cutoff=8.0, coulomb=True, electrostatics="pme", pme_order=4, pme_spacing=1.0, tail=False, parts=5)


def test_stages_run_in_order_and_are_summarised():
    print("case equilibrate: ramp, NPT, NVE", flush=True)
    s = H.melt_case(chains=8, seed=31)
    rows = []

    def progress(index, total, row):
        rows.append((index, row))
        return True

    stages = [ca.Stage("nvt", 1.0, 300.0, 600.0), ca.Stage("npt", 1.0, 600.0, -1.0, 10000.0), ca.Stage("nve", 0.5, 600.0)]
    o = ca.EquilibrateOptions(stages=stages, dt=0.5, tau_t=50, tau_p=200, seed=3, energy=OPTS, progress=progress, thermo_ps=0.01)
    v0 = abs(np.linalg.det(s.cell[3:].reshape(3, 3)))
    out = ca.equilibrate(s, o)
    assert len(out) == 3 and [round(r["ps"], 6) for r in out] == [1.0, 1.0, 0.5]
    steps = [r["step"] for _, r in rows]
    assert steps == sorted(set(steps)) and steps[0] == 0 and steps[-1] == 2000 + 2000 + 1000
    # a row every 0.01 ps = 20 steps of 0.5 fs; the first row of a later stage repeats the last one of the stage before
    assert [i for i, _ in rows].count(0) == 101 and [i for i, _ in rows].count(1) == 100 and [i for i, _ in rows].count(2) == 50
    # the ramp: the second half of a 300 -> 600 K ramp averages near 525 K
    print("summaries:", out, flush=True)
    assert 440 < out[0]["temperature"] < 610
    assert abs(np.linalg.det(s.cell[3:].reshape(3, 3))) != v0
    for k, r in enumerate(out):
        mine = [row for i, row in rows if i == k]
        kept = mine[len(mine) // 2:]
        assert abs(r["potential"] - np.mean([q["potential"] for q in kept])) <= 1e-9 * abs(r["potential"])
        assert r["density_end"] == mine[-1]["density"]
