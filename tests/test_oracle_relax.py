"""The oracle's minimiser (oracle/orc_relax.py, caps::minimise of core/src/relax.cpp:65-282) on problems with a known
minimum: a quadratic bowl (every method reaches its centre), and harmonic chains, where the four methods — different
directions and step rules — must end at one and the same minimum energy."""
import numpy as np
import pytest

import caps_amd as ca
import helpers as H
import orc_relax as R


@pytest.mark.parametrize("method", ["sd", "cg", "lbfgs", "fire"])
def test_quadratic_bowl(method):
    rng = np.random.default_rng(1)
    k = rng.uniform(50, 400, 3 * 40)
    c = rng.normal(0, 1, k.size)

    def ev(x):
        return float((k * (x - c) ** 2).sum()), -2 * k * (x - c)

    x, e, fmax, it, nev, stop = R.minimise(ev, np.zeros_like(c), method, ftol=1e-3, etol=0.0, max_iterations=20000)
    # E = sum f^2 / (4 k) once every force is below 1e-3
    assert stop == "force" and fmax < 1e-3 and e < k.size * 1e-6 / (4 * k.min())
    assert np.abs(x - c).max() < 1e-3 / 2 / k.min() * 1.01


def test_bonded_chains_every_method_finds_the_same_minimum():
    s = H.melt_case(chains=2, seed=4)
    s = ca.System(x=s.x + np.random.default_rng(2).normal(0, 0.05, s.x.shape), cell=s.cell, ff=s.ff)

    def ev(x):
        e, f, _ = H.oracle_bonded(ca.System(x=x.reshape(-1, 3), cell=s.cell, ff=s.ff))
        return float(e[0] + e[1]), f.reshape(-1)

    e0 = ev(s.x.reshape(-1))[0]
    ends = {}
    for method in ("sd", "cg", "lbfgs", "fire"):
        x, e, fmax, it, nev, stop = R.minimise(ev, s.x, method, ftol=0.05, max_iterations=20000)
        assert stop == "force" and fmax < 0.05
        ends[method] = e
    assert max(ends.values()) < 0.02 * e0
    assert max(ends.values()) - min(ends.values()) < 0.01
