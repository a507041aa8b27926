"""The GPU minimiser (caps::minimise, core/src/relax.cpp:65-282) through the C ABI against the oracle's numpy
restatement (oracle/orc_relax.py) driven by the FP64 oracle forces: on the bonded terms (FP64 on both sides) each method
follows the oracle's path — same iterations, evaluations and stop, positions to 1e-8 Å — and with PME's reciprocal part
(FP32 mesh) it ends at the oracle's minimum; held atoms do not move."""
import numpy as np
import pytest

import caps_amd as ca
import helpers as H
import orc_relax as R
import orcpy as O

pytestmark = pytest.mark.gpu

BONDED = ca.EnergyOptions(cutoff=8.0, coulomb=False, parts=1, tail=False)
PME = ca.EnergyOptions(cutoff=8.0, coulomb=True, electrostatics="pme", pme_order=4, pme_spacing=1.0, tail=False, parts=5)


def _perturbed(chains=2, seed=4):
    s = H.melt_case(chains=chains, seed=seed)
    return ca.System(x=s.x + np.random.default_rng(2).normal(0, 0.05, s.x.shape), cell=s.cell, ff=s.ff)


def _oracle_eval(s, pme):
    beta = O.orc().orc_ewald_beta(8.0, 1e-5)
    grid = np.zeros(3, np.int32)
    O.orc().orc_pme_grid(np.ascontiguousarray(s.cell, np.float64), 1.0, 4, grid)
    q = s.ff.charge

    def ev(x):
        sx = ca.System(x=x.reshape(-1, 3), cell=s.cell, ff=s.ff)
        e, f, _ = H.oracle_bonded(sx)
        en, fo = float(e[0] + e[1]), f.reshape(-1)
        if pme:
            er, fr, _ = H.oracle_pme(sx, beta, grid, 4)
            en += er - 332.06371 * beta / np.sqrt(np.pi) * float((q * q).sum())
            fo = fo + fr.reshape(-1)
        return en, fo
    return ev


@pytest.mark.parametrize("method", ["sd", "cg", "lbfgs", "fire"])
def test_minimiser_follows_the_oracle_on_bonded_terms(method):
    print("case minimise %s bonded" % method, flush=True)
    s = _perturbed()
    x0 = s.x.copy()
    ev = ca.Evaluator(s.ff, BONDED)
    st = ca.minimise(ev, s, ca.RelaxOptions(method=method, ftol=0.05, max_iterations=60))
    xo, eo, fo, ito, nevo, stopo = R.minimise(_oracle_eval(s, False), x0, method, ftol=0.05, max_iterations=60)
    print("GPU: %d iterations, %d evaluations, E %.8f, fmax %.4g, %s; oracle: %d, %d, %.8f, %.4g, %s"
          % (st.iterations, st.evaluations, st.energy, st.fmax, st.stopped_by, ito, nevo, eo, fo, stopo), flush=True)
    assert (st.iterations, st.evaluations, st.stopped_by) == (ito, nevo, stopo)
    assert abs(st.energy - eo) <= 1e-9 * max(1.0, abs(eo))
    assert np.abs(s.x.reshape(-1) - xo).max() < 1e-8
    # the result is the context's state
    x, _, _ = ev.get_state()
    assert np.array_equal(x, s.x)


@pytest.mark.parametrize("method", ["lbfgs", "fire"])
def test_minimiser_with_pme_reaches_the_oracles_minimum(method):
    print("case minimise %s bonded + PME" % method, flush=True)
    s = _perturbed(chains=4, seed=5)
    x0 = s.x.copy()
    ev = ca.Evaluator(s.ff, PME)
    st = ca.minimise(ev, s, ca.RelaxOptions(method=method, ftol=0.2, max_iterations=2000))
    xo, eo, fo, ito, nevo, stopo = R.minimise(_oracle_eval(s, True), x0, method, ftol=0.2, max_iterations=2000)
    print("GPU: E %.6f after %d (%s); oracle E %.6f after %d (%s)" % (st.energy, st.iterations, st.stopped_by, eo, ito, stopo), flush=True)
    # FP32 mesh forces: the paths part slowly, both end at (nearly) one minimum; FIRE is still moving after 2000 steps
    assert st.stopped_by == stopo
    assert abs(st.energy - eo) <= 2e-3 * abs(eo)


def test_held_atoms_do_not_move():
    print("case minimise with held atoms", flush=True)
    s = _perturbed()
    x0 = s.x.copy()
    fixed = np.zeros(s.ff.n, np.uint8)
    fixed[::3] = 1
    ca.minimise(ca.Evaluator(s.ff, BONDED), s, ca.RelaxOptions(method="lbfgs", ftol=0.05, fixed=fixed))
    assert np.array_equal(s.x[fixed == 1], x0[fixed == 1])
    assert not np.array_equal(s.x[fixed == 0], x0[fixed == 0])
