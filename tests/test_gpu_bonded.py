"""GPU parity of the bonded terms (harmonic bonds and angles, FENE + WCA; field.cpp:1980-2003, 2150-2180) through the C
ABI, against the oracle's FP64 restatement. Both sides are FP64; only the summation order differs."""
import numpy as np
import pytest

import caps_amd as ca
import helpers as H

pytestmark = pytest.mark.gpu


def _bonded_only(s):
    ev = ca.Evaluator(s.ff, ca.EnergyOptions(cutoff=10.0, coulomb=False, parts=1, tail=False))
    return ev.compute(s.x, s.cell)


@pytest.mark.parametrize("which", ["melt", "kg"])
def test_bonded_terms_match_the_oracle(which):
    s = H.melt_case(chains=16) if which == "melt" else ca.kremer_grest(chains=40, beads=50, density=0.85, seed=5)
    print("case", which, s.ff.n, flush=True)
    t, f = _bonded_only(s)
    e, fo, wo = H.oracle_bonded(s)
    assert t.bond == pytest.approx(e[0], rel=1e-12, abs=1e-9)
    assert t.angle == pytest.approx(e[1], rel=1e-12, abs=1e-9)
    assert t.virial == pytest.approx(e[2], rel=1e-10, abs=1e-8)
    assert np.abs(np.array(t.w) - wo).max() <= 1e-10 * np.abs(wo).max() + 1e-8
    assert np.abs(f - fo).max() <= 1e-10 * np.abs(fo).max() + 1e-10
    assert t.vdw == 0 and t.coulomb == 0


def test_fene_past_r0_is_a_field_error():
    s = ca.kremer_grest(chains=20, beads=10, density=0.85, seed=2)
    s.x[1] = s.x[0] + np.array([1.6, 0, 0])
    print("case fene past R0", flush=True)
    with pytest.raises(ca.FieldError):
        _bonded_only(s)
    # the context stays usable and the flag is cleared
    s2 = ca.kremer_grest(chains=20, beads=10, density=0.85, seed=2)
    ev = ca.Evaluator(s2.ff, ca.EnergyOptions(cutoff=10.0, coulomb=False, parts=1, tail=False))
    with pytest.raises(ca.FieldError):
        ev.compute(s.x, s.cell)
    t, f = ev.compute(s2.x, s2.cell)
    assert np.isfinite(f).all()


@pytest.mark.parametrize("which", ["262k melt", "kg 10k"])
def test_bonded_terms_at_full_size(which):
    # the 262,144-atom melt (bonds and angles) and the 10,000-bead Kremer-Grest melt (FENE + WCA) of the benchmark
    # configurations
    s = ca.build_melt(2048, seed=1) if which == "262k melt" else ca.kremer_grest(chains=200, beads=50, density=0.85, seed=1)
    print("case full-size", which, s.ff.n, flush=True)
    t, f = _bonded_only(s)
    e, fo, wo = H.oracle_bonded(s)
    # both FP64; the sums over ~2.6e5 terms run in different orders, so the energies agree to about N·ε ≈ 3e-11
    assert t.bond == pytest.approx(e[0], rel=1e-10, abs=1e-9)
    assert t.angle == pytest.approx(e[1], rel=1e-10, abs=1e-9)
    assert t.virial == pytest.approx(e[2], rel=1e-10, abs=1e-8)
    assert np.abs(f - fo).max() <= 1e-10 * np.abs(fo).max() + 1e-10
