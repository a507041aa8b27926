"""The oracle's pair sum, tail correction and bonded terms (oracle/orc_forces.c) pinned against the reference's own
cluster kernel (oracle/_ref) with tests/nbnxm_check.cpp's cases and bars, and against closed forms: numerical
integrals of the tail, central differences of the energy, and the virial as the strain derivative."""
import numpy as np
import pytest
from scipy import integrate

import caps_amd as ca
import helpers as H
import orcpy as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="the reference build oracle/_ref is absent (no reference sources)")

# tests/nbnxm_check.cpp:59-60: name, 9-6, coulomb (1 Ewald, 2 DSF), κ, shifted LJ
KERNEL_CASES = [("12-6 Ewald", False, 1, 0.32, False), ("9-6 DSF shifted", True, 2, 0.2, True), ("12-6 Ewald erfc", False, 1, 0.6, False),
                ("9-6 no charges", True, 0, 0.0, False)]


@needs_ref
@pytest.mark.parametrize("name,lj96,coul,kappa,shift", KERNEL_CASES, ids=[c[0] for c in KERNEL_CASES])
def test_direct_sum_matches_the_reference_kernel(name, lj96, coul, kappa, shift):
    s = H.triclinic_case(n=1100)
    s.ff.lj96 = lj96
    rc, skin = 8.0, 2.0
    dsf = H.dsf_constants(kappa, rc) if coul == 2 else (0.0, 0.0)
    (ev, ec), f, w = H.direct_pairs(s, rc, coul, kappa, shift=shift, with_excl=True, dsf=dsf)
    (rv, rcl), fr, wr = H.ref_pairs(s, rc, skin, coul, kappa, shift=shift, dsf=dsf)
    # nbnxm_check.cpp:176-183: relative errors of the FP32 kernel against the FP64 direct sum, the Coulomb error relative
    # to the Coulomb energy without the excluded pairs' corrections
    assert abs(rv - ev) / abs(ev) < 2e-6
    if coul:
        ec0 = H.direct_pairs(s, rc, coul, kappa, shift=shift, with_excl=False, dsf=dsf)[0][1]
        assert abs(rcl - ec) / abs(ec0) < (2e-5 if kappa * rc > 4 else 2e-6)
    else:
        assert ec == 0 and rcl == 0
    assert np.sqrt(((fr - f) ** 2).sum() / (f ** 2).sum()) < 2e-6
    assert np.abs(wr - w).max() / (abs(w[:3].sum()) / 3) < 2e-5


def test_direct_sum_forces_are_minus_the_energy_gradient():
    s = H.triclinic_case(n=200)
    rc, kappa = 8.0, 0.32
    (ev, ec), f, w = H.direct_pairs(s, rc, 1, kappa, shift=True, with_excl=True)
    h = 1e-5
    for i, k in [(3, 0), (77, 1), (150, 2)]:
        xp, xm = s.x.copy(), s.x.copy()
        xp[i, k] += h
        xm[i, k] -= h
        ep = sum(H.direct_pairs(H.ca.System(x=xp, cell=s.cell, ff=s.ff), rc, 1, kappa, shift=True)[0])
        em = sum(H.direct_pairs(H.ca.System(x=xm, cell=s.cell, ff=s.ff), rc, 1, kappa, shift=True)[0])
        assert f[i, k] == pytest.approx(-(ep - em) / (2 * h), rel=1e-5, abs=1e-6)


@pytest.mark.parametrize("lj96", [False, True])
def test_tail_correction_is_the_integral_beyond_the_cutoff(lj96):
    nt = 2
    eps = np.array([[0.1, 0.07], [0.07, 0.05]])
    sig = np.array([[3.4, 3.1], [3.1, 2.8]])
    cnt = np.array([300.0, 200.0])
    rc, V = 9.0, 40000.0
    out = np.zeros(2)
    O.orc().orc_tail(nt, eps.reshape(-1).copy(), (sig ** 6).reshape(-1).copy(), cnt, 2 if lj96 else 1, rc, V, out)
    e = wv = 0.0
    for a in range(nt):
        for b in range(nt):
            if lj96:
                U = lambda r: eps[a, b] * (2 * (sig[a, b] / r) ** 9 - 3 * (sig[a, b] / r) ** 6)
                dU = lambda r: eps[a, b] * (-18 * sig[a, b] ** 9 / r ** 10 + 18 * sig[a, b] ** 6 / r ** 7)
            else:
                U = lambda r: 4 * eps[a, b] * ((sig[a, b] / r) ** 12 - (sig[a, b] / r) ** 6)
                dU = lambda r: 4 * eps[a, b] * (-12 * sig[a, b] ** 12 / r ** 13 + 6 * sig[a, b] ** 6 / r ** 7)
            e += 2 * np.pi * cnt[a] * cnt[b] / V * integrate.quad(lambda r: r * r * U(r), rc, np.inf)[0]
            wv += -2 * np.pi / 3 * cnt[a] * cnt[b] / V * integrate.quad(lambda r: r ** 3 * dU(r), rc, np.inf)[0] * 3
    assert out[0] == pytest.approx(e, rel=1e-9)
    assert out[1] == pytest.approx(wv, rel=1e-9)


def _bonded_energy(s):
    e = H.oracle_bonded(s)[0]   # bond (with FENE), angle, scalar virial
    return e[0] + e[1]


@pytest.mark.parametrize("which", ["melt", "kg"])
def test_bonded_forces_and_virial_are_derivatives_of_the_energy(which):
    s = H.melt_case(chains=2) if which == "melt" else ca.kremer_grest(chains=4, beads=20, density=0.85, seed=3)
    e, f, w = H.oracle_bonded(s)
    h = 1e-6
    rng = np.random.default_rng(1)
    for i in rng.choice(len(s.x), 6, replace=False):
        for k in range(3):
            xp, xm = s.x.copy(), s.x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            d = (_bonded_energy(ca.System(x=xp, cell=s.cell, ff=s.ff)) - _bonded_energy(ca.System(x=xm, cell=s.cell, ff=s.ff))) / (2 * h)
            assert f[i, k] == pytest.approx(-d, rel=1e-6, abs=1e-6)
    # the trace of the virial is −dE/dλ under x → λx, cell → λ·cell (bonded terms see only in-chain vectors)
    def scaled(lam):
        c = s.cell.copy()
        c[3:] *= lam
        return _bonded_energy(ca.System(x=s.x * lam, cell=c, ff=s.ff))
    eps = 1e-6
    assert w[:3].sum() == pytest.approx(-(scaled(1 + eps) - scaled(1 - eps)) / (2 * eps), rel=1e-5, abs=1e-6)
    assert e[2] == pytest.approx(w[:3].sum(), rel=1e-12, abs=1e-12)


def test_fene_past_r0_is_a_field_error():
    s = ca.kremer_grest(chains=20, beads=10, density=0.85, seed=2)   # L ≈ 6.2σ: the stretched bond is its own nearest image
    s.x[1] = s.x[0] + np.array([1.6, 0, 0])   # R0 = 1.5
    with pytest.raises(ca.FieldError):
        H.oracle_bonded(s)
