"""The oracle's k-space restatement (oracle/orc_pme.c) pinned by tests/test_kspace.cpp's golden values and bars (FFT
against the direct transform, good sizes, the Madelung constant of rock salt, PME against the plain Ewald sum), by
derivatives of the energy, and against the reference's own Fortran PME compiled into oracle/_ref."""
import numpy as np
import pytest

import caps_amd as ca
import helpers as H
import orcpy as O

KC = 332.06371
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="the reference build oracle/_ref is absent (no reference sources)")


def test_fft_matches_the_direct_transform():
    # test_kspace.cpp:61-83: radix 7 alone, after 2, 3, 5 and twice; the transform's sign is +1 (exp(+2πi mk/n))
    rng = np.random.default_rng(3)
    for k1 in (6, 7, 10, 14, 21, 35, 36, 45, 49, 50, 56):
        k2, k3 = 5, 4
        a = rng.uniform(-1, 1, (k3, k2, k1)) + 1j * rng.uniform(-1, 1, (k3, k2, k1))
        re, im = np.ascontiguousarray(a.real).reshape(-1), np.ascontiguousarray(a.imag).reshape(-1)
        O.orc().orc_fft3d(k1, k2, k3, +1, re, im)
        ref = np.fft.ifftn(a) * a.size   # numpy's inverse has the + sign
        assert np.abs((re + 1j * im).reshape(a.shape) - ref).max() < 1e-10, k1


def test_good_sizes():
    # test_kspace.cpp:80-83
    for n, g in ((31, 32), (49, 49), (13, 14), (43, 45)):
        assert O.orc().orc_good_size(n) == g
    if O.ref_available():
        for n in range(2, 600):
            assert O.orc().orc_good_size(n) == O.ref().ref_good_size(n), n


def _rock_salt(cells, a0):
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    x, q = [], []
    for i in range(cells):
        for j in range(cells):
            for k in range(cells):
                for b in base:
                    x.append((np.array([i, j, k]) + b) * a0)
                    q.append(1.0)
                    x.append((np.array([i, j, k]) + b + [0.5, 0, 0]) * a0)
                    q.append(-1.0)
    return np.array(x), np.array(q)


def _ewald_total(x, q, cell, beta, rc, recip):
    """real space (direct sum with erfc) + reciprocal + self"""
    n = len(q)
    ff = ca.ForceField(type=np.zeros(n, np.int32), charge=q, mass=np.ones(n), eps=np.zeros((1, 1)), sigma=np.ones((1, 1)),
                       excl_offset=np.zeros(n + 1, np.int32), excl_index=np.zeros(1, np.int32))
    (_, er), _, _ = H.direct_pairs(ca.System(x=x, cell=cell, ff=ff), rc, 1, beta, with_excl=False)
    return er - KC * beta / np.sqrt(np.pi) * (q * q).sum() + recip


def test_madelung_constant_of_rock_salt():
    # test_kspace.cpp:87-104: E per ion pair = −M kC / r0, r0 = a0/2, M = 1.747565
    a0 = 5.64
    x, q = _rock_salt(2, a0)
    L = 2 * a0
    cell = np.array([0, 0, 0, L, 0, 0, 0, L, 0, 0, 0, L], float)
    rc = 5.5
    beta = O.orc().orc_ewald_beta(rc, 1e-8)
    f, vir = np.zeros(x.size), np.zeros(6)
    er = O.orc().orc_ewald_recip(x.reshape(-1).copy(), q, len(q), cell, np.array([12, 12, 12], np.int32), beta, KC, f, vir)
    e = _ewald_total(x, q, cell, beta, rc, er)
    M = -e / (len(q) / 2) * (a0 / 2) / KC
    assert M == pytest.approx(1.747565, abs=2e-5)
    grid = np.zeros(3, np.int32)
    O.orc().orc_pme_grid(cell, 0.3, 6, grid)
    ep = O.orc().orc_pme(x.reshape(-1).copy(), q, len(q), cell, grid, 6, beta, KC, np.zeros(x.size), np.zeros(6))
    assert _ewald_total(x, q, cell, beta, rc, ep) == pytest.approx(e, rel=1e-4)


def _triclinic_charges(n=200, seed=11):
    rng = np.random.default_rng(seed)
    cell = np.array([0, 0, 0, 20, 0, 0, 3, 19, 0, -2, 4, 21], float)
    s = rng.uniform(0, 1, (n, 3))
    x = s @ cell[3:].reshape(3, 3)
    q = np.where(np.arange(n) % 2 == 1, 0.4, -0.4)
    return x, q, cell


def test_pme_matches_the_ewald_sum_for_forces_energy_and_virial():
    # test_kspace.cpp:106-158
    x, q, cell = _triclinic_charges()
    beta = O.orc().orc_ewald_beta(9.0, 1e-5)
    kmax = np.array([14, 14, 14], np.int32)
    fe, ve = np.zeros(x.size), np.zeros(6)
    ee = O.orc().orc_ewald_recip(x.reshape(-1).copy(), q, len(q), cell, kmax, beta, KC, fe, ve)
    grid = np.zeros(3, np.int32)
    O.orc().orc_pme_grid(cell, 0.8, 6, grid)
    fp, vp = np.zeros(x.size), np.zeros(6)
    ep = O.orc().orc_pme(x.reshape(-1).copy(), q, len(q), cell, grid, 6, beta, KC, fp, vp)
    assert ep == pytest.approx(ee, abs=1e-4 * abs(ee) + 1e-4)
    assert np.sqrt(((fp - fe) ** 2).sum() / (fe ** 2).sum()) < 1e-3
    assert np.abs(vp - ve).max() < 1e-3 * abs(ee) + 1e-3

    def recip(y, c):
        return O.orc().orc_ewald_recip(y.reshape(-1).copy(), q, len(q), c, kmax, beta, KC, np.zeros(y.size), np.zeros(6))
    h = 1e-5
    for k in (0, 7, 301):
        xp, xm = x.reshape(-1).copy(), x.reshape(-1).copy()
        xp[k] += h
        xm[k] -= h
        dE = (recip(xp, cell) - recip(xm, cell)) / (2 * h)
        assert fe[k] == pytest.approx(-dE, abs=1e-5 * (1 + abs(dE)))
    eps = 1e-5

    def scaled(lam):
        c = cell.copy()
        c[3:] *= lam
        return recip(x * lam, c)
    assert ve[:3].sum() == pytest.approx(-(scaled(1 + eps) - scaled(1 - eps)) / (2 * eps), abs=1e-5 * abs(ee) + 1e-6)


@needs_ref
@pytest.mark.parametrize("order", [4, 5])
def test_pme_matches_the_reference_fortran(order):
    x, q, cell = _triclinic_charges(n=300, seed=5)
    beta = O.orc().orc_ewald_beta(9.0, 1e-5)
    assert beta == O.ref().ref_ewald_beta(9.0, 1e-5)
    grid, gref = np.zeros(3, np.int32), np.zeros(3, np.int32)
    O.orc().orc_pme_grid(cell, 1.0, order, grid)
    O.ref().ref_pme_grid(cell, beta, 1.0, order, gref)
    assert np.array_equal(grid, gref)
    f, v = np.zeros(x.size), np.zeros(6)
    e = O.orc().orc_pme(x.reshape(-1).copy(), q, len(q), cell, grid, order, beta, KC, f, v)
    fr, vr = np.zeros(x.size), np.zeros(6)
    er = O.ref().ref_pme(x.reshape(-1).copy(), q, len(q), cell, grid, order, beta, 1, fr, vr)
    assert e == pytest.approx(er, rel=1e-12)
    assert np.abs(f - fr).max() < 1e-10 * np.abs(fr).max()
    assert np.abs(v - vr).max() < 1e-10 * np.abs(vr).max()


def test_ewald_beta_meets_the_tolerance():
    from scipy.special import erfc
    for rc, tol in ((9.0, 1e-5), (12.0, 1e-5), (10.0, 1e-6)):
        b = O.orc().orc_ewald_beta(rc, tol)
        assert erfc(b * rc) == pytest.approx(tol, rel=1e-9)
