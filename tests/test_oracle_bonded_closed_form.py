"""The oracle's bonded terms against closed forms at chosen geometries, pinning the conventions of
core/src/field.cpp:1980-2003 (harmonic bond E = k (r - r0)^2 and angle E = k_theta (theta - theta0)^2, no factor 1/2) and
:2150-2180 (FENE -1/2 K R0^2 ln(1 - (r/R0)^2) plus the WCA repulsion 4 eps[(s/r)^12 - (s/r)^6] + eps below 2^(1/6) s, the
FENE energy counted with the bonds), values worked out by hand rather than by the oracle's own formulas: a dropped
1/2, a missing +eps or a wrong WCA switch point fails one of them."""
import math

import numpy as np

import helpers as H
import caps_amd as ca

BIG = np.array([0, 0, 0, 50.0, 0, 0, 0, 50.0, 0, 0, 0, 50.0])


def _system(x, bonds=(), bond_param=(), angles=(), angle_param=(), fene=(), fene_param=()):
    n = len(x)
    ff = ca.ForceField(type=np.zeros(n, np.int32), charge=np.zeros(n), mass=np.ones(n), eps=np.zeros((1, 1)), sigma=np.ones((1, 1)),
                       excl_offset=np.zeros(n + 1, np.int32), excl_index=np.zeros(0, np.int32),
                       bond_atoms=np.array(bonds, np.int32).reshape(-1, 2), bond_param=np.array(bond_param, float).reshape(-1, 2),
                       angle_atoms=np.array(angles, np.int32).reshape(-1, 3), angle_param=np.array(angle_param, float).reshape(-1, 2),
                       fene_atoms=np.array(fene, np.int32).reshape(-1, 2), fene_param=np.array(fene_param, float).reshape(-1, 4))
    return ca.System(x=np.array(x, float) + 10.0, cell=BIG, ff=ff)


def test_harmonic_bond_has_no_half():
    # r = 2.0 A (a 3-4-5 triangle scaled: d = (1.2, 1.6, 0)), k = 300, r0 = 1.5: E = 300 * 0.25 = 75 kcal/mol exactly,
    # |F| = 2 k (r - r0) = 300 kcal/mol/A along d, pulling the atoms together
    s = _system([[0, 0, 0], [1.2, 1.6, 0]], bonds=[[0, 1]], bond_param=[[300.0, 1.5]])
    e, f, w = H.oracle_bonded(s)
    assert abs(e[0] - 75.0) < 1e-12 and e[1] == 0
    assert np.allclose(f[1], [-300 * 0.6, -300 * 0.8, 0], atol=1e-10) and np.allclose(f[0], -f[1], atol=1e-12)
    # virial d (x) f of the pair: d = (1.2, 1.6, 0), f_j = (-180, -240, 0) -> w_xx = -216, w_yy = -384, w_xy = -288
    assert np.allclose(w, [-216.0, -384.0, 0.0, -288.0, 0.0, 0.0], atol=1e-9)


def test_harmonic_angle_has_no_half():
    # theta = 90 deg exactly (u along x, v along y), theta0 = 120 deg, k_theta = 60: E = 60 (pi/6)^2;
    # |F_i| = |dE/dtheta| / |u| = 2 * 60 * (pi/6) / 2 A, along -y (the angle opens as i moves away from k)
    th0 = 2 * math.pi / 3
    s = _system([[2.0, 0, 0], [0, 0, 0], [0, 3.0, 0]], angles=[[0, 1, 2]], angle_param=[[60.0, th0]])
    e, f, _ = H.oracle_bonded(s)
    assert abs(e[1] - 60.0 * (math.pi / 6) ** 2) < 1e-10 and e[0] == 0
    g = 2 * 60.0 * (math.pi / 6)
    assert np.allclose(f[0], [0, -g / 2.0, 0], atol=1e-9)
    assert np.allclose(f[2], [-g / 3.0, 0, 0], atol=1e-9)
    assert np.allclose(f.sum(0), 0, atol=1e-10)


def test_fene_with_and_without_the_wca_part():
    K, R0, eps, sig = 30.0, 1.5, 1.0, 1.0
    # r = 1.0 < 2^(1/6): (s/r)^6 = 1, so the WCA part is 4 (1 - 1) + eps = eps exactly; FENE -1/2 K R0^2 ln(1 - 4/9)
    s = _system([[0, 0, 0], [0, 0, 1.0]], fene=[[0, 1]], fene_param=[[K, R0, eps, sig]])
    e, f, _ = H.oracle_bonded(s)
    assert abs(e[0] - (-0.5 * K * R0 * R0 * math.log(1 - 4.0 / 9.0) + 1.0)) < 1e-12
    # dE/dr = K r / (1 - (r/R0)^2) - 24 eps (2 - 1) / r = 30 / (5/9) - 24 = 30; force on j = -30 along +z
    assert np.allclose(f[1], [0, 0, -30.0], atol=1e-10)
    # r = 1.2 > 2^(1/6) = 1.1225: no WCA part at all (not even the +eps)
    s = _system([[0, 0, 0], [0, 1.2, 0]], fene=[[0, 1]], fene_param=[[K, R0, eps, sig]])
    e, f, _ = H.oracle_bonded(s)
    assert abs(e[0] - (-0.5 * K * R0 * R0 * math.log(1 - 0.64))) < 1e-12
    assert np.allclose(f[1], [0, -K * 1.2 / 0.36, 0], atol=1e-9)
    # the switch sits at 2^(1/6) sigma: just inside, the WCA part is ~0 (continuous), and the +eps shift is what makes it so
    r = 2 ** (1 / 6) * sig * (1 - 1e-9)
    s = _system([[0, 0, 0], [r, 0, 0]], fene=[[0, 1]], fene_param=[[K, R0, eps, sig]])
    e_in, _, _ = H.oracle_bonded(s)
    assert abs(e_in[0] - (-0.5 * K * R0 * R0 * math.log(1 - (r / R0) ** 2))) < 1e-6


def test_fene_past_r0_is_refused():
    s = _system([[0, 0, 0], [1.5, 0, 0]], fene=[[0, 1]], fene_param=[[30.0, 1.5, 1.0, 1.0]])
    try:
        H.oracle_bonded(s)
    except ca.FieldError:
        return
    raise AssertionError("a FENE bond at R0 must be refused")
