"""The oracle's random numbers, thermostat draw, Kremer–Grest builder and MLP (oracle/orc_dyn.c), pinned by the C++
standard's mt19937_64 check value, the reference's own rng.hpp streams (compiled into oracle/_ref), distribution moments,
Bussi's limits and mean, tests/test_provenance.cpp's KG invariants, and central differences of the MLP loss."""
import numpy as np
import pytest

import caps_amd as ca
import orcpy as O

KB = 0.0019872041
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="the reference build oracle/_ref is absent (no reference sources)")


def test_mt19937_64_check_value():
    # [rand.predef]: the 10000th consecutive invocation of a default-constructed mt19937_64 returns 9981545732273789042
    r = O.Rng(5489)
    for _ in range(9999):
        r.next()
    assert r.next() == 9981545732273789042


def test_uniform01_takes_the_top_53_bits():
    # rng.hpp:26-33
    a, b = O.Rng(42), O.Rng(42)
    for _ in range(100):
        assert a.uniform01() == (b.next() >> 11) * 2.0 ** -53


def test_normal_and_gamma_moments():
    r = O.Rng(7)
    z = np.array([r.normal() for _ in range(200000)])
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1) < 0.01
    # polar method: the draws come in pairs sharing one radius, u·f and v·f with u² + v² < 1
    for shape, scale in ((0.7, 2.0), (12.5, 2.0), (1500.0, 2.0)):
        g = np.array([r.gamma(shape, scale) for _ in range(100000)])
        assert g.mean() == pytest.approx(shape * scale, rel=0.02)
        assert g.var() == pytest.approx(shape * scale * scale, rel=0.05)
        assert (g > 0).all()


def test_bussi_limits_and_mean():
    k, kt, ndof = 120.0, 150.0, 300.0
    # c = 1 (τ → ∞): the kinetic energy stays; c = 0: it is redrawn from the canonical distribution
    assert O.orc().orc_bussi_alpha(k, kt, ndof, 1.0, 0.3, 250.0) == pytest.approx(1.0, abs=1e-15)
    r1, sm = 0.7, 280.0
    assert O.orc().orc_bussi_alpha(k, kt, ndof, 0.0, r1, sm) ** 2 * k == pytest.approx(kt * (r1 * r1 + sm) / ndof, rel=1e-12)
    # E[K_new] = c K + (1 − c) K_target (Bussi, Donadio & Parrinello 2007, eq. A7)
    rng = O.Rng(9)
    c = 0.8
    knew = []
    for _ in range(20000):
        r1 = rng.normal()
        sm = rng.gamma((ndof - 1) / 2, 2.0)
        knew.append(O.orc().orc_bussi_alpha(k, kt, ndof, c, r1, sm) ** 2 * k)
    assert np.mean(knew) == pytest.approx(c * k + (1 - c) * kt, rel=0.01)


def test_initial_velocities_have_no_drift_and_the_temperature():
    n = 500
    m = np.where(np.arange(n) % 3 == 0, 12.011, 1.008)
    v = np.zeros(3 * n)
    O.orc().orc_initial_velocities(O.Rng(3).ptr(), m, n, 300.0, v)
    p = (m[:, None] * v.reshape(-1, 3)).sum(0)
    assert np.abs(p).max() < 1e-10
    ke = O.orc().orc_kinetic_energy(v, m, n)
    assert 2 * ke / ((3 * n - 3) * KB) == pytest.approx(300.0, rel=1e-12)


def test_kremer_grest_chains():
    # test_provenance.cpp:315-335: 10 chains × 30 beads, k_θ = 1.5
    x = np.zeros(3 * 300)
    box = O.orc().orc_kremer_grest(10, 30, 0.85, 1.5, 0.97, 1, x)
    assert box == pytest.approx(np.cbrt(300 / 0.85), abs=1e-9)
    x = x.reshape(10, 30, 3)
    d = np.diff(x, axis=1)
    assert np.abs(np.linalg.norm(d, axis=2) - 0.97).max() < 1e-9
    cos = (d[:, :-1] * d[:, 1:]).sum(2) / 0.97 ** 2
    assert cos.min() >= -0.5 - 1e-9   # no folding back past 120° between bonds


def test_the_products_kg_builder_is_the_oracles():
    s = ca.kremer_grest(chains=6, beads=25, density=0.85, k_theta=1.5, seed=4)
    x = np.zeros(3 * 150)
    box = O.orc().orc_kremer_grest(6, 25, 0.85, 1.5, 0.97, 4, x)
    assert s.cell[3] == box
    assert np.array_equal(s.x.reshape(-1), x)


def _mlp_params(rng):
    return [rng.normal(0, 0.1, (128, 96)), rng.normal(0, 0.1, 128), rng.normal(0, 0.1, (128, 128)), rng.normal(0, 0.1, 128),
            rng.normal(0, 0.1, (3, 128)), rng.normal(0, 0.1, 3)]


def _loss(P, x, t):
    y = np.zeros((len(x), 3))
    O.orc().orc_mlp_forward(*[np.ascontiguousarray(p).reshape(-1) for p in P], np.ascontiguousarray(x).reshape(-1), len(x), y.reshape(-1))
    return 0.5 * ((y - t) ** 2).sum() / len(x), y


def test_mlp_forward_is_silu_layers_and_backward_its_gradient():
    rng = np.random.default_rng(0)
    P = _mlp_params(rng)
    x, t = rng.normal(0, 1, (16, 96)), rng.normal(0, 1, (16, 3))
    silu = lambda z: z / (1 + np.exp(-z))
    h1 = silu(x @ P[0].T + P[1])
    h2 = silu(h1 @ P[2].T + P[3])
    L, y = _loss(P, x, t)
    assert np.abs(y - (h2 @ P[4].T + P[5])).max() < 1e-12
    G = [np.zeros_like(p).reshape(-1) for p in P]
    Lb = O.orc().orc_mlp_backward(*[np.ascontiguousarray(p).reshape(-1) for p in P], np.ascontiguousarray(x).reshape(-1),
                                  np.ascontiguousarray(t).reshape(-1), len(x), *G)
    assert Lb == pytest.approx(L, rel=1e-12)
    h = 1e-6
    for li, idx in ((0, (5, 17)), (1, (9,)), (2, (3, 100)), (3, (64,)), (4, (2, 7)), (5, (1,))):
        Pp = [p.copy() for p in P]
        Pm = [p.copy() for p in P]
        Pp[li][idx] += h
        Pm[li][idx] -= h
        fd = (_loss(Pp, x, t)[0] - _loss(Pm, x, t)[0]) / (2 * h)
        flat = np.ravel_multi_index(idx, P[li].shape) if len(idx) > 1 else idx[0]
        assert G[li][flat] == pytest.approx(fd, rel=1e-6, abs=1e-9)


@needs_ref
def test_normal_stream_is_the_references():
    # caps::Normal (rng.hpp:70-89), the polar method with its spare, compiled from the reference header
    n = 5001
    ref = np.zeros(n)
    O.ref().ref_rng_normal(123, n, ref)
    r = O.Rng(123)
    assert np.array_equal(np.array([r.normal() for _ in range(n)]), ref)


@needs_ref
def test_bussi_draws_are_the_references():
    # dynamics.cpp:466-477: r1 from the run's normal generator, then a fresh caps::Gamma((ndof − 1)/2, 2) per step
    steps, ndof = 300, 3.0 * 512 - 3
    ref = np.zeros(2 * steps)
    O.ref().ref_rng_bussi(77, ndof, steps, ref)
    r = O.Rng(77)
    got = []
    for _ in range(steps):
        got.append(r.normal())
        got.append(r.gamma((ndof - 1) / 2, 2.0))
    assert np.array_equal(np.array(got), ref)


@needs_ref
def test_uniform_stream_is_the_references():
    ref = np.zeros(1000)
    O.ref().ref_rng_uniform(5, 0.0, 1.0, 1000, ref)
    r = O.Rng(5)
    assert np.array_equal(np.array([r.uniform01() for _ in range(1000)]), ref)


def test_descriptor_bins_are_hat_functions_with_the_cosine_cutoff():
    # two atoms at distance r: only the two bins around r/Δ (Δ = 8/23) hold q_j·f(r)·(1 − frac, frac); the vector
    # channels carry the same times the unit vector
    L = 30.0
    cell = np.array([0, 0, 0, L, 0, 0, 0, L, 0, 0, 0, L], float)
    for r in (0.9, 3.3, 5.0, 7.95):
        x = np.array([[5.0, 5.0, 5.0], [5.0 + r * 0.6, 5.0 + r * 0.8, 5.0]])
        q = np.array([0.3, -0.7])
        out = np.zeros(2 * 96)
        O.orc().orc_descriptors(x.reshape(-1).copy(), q, 2, cell, out)
        d0 = out[:96]
        u = r / (8.0 / 23.0)
        k0, fr = int(np.floor(u)), u - np.floor(u)
        fc = 0.5 * (1 + np.cos(np.pi * r / 8.0))
        expect = np.zeros(24)
        expect[k0] = q[1] * fc * (1 - fr)
        if k0 + 1 < 24:
            expect[k0 + 1] = q[1] * fc * fr
        assert np.abs(d0[:24] - expect).max() < 1e-12
        assert np.abs(d0[24:48] - 0.6 * expect).max() < 1e-12 and np.abs(d0[48:72] - 0.8 * expect).max() < 1e-12
        assert np.abs(d0[72:]).max() < 1e-12


def test_descriptor_sums_are_the_cutoff_weighted_charges():
    # the hat functions sum to one on [0, 8): Σ_k scalar channel = Σ_j q_j f(r_ij), Σ_k vector channel = Σ_j q_j f d̂
    rng = np.random.default_rng(2)
    L = 25.0
    n = 150
    x = rng.uniform(0, L, (n, 3))
    q = rng.normal(0, 0.3, n)
    cell = np.array([0, 0, 0, L, 0, 0, 0, L, 0, 0, 0, L], float)
    out = np.zeros(n * 96)
    O.orc().orc_descriptors(x.reshape(-1).copy(), q, n, cell, out)
    D = out.reshape(n, 96)
    for i in (0, 17, 99):
        d = x - x[i]
        d -= L * np.round(d / L)
        r = np.linalg.norm(d, axis=1)
        m = (r > 1e-3) & (r < 8.0)
        fc = 0.5 * (1 + np.cos(np.pi * r[m] / 8.0))
        assert D[i, :24].sum() == pytest.approx((q[m] * fc).sum(), abs=1e-12)
        for c in range(3):
            assert D[i, 24 + 24 * c:48 + 24 * c].sum() == pytest.approx((q[m] * fc * d[m, c] / r[m]).sum(), abs=1e-12)
