"""The oracle against the committed golden vectors of tests/golden/ (written by tests/golden/make_golden.py from the
reference's own code in oracle/_ref), so the oracle stays pinned where the reference build is absent."""
import os

import numpy as np
import pytest

import caps_amd as ca
import helpers as H
import orcpy as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def _system(d):
    ff = ca.ForceField(type=d["type"], charge=d["charge"], mass=np.full(len(d["type"]), 12.0), eps=d["eps"], sigma=d["sigma"],
                       excl_offset=d["excl_offset"], excl_index=d["excl_index"], p14_offset=d["p14_offset"], p14_index=d["p14_index"],
                       coul14=0.5)
    return ca.System(x=d["x"], cell=d["cell"], ff=ff)


def test_list_and_prune_match_the_golden_vectors():
    d = _load("list_triclinic300.npz")
    o = O.orc_list(d["x"], d["cell"], float(d["rlist"]), d["excl_offset"], d["excl_index"], d["p14_offset"], d["p14_index"], True)
    try:
        for k in O.LIST_FIELDS:
            assert np.array_equal(o[k], d[k]), k
        X, Y, Z, pm, pe, p14 = O.orc_positions_prune(o, d["x_moved"], float(d["prune_radius"]))
        assert np.array_equal(X, d["X"]) and np.array_equal(Y, d["Y"]) and np.array_equal(Z, d["Z"])
        keep = (pm | pe) != 0
        assert np.array_equal(o["cj"][keep], d["p_cj"]) and np.array_equal(pm[keep], d["p_mask"])
        assert np.array_equal(pe[keep], d["p_ex"]) and np.array_equal(p14[keep], d["p_ex14"])
    finally:
        O.orc_list_free(o)


def test_direct_sum_meets_the_reference_kernel_bars():
    d = _load("kernel_triclinic300.npz")
    s = _system(d)
    (ev, ec), f, w = H.direct_pairs(s, float(d["rc"]), 1, float(d["kappa"]), with_excl=True)
    ec0 = H.direct_pairs(s, float(d["rc"]), 1, float(d["kappa"]), with_excl=False)[0][1]
    assert abs(float(d["evdw"]) - ev) / abs(ev) < 2e-6
    assert abs(float(d["ecoul"]) - ec) / abs(ec0) < 2e-6
    assert np.sqrt(((d["f"] - f) ** 2).sum() / (f ** 2).sum()) < 2e-6
    assert np.abs(d["w"] - w).max() / (abs(w[:3].sum()) / 3) < 2e-5


def test_pme_matches_the_golden_vectors():
    d = _load("pme_triclinic300.npz")
    s = _system(d)
    e, f, v = H.oracle_pme(s, float(d["beta"]), d["grid"].astype(np.int32), int(d["order"]))
    assert e == pytest.approx(float(d["energy"]), rel=1e-12)
    assert np.abs(f - d["f"]).max() < 1e-10 * np.abs(d["f"]).max()
    assert np.abs(v - d["vir"]).max() < 1e-10 * np.abs(d["vir"]).max()


def test_random_streams_match_the_golden_vectors():
    d = _load("rng.npz")
    r = O.Rng(int(d["normal_seed"]))
    assert np.array_equal(np.array([r.normal() for _ in range(len(d["normal"]))]), d["normal"])
    r = O.Rng(int(d["bussi_seed"]))
    nd = float(d["bussi_ndof"])
    got = []
    for _ in range(len(d["bussi"]) // 2):
        got += [r.normal(), r.gamma((nd - 1) / 2, 2.0)]
    assert np.array_equal(np.array(got), d["bussi"])


@pytest.mark.parametrize("tag", ["flexible", "stiff"])
def test_kremer_grest_chains_match_the_golden_vectors(tag):
    # the reference builder's bead positions (kremer_grest.cpp:15-83, built alone by oracle/Makefile): the oracle's
    # restatement and the library's host builder draw the same mt19937_64 stream in the same order, bit for bit
    d = _load("kg.npz")
    chains, beads, density, k_theta, bond, seed = d[tag + "_args"]
    ref = d[tag + "_x"]
    x = np.zeros(ref.size)
    O.orc().orc_kremer_grest(int(chains), int(beads), density, k_theta, bond, int(seed), x)
    assert np.array_equal(x.reshape(-1, 3), ref)
    s = ca.kremer_grest(chains=int(chains), beads=int(beads), density=density, k_theta=k_theta, bond=bond, seed=int(seed))
    assert np.array_equal(s.x, ref)


def test_initial_velocities_draw_the_reference_stream_in_atom_order():
    # dynamics.cpp:207-208, 231-247: v[3i + k] = sd_i · gauss(rng) for i, then k, from the run's mt19937_64(seed) and
    # caps::Normal; the centre-of-mass momentum removed; the whole set scaled to exactly T over ndof = 3N - 3. The draws are
    # the reference's own normal stream of seed 123 (golden), so the order of consumption is what is pinned here.
    d = _load("rng.npz")
    g = d["normal"]
    n, T = 600, 300.0
    m = 1.0 + 15.0 * (np.arange(n) % 5) / 4.0
    kB, kAcc = 0.0019872041, 4.184e-4
    v = np.sqrt(kB * T * kAcc / m)[:, None] * g[: 3 * n].reshape(n, 3)
    v -= (m[:, None] * v).sum(0) / m.sum()
    t0 = (m[:, None] * v * v).sum() / kAcc / ((3 * n - 3) * kB)
    v *= np.sqrt(T / t0)
    out = np.zeros(3 * n)
    O.orc().orc_initial_velocities(O.Rng(int(d["normal_seed"])).ptr(), m, n, T, out)
    assert np.allclose(out.reshape(n, 3), v, rtol=1e-12, atol=1e-15)
