"""GPU parity of PME's reciprocal part (spline, spread, 3-D FFT, convolution, gather; caps_kspace.f90:625-715) with the
Ewald self and background terms (field.cpp:2442-2467), through the C ABI (parts = 4), against the oracle's FP64 SPME.
The mesh is FP32 on the GPU: forces within 1e-5 relative RMS, energy and virial within 1e-5 relative."""
import numpy as np
import pytest

import caps_amd as ca
import helpers as H
import orcpy as O

pytestmark = pytest.mark.gpu
KC = 332.06371


def _oracle_recip(s, rc, rtol, spacing, order):
    beta = O.orc().orc_ewald_beta(rc, rtol)
    grid = np.zeros(3, np.int32)
    O.orc().orc_pme_grid(np.ascontiguousarray(s.cell, np.float64), spacing, order, grid)
    e, f, v = H.oracle_pme(s, beta, grid, order)
    q = s.ff.charge
    e -= KC * beta / np.sqrt(np.pi) * (q * q).sum()
    qt = q.sum()
    if abs(qt) > 1e-10:
        V = abs(np.linalg.det(s.cell[3:].reshape(3, 3)))
        eb = -KC * np.pi * qt * qt / (2 * V * beta * beta)
        e += eb
        v = v.copy()
        v[:3] += eb
    return e, f, v, grid


# the grids: 27 (radix 3), 24 (8 · 3); spacing 0.7 on the triclinic cell gives 35 × 32 × 35 (radices 5, 7, 8, 4) and
# 0.45 on the melt 56 (8 · 7); on the melt 0.70 gives 36 and 1.26 gives 20 (9 and 5 tiles a side: the block spread's
# last block along each axis one tile deep)
CASES = [("melt order 4", "melt", 4, 1.0), ("melt order 5", "melt", 5, 1.0), ("triclinic order 5", "tri", 5, 1.0),
         ("triclinic order 6", "tri", 6, 1.0), ("triclinic order 5 radix 5 and 7", "tri", 5, 0.7), ("melt order 8 radix 7", "melt", 8, 0.45),
         ("melt order 5 odd tile count 36", "melt", 5, 0.70), ("melt order 4 odd tile count 20", "melt", 4, 1.26)]


@pytest.mark.parametrize("name,which,order,spacing", CASES, ids=[c[0] for c in CASES])
def test_reciprocal_part_matches_the_oracle(name, which, order, spacing):
    s = H.melt_case(chains=16) if which == "melt" else H.triclinic_case(n=1100)
    print("case", name, s.ff.n, flush=True)
    rc = 10.0
    op = ca.EnergyOptions(cutoff=rc, coulomb=True, electrostatics="pme", ewald_rtol=1e-5, pme_order=order, pme_spacing=spacing,
                          tail=False, parts=4)
    t, f = ca.Evaluator(s.ff, op).compute(s.x, s.cell)
    e, fo, vo, grid = _oracle_recip(s, rc, 1e-5, spacing, order)
    print("grid", grid, "E", t.coulomb, e, flush=True)
    assert t.coulomb == pytest.approx(e, rel=1e-5)
    assert np.sqrt(((f - fo) ** 2).sum() / (fo ** 2).sum()) < 1e-5
    assert np.abs(np.array(t.w) - vo).max() < 1e-5 * np.abs(vo).max()
    assert t.vdw == 0 and t.bond == 0


def test_reciprocal_forces_follow_the_atoms():
    # a rigid translation by a lattice vector changes nothing; by half a grid step, forces move with the atoms
    s = H.triclinic_case(n=600)
    op = ca.EnergyOptions(cutoff=9.0, coulomb=True, electrostatics="pme", pme_order=5, tail=False, parts=4)
    ev = ca.Evaluator(s.ff, op)
    t0, f0 = ev.compute(s.x, s.cell)
    t1, f1 = ev.compute(s.x + s.cell[3:6], s.cell)
    assert t1.coulomb == pytest.approx(t0.coulomb, rel=1e-6)
    assert np.sqrt(((f1 - f0) ** 2).sum() / (f0 ** 2).sum()) < 1e-5


@pytest.mark.parametrize("case,spacing", [("tri", 0.7), ("melt", 0.45)], ids=["35x32x35", "56^3"])
def test_every_transform_tile_shape_gives_the_same_forces(case, spacing):
    # the tile shape (lines per workgroup) only regroups the lines: forces bit-identical, energies to summation order;
    # the tuner keeps one of the candidates and the forces stay the same
    print("case transform tiles", case, flush=True)
    s = H.triclinic_case(n=900) if case == "tri" else H.melt_case(chains=16)
    op = ca.EnergyOptions(cutoff=9.0, coulomb=True, electrostatics="pme", pme_order=5, pme_spacing=spacing, tail=False, parts=4)
    ev = ca.Evaluator(s.ff, op)
    t0, f0 = ev.compute(s.x, s.cell)
    for lx, ly in [(1, 1), (2, 4), (16, 8), (4, 16), (0, 8)]:
        assert ev.set_tiles(lx, ly) == (lx, ly)
        t, f = ev.compute(s.x, s.cell)
        assert np.array_equal(f, f0), (lx, ly)
        assert t.coulomb == pytest.approx(t0.coulomb, rel=1e-12, abs=1e-9)
    lx, ly = ev.tune()
    print("tuned tiles", lx, ly, flush=True)
    assert lx in (0, 2, 4, 8, 16) and ly in (4, 8, 16)
    t, f = ev.compute(s.x, s.cell)
    assert np.array_equal(f, f0)
    with pytest.raises(ValueError, match="powers of two"):
        ev.set_tiles(3, 8)


@pytest.mark.parametrize("chains,order", [(256, 4), (2048, 5)], ids=["32k order 4 on 64^3", "262k order 5 on 128^3"])
def test_reciprocal_part_at_full_size(chains, order):
    # the 32,768- and 262,144-atom melts of the benchmark workloads, against the oracle's FP64 PME (seconds on the CPU)
    s = ca.build_melt(chains, seed=1)
    print("case full-size melt", s.ff.n, "order", order, flush=True)
    rc = 12.0
    op = ca.EnergyOptions(cutoff=rc, coulomb=True, electrostatics="pme", ewald_rtol=1e-5, pme_order=order, pme_spacing=1.0,
                          tail=False, parts=4)
    t, f = ca.Evaluator(s.ff, op).compute(s.x, s.cell)
    e, fo, vo, grid = _oracle_recip(s, rc, 1e-5, 1.0, order)
    err = np.sqrt(((f - fo) ** 2).sum() / (fo ** 2).sum())
    print("grid", grid, "E", t.coulomb, e, "force rel RMS", err, flush=True)
    assert t.coulomb == pytest.approx(e, rel=1e-5)
    assert err < 1e-5
    assert np.abs(np.array(t.w) - vo).max() < 1e-5 * np.abs(vo).max()


@pytest.mark.parametrize("case,order", [("tri", 5), ("tri", 3), ("32k", 4), ("32k", 6), ("32k", 8), ("262k", 5)],
                         ids=["triclinic order 5 on 24^3", "triclinic order 3", "32k order 4 on 64^3", "32k order 6", "32k order 8",
                              "262k order 5 on 128^3"])
def test_gather_by_runs_gives_the_same_forces(case, order):
    # the gather by runs of 64 sorted atoms (one lane each) runs the per-tile gather's arithmetic: forces bit-identical;
    # the tuner keeps one of the two and the forces stay the same
    s = H.triclinic_case(n=1100) if case == "tri" else ca.build_melt(256 if case == "32k" else 2048, seed=1)
    print("case gather shapes", case, s.ff.n, "order", order, flush=True)
    op = ca.EnergyOptions(cutoff=9.0 if case == "tri" else 12.0, coulomb=True, electrostatics="pme", ewald_rtol=1e-5, pme_order=order,
                          pme_spacing=1.0, tail=False, parts=4)
    ev = ca.Evaluator(s.ff, op)
    assert ev.set_gather(0) == 0
    t0, f0 = ev.compute(s.x, s.cell)
    assert ev.set_gather(1) == 1 and ev.gather_shape() == 1
    t1, f1 = ev.compute(s.x, s.cell)
    assert np.array_equal(f1, f0)
    assert t1.coulomb == t0.coulomb
    ev.tune()
    print("tuned gather shape", ev.gather_shape(), flush=True)
    assert ev.gather_shape() in (0, 1)
    t2, f2 = ev.compute(s.x, s.cell)
    assert np.array_equal(f2, f0)
    with pytest.raises(ValueError, match="shape"):
        ev.set_gather(2)


@pytest.mark.parametrize("shape", [0, 1], ids=["gather by tiles", "gather by runs"])
def test_a_tile_past_its_bucket_capacity(shape):
    # 80 of 300 atoms inside one 4x4x4 tile of a 20^3 mesh (the tile buckets hold 32 atoms here): the atoms past the cap
    # go through the overflow list; forces and energy as the oracle's, and the same on a second evaluation (the overflow
    # count is cleared for it)
    rng = np.random.default_rng(5)
    n, L = 300, 20.0
    x = rng.random((n, 3)) * L
    x[:80] = 6.3 + rng.random((80, 3)) * 0.8
    q = rng.normal(0.0, 0.3, n)
    q -= q.mean()
    ff = ca.ForceField(type=np.zeros(n, np.int32), charge=q, mass=np.full(n, 12.0), eps=np.full((1, 1), 0.1),
                       sigma=np.full((1, 1), 3.0), excl_offset=np.zeros(n + 1, np.int32), excl_index=np.zeros(1, np.int32))
    s = ca.System(x=x, cell=np.array([0, 0, 0, L, 0, 0, 0, L, 0, 0, 0, L], float), ff=ff)
    print("case overflowing tile, gather shape", shape, flush=True)
    op = ca.EnergyOptions(cutoff=9.0, coulomb=True, electrostatics="pme", ewald_rtol=1e-5, pme_order=5, pme_spacing=1.0, tail=False,
                          parts=4)
    ev = ca.Evaluator(s.ff, op)
    ev.set_gather(shape)
    t, f = ev.compute(s.x, s.cell)
    e, fo, vo, grid = _oracle_recip(s, 9.0, 1e-5, 1.0, 5)
    err = np.sqrt(((f - fo) ** 2).sum() / (fo ** 2).sum())
    print("grid", grid, "E", t.coulomb, e, "force rel RMS", err, flush=True)
    assert t.coulomb == pytest.approx(e, rel=1e-5)
    assert err < 1e-5
    t2, f2 = ev.compute(s.x, s.cell)
    assert np.array_equal(f2, f) and t2.coulomb == t.coulomb
