"""LAMMPS data files (atom_style full, units real) through the C ABI, host code only: read_lammps_data /
write_lammps_data (core/src/io_lammps.cpp:58-297, :505-583) with this path's force field in the coefficient sections.
A system written and read back is the same system to the last bit; a hand-written file in LAMMPS's own layout reads to
the numbers it states; a general cell is rotated into LAMMPS's frame without changing any distance; the checkpoint
helpers follow caps md's --checkpoint-every / --resume (cli/main.cpp:7809-7846)."""
import os

import numpy as np
import pytest

import caps_amd as ca


def _same(a, b):
    fa, fb = a.ff, b.ff
    assert np.array_equal(a.x, b.x) and np.array_equal(a.cell, b.cell)
    for k in ("type", "charge", "mass", "eps", "sigma", "excl_offset", "excl_index", "bond_atoms", "bond_param", "angle_atoms",
              "angle_param", "fene_atoms", "fene_param"):
        u, w = getattr(fa, k), getattr(fb, k)
        assert u.shape == w.shape and np.array_equal(u, w), k


def test_melt_round_trip_is_exact(tmp_path):
    s = ca.build_melt(4, seed=3, density=0.9)
    # atoms outside the box and across it: the unwrapped chains of a melt, pushed further
    s.x[::7] += s.cell[3:6] * 2.0 - s.cell[6:9]
    s.v = np.random.default_rng(1).normal(0, 0.01, s.x.shape)
    p = str(tmp_path / "melt.data")
    ca.write_lammps(p, s, "melt of 4 chains")
    assert not os.path.exists(p + ".tmp")
    r = ca.read_lammps(p)
    _same(s, r)
    assert np.array_equal(r.v, s.v)
    assert r.title.startswith("CAPS-AMD") and "melt of 4 chains" in r.title and "atom_style full" in r.title
    text = open(p).read()
    assert "Pair Coeffs  # lj/cut" in text and "Bond Coeffs  # harmonic" in text and "Angle Coeffs  # harmonic" in text
    # most atoms in the box carry zero image flags; the ones pushed out carry them
    rows = text.split("Atoms  # full")[1].split("Velocities")[0].split("\n")
    flags = [tuple(int(q) for q in r.split()[7:10]) for r in rows if len(r.split()) == 10]
    assert len(flags) == s.ff.n and any(f != (0, 0, 0) for f in flags)


def test_kremer_grest_round_trip_keeps_fene_and_1_2_exclusions(tmp_path):
    s = ca.kremer_grest(chains=5, beads=20, density=0.85, seed=2)
    assert len(s.ff.fene_atoms) > 0 and len(s.ff.angle_atoms) == 0
    p = str(tmp_path / "kg.data")
    ca.write_lammps(p, s)
    assert "Bond Coeffs  # fene" in open(p).read()
    r = ca.read_lammps(p)
    _same(s, r)
    assert r.v is None
    # with bending terms added, 1-3 partners are excluded too unless asked not to be
    f = s.ff
    tri = np.array([[c * 20 + k, c * 20 + k + 1, c * 20 + k + 2] for c in range(5) for k in range(18)], np.int32)
    f.angle_atoms, f.angle_param = tri, np.tile([1.5, np.pi], (len(tri), 1))
    ca.write_lammps(p, s)
    r12, r13 = ca.read_lammps(p, exclude13=False), ca.read_lammps(p)
    assert np.array_equal(r12.ff.excl_index, f.excl_index) and np.array_equal(r12.ff.angle_param, f.angle_param)
    assert r13.ff.excl_index.size == f.excl_index.size + 2 * len(tri)


def test_reads_a_lammps_file_as_written_by_hand(tmp_path):
    p = tmp_path / "hand.data"
    p.write_text("""LAMMPS data file via write_data, a hand-made one

4 atoms
2 atom types
2 bonds
1 bond types
1 angles
1 angle types

0.0 20.0 xlo xhi
-1.0 19.0 ylo yhi
0.0 30.0 zlo zhi
2.0 0.0 0.0 xy xz yz

Masses

1 12.011
2 1.008  # H

Pair Coeffs # lj/cut

1 0.1 3.5
2 0.03 2.5

Bond Coeffs # harmonic

1 340.0 1.09

Angle Coeffs # harmonic

1 35.0 109.5

Atoms # full

10 1 1 -0.24 1.0 2.0 3.0 0 0 0
11 1 2 0.06 2.0 2.0 3.0 1 0 -1
12 1 2 0.06 1.0 3.0 3.0
13 2 2 0.12 5.0 5.0 5.0 0 0 0

Velocities

10 0.001 0.002 0.003
13 -0.1 0 0

Bonds

1 1 10 11
2 1 10 12

Angles

1 1 11 10 12
""")
    s = ca.read_lammps(str(p))
    f = s.ff
    assert f.n == 4 and f.ntypes == 2 and list(f.type) == [0, 1, 1, 1]
    assert np.allclose(s.cell, [0, -1, 0, 20, 0, 0, 2, 20, 0, 0, 0, 30])
    # image flags (1, 0, -1): + a - c
    assert np.allclose(s.x[1], [2.0 + 20.0, 2.0, 3.0 - 30.0])
    assert np.allclose(f.eps, [[0.1, np.sqrt(0.1 * 0.03)], [np.sqrt(0.1 * 0.03), 0.03]])
    assert np.allclose(f.sigma[0, 1], np.sqrt(3.5 * 2.5))
    assert np.array_equal(f.bond_atoms, [[0, 1], [0, 2]]) and np.allclose(f.bond_param, [[340.0, 1.09]] * 2)
    assert np.array_equal(f.angle_atoms, [[1, 0, 2]]) and np.isclose(f.angle_param[0, 1], np.radians(109.5))
    # 1-2 and 1-3 partners: atom 0 with 1, 2; atoms 1 and 2 with each other
    ex = [list(f.excl_index[f.excl_offset[i]:f.excl_offset[i + 1]]) for i in range(4)]
    assert ex == [[1, 2], [0, 2], [0, 1], []]
    assert np.allclose(s.v, [[0.001, 0.002, 0.003], [0, 0, 0], [0, 0, 0], [-0.1, 0, 0]])
    assert np.allclose(f.mass, [12.011, 1.008, 1.008, 1.008]) and np.isclose(f.charge.sum(), 0.0)


def test_errors_say_what_is_wrong(tmp_path):
    with pytest.raises(ValueError, match="cannot open"):
        ca.read_lammps(str(tmp_path / "missing.data"))
    p = tmp_path / "bad.data"
    p.write_text("t\n\n1 atoms\n1 atom types\n0 1 xlo xhi\n0 1 ylo yhi\n0 1 zlo zhi\n\nAtoms\n\n1 1 1 0.0 0.5 0.5\n")
    with pytest.raises(ValueError, match="7 or 10 columns"):
        ca.read_lammps(str(p))
    p.write_text("t\n\n1 atoms\n1 atom types\n0 1 xlo xhi\n0 1 ylo yhi\n0 1 zlo zhi\n\nAtoms\n\n1 1 1 0.0 0.5 0.5 0.5\n")
    with pytest.raises(ValueError, match="no mass"):
        ca.read_lammps(str(p))
    p.write_text("t\n\n2 atoms\n1 atom types\n0 1 xlo xhi\n0 1 ylo yhi\n0 1 zlo zhi\n\nMasses\n\n1 1.0\n\nAtoms\n\n1 1 1 0.0 0.5 0.5 0.5\n")
    with pytest.raises(ValueError, match="header says 2 atoms"):
        ca.read_lammps(str(p))


def test_a_general_cell_is_rotated_into_the_lammps_frame(tmp_path):
    s = ca.build_melt(2, seed=5, density=0.8)
    th = 0.4
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]]) @ \
        np.array([[1, 0, 0], [0, np.cos(0.3), -np.sin(0.3)], [0, np.sin(0.3), np.cos(0.3)]])
    o = s.cell[:3]
    s.x = (s.x - o) @ R.T + o
    s.cell = np.concatenate([o, (s.cell[3:].reshape(3, 3) @ R.T).reshape(-1)])
    p = str(tmp_path / "rot.data")
    ca.write_lammps(p, s, velocities=False)
    r = ca.read_lammps(p)
    c = r.cell[3:].reshape(3, 3)
    assert c[0, 1] == 0 and c[0, 2] == 0 and c[1, 2] == 0 and np.all(np.diag(c) > 0)
    # distances and the volume unchanged
    d0 = np.linalg.norm(s.x[1:] - s.x[:-1], axis=1)
    d1 = np.linalg.norm(r.x[1:] - r.x[:-1], axis=1)
    assert np.allclose(d0, d1, rtol=0, atol=1e-9)
    assert np.isclose(abs(np.linalg.det(c)), abs(np.linalg.det(s.cell[3:].reshape(3, 3))), rtol=1e-12)


def test_checkpoint_files_and_resume_options(tmp_path):
    s = ca.kremer_grest(chains=3, beads=10, seed=4)
    s.v = np.zeros_like(s.x)
    p = str(tmp_path / "run.restart.data")
    ca.write_checkpoint(p, s, 400, 1000)
    assert os.path.exists(str(tmp_path / "run.restart.settings.json"))
    r = ca.read_lammps(p)
    assert ca.checkpoint_step(r) == 400
    # the run settings come back from the file beside it
    assert r.ff.lj_shift and r.ff.cutoff == 2.0 ** (1.0 / 6.0) and r.ff.coul14 == 1.0
    o = ca.resume_options(r, ca.DynamicsOptions(steps=1000, new_velocities=True), 1000)
    assert (o.step_offset, o.steps, o.new_velocities) == (400, 600, False)
    with pytest.raises(ValueError, match="must be larger"):
        ca.resume_options(r, ca.DynamicsOptions(), 400)
    ca.write_lammps(p, s, "not a checkpoint")
    with pytest.raises(ValueError, match="not a checkpoint"):
        ca.checkpoint_step(ca.read_lammps(p))
