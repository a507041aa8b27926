"""The C ABI (include/*.h) without a GPU: the library loads, exports every declared entry point, rejects bad arguments
with the documented codes, and its host-side system builders produce the workloads' shapes."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import caps_amd as ca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(capsgpu_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_every_declared_entry_point_is_exported():
    names = _declared()
    assert "capsgpu_compute" in names and "capsgpu_run_dynamics" in names
    L = ca.lib()
    for nm in names:
        assert hasattr(L, nm), nm


def test_null_arguments_are_einval():
    L = ca.lib()
    out = C.c_void_p()
    assert L.capsgpu_create(None, None, C.byref(out)) == -1
    assert b"null" in L.capsgpu_last_error()
    assert L.capsgpu_compute(None, None, None, None, None, 0) == -1
    assert L.capsgpu_destroy(None) == 0


def test_bad_options_are_rejected_before_the_device_is_touched():
    s = ca.kremer_grest(chains=2, beads=10)
    with pytest.raises(ValueError, match="parts"):
        ca.Evaluator(s.ff, ca.EnergyOptions(parts=9))
    with pytest.raises(ValueError, match="reciprocal"):
        ca.Evaluator(s.ff, ca.EnergyOptions(parts=4, electrostatics="dsf"))


def test_melt_builder_shapes():
    s = ca.build_melt(4, seed=3)
    ff = s.ff
    n = 4 * 128
    assert ff.n == n and s.x.shape == (n, 3)
    # per chain: CH3-(CH2)40-CH3, 42 carbons and 86 hydrogens, neutral
    q = ff.charge.reshape(4, 128)
    assert np.abs(q.sum(1)).max() < 1e-12
    m = ff.mass.reshape(4, 128)
    assert ((m == 12.011).sum(1) == 42).all() and ((m == 1.008).sum(1) == 86).all()
    # 1.0 g/cm³ in a cubic cell
    L = s.cell[3]
    assert np.allclose(s.cell[[4, 5, 6, 8, 9, 10]], 0) and s.cell[7] == L and s.cell[11] == L
    assert ff.mass.sum() / 0.602214076 / L ** 3 == pytest.approx(1.0, rel=1e-12)
    # bonds at their rest lengths; exclusions are the 1-2 and 1-3 partners, sorted
    b = ff.bond_atoms
    r = np.linalg.norm(s.x[b[:, 1]] - s.x[b[:, 0]], axis=1)
    assert np.abs(r - ff.bond_param[:, 1]).max() < 0.02
    eo, ei = ff.excl_offset, ff.excl_index
    for i in range(0, n, 37):
        part = ei[eo[i]:eo[i + 1]]
        assert (np.diff(part) > 0).all() and i not in part
    # every carbon has four neighbours, so C(4, 2) = 6 angles with it as the vertex
    assert len(ff.angle_atoms) == 4 * 42 * 6 and len(ff.bond_atoms) == 4 * 127


def test_batch_arguments_are_rejected_without_a_device():
    L = ca.lib()
    out = C.c_void_p()
    assert L.capsgpu_batch_create(None, 3, C.byref(out)) == -1
    assert b"at least one replica" in L.capsgpu_last_error()
    assert L.capsgpu_batch_create(None, 0, None) == -1
    assert L.capsgpu_batch_run(None, None, None, None, None) == -1
    assert b"null" in L.capsgpu_last_error()
    assert L.capsgpu_batch_destroy(None) == 0
    # a context named twice is refused before anything runs
    fake = (C.c_void_p * 2)(C.c_void_p(0x1000), C.c_void_p(0x1000))
    assert L.capsgpu_batch_create(C.cast(fake, C.c_void_p), 2, C.byref(out)) == -1
    assert b"twice" in L.capsgpu_last_error()
    with pytest.raises(ValueError, match="mode"):
        ca.Batch.run(object.__new__(ca.Batch), [], [], ca.BatchOptions(mode="fast"))


def test_tile_entry_points_reject_bad_arguments_without_a_device():
    L = ca.lib()
    assert L.capsgpu_tune(None, None) == -1
    assert b"null" in L.capsgpu_last_error()
    assert L.capsgpu_set_tiles(None, 8, 16, None) == -1


def test_gather_shape_entry_points_reject_bad_arguments_without_a_device():
    L = ca.lib()
    assert L.capsgpu_set_gather(None, 1, None) == -1
    assert b"null" in L.capsgpu_last_error()
    assert L.capsgpu_get_gather(None, None) == -1
    assert b"null" in L.capsgpu_last_error()
