"""The allocation-free evaluation (capsgpu_compute_noalloc, no reference counterpart) through the C ABI: forces and
energy terms bit-identical to capsgpu_compute's on two calls in a row, the second on a smaller system (the shared
scratch block keeps the first call's contents), and a clear error for a system beyond the block."""
import ctypes as C

import numpy as np
import pytest

import caps_amd as ca
import helpers as H

pytestmark = pytest.mark.gpu

OPTS = ca.EnergyOptions(cutoff=10.0, coulomb=True, electrostatics="pme", ewald_rtol=1e-5, pme_order=5, pme_spacing=1.0, tail=False,
                        parts=5)


def _both(s, opts=OPTS):
    L = ca.lib()
    ev_a, ev_b = ca.Evaluator(s.ff, opts), ca.Evaluator(s.ff, opts)
    x = np.ascontiguousarray(s.x, np.float64).reshape(-1)
    cell = np.ascontiguousarray(s.cell, np.float64)
    out = []
    for ev, fn in ((ev_a, L.capsgpu_compute), (ev_b, L.capsgpu_compute_noalloc)):
        f, t = np.zeros_like(x), np.zeros(ca.NTERMS)
        ca._check(fn(ev._h, ca._p(x), ca._p(cell), ca._p(f), ca._p(t), 1))
        out.append((f, t))
    ev_a.close(), ev_b.close()
    return out


@pytest.mark.parametrize("chains", [64, 12])
def test_noalloc_matches_compute(chains):
    # 8,192 then 1,536 atoms: the second call runs on the block the first one filled
    print("case noalloc melt %d chains" % chains, flush=True)
    s = H.melt_case(chains=chains, seed=5)
    (fa, ta), (fb, tb) = _both(s)
    assert np.array_equal(fa, fb), np.abs(fa - fb).max()
    assert np.array_equal(ta, tb), (ta, tb)
    assert np.abs(fa).max() > 0


def test_noalloc_twice_on_one_context_then_compute_again():
    print("case noalloc repeated on one context", flush=True)
    s = H.melt_case(chains=10, seed=6)
    L = ca.lib()
    ev = ca.Evaluator(s.ff, OPTS)
    x = np.ascontiguousarray(s.x, np.float64).reshape(-1)
    cell = np.ascontiguousarray(s.cell, np.float64)
    res = []
    for fn in (L.capsgpu_compute, L.capsgpu_compute_noalloc, L.capsgpu_compute_noalloc, L.capsgpu_compute):
        f, t = np.zeros_like(x), np.zeros(ca.NTERMS)
        ca._check(fn(ev._h, ca._p(x), ca._p(cell), ca._p(f), ca._p(t), 1))
        res.append((f, t))
    ev.close()
    for f, t in res[1:]:
        assert np.array_equal(f, res[0][0]) and np.array_equal(t, res[0][1])


def test_noalloc_refuses_the_pair_path_and_oversized_systems():
    print("case noalloc refusals", flush=True)
    L = ca.lib()
    s = H.melt_case(chains=2)
    ev = ca.Evaluator(s.ff, ca.EnergyOptions(cutoff=8.0, coulomb=False, parts=3, tail=False))
    x = np.ascontiguousarray(s.x, np.float64).reshape(-1)
    f, t = np.zeros_like(x), np.zeros(ca.NTERMS)
    rc = L.capsgpu_compute_noalloc(ev._h, ca._p(x), ca._p(np.ascontiguousarray(s.cell)), ca._p(f), ca._p(t), 1)
    assert rc == -1 and b"parts" in L.capsgpu_last_error()
    ev.close()
    # a 512³ mesh (0.5 GB of mesh alone) is beyond the block: refused before any work
    big = np.array([0, 0, 0, 510.0, 0, 0, 0, 510.0, 0, 0, 0, 510.0])
    ev = ca.Evaluator(s.ff, OPTS)
    rc = L.capsgpu_compute_noalloc(ev._h, ca._p(x), ca._p(big), ca._p(f), ca._p(t), 1)
    assert rc == -1 and b"MiB" in L.capsgpu_last_error()
    # the context still evaluates through the allocating entry point
    ca._check(L.capsgpu_compute(ev._h, ca._p(x), ca._p(np.ascontiguousarray(s.cell)), ca._p(f), ca._p(t), 1))
    ev.close()
