"""The oracle's cluster pair list (oracle/orc_list.c) pinned against the reference build (oracle/_ref, compiled from
core/src/nbnxm.cpp) and against a brute-force pair count, as tests/nbnxm_check.cpp:78-119 checks the reference."""
import numpy as np
import pytest

import helpers as H
import orcpy as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="the reference build oracle/_ref is absent (no reference sources)")


def _cases():
    tri = H.triclinic_case(n=1100)
    melt = H.melt_case(chains=16)
    return [("triclinic", tri, 10.0), ("melt", melt, 14.0)]


@needs_ref
@pytest.mark.parametrize("name", ["triclinic", "melt"])
def test_list_is_bit_exact_against_the_reference(name):
    s, rl = {k: (v, r) for k, v, r in _cases()}[name]
    ff = s.ff
    o = O.orc_list(s.x, s.cell, rl, ff.excl_offset, ff.excl_index, ff.p14_offset, ff.p14_index, True)
    r = O.ref_list(s.x, s.cell, rl, ff.excl_offset, ff.excl_index, ff.p14_offset, ff.p14_index, True)
    try:
        for k in O.LIST_FIELDS:
            assert o[k].shape == r[k].shape, k
            assert np.array_equal(o[k], r[k]), k
    finally:
        O.orc_list_free(o)
        O.ref_list_free(r)


@needs_ref
def test_prune_is_bit_exact_against_the_reference():
    s = H.triclinic_case(n=1100)
    ff = s.ff
    rng = np.random.default_rng(3)
    x1 = s.x + rng.normal(0, 0.15, s.x.shape)   # moved atoms: the prune sees other distances than the build
    o = O.orc_list(s.x, s.cell, 10.0, ff.excl_offset, ff.excl_index, ff.p14_offset, ff.p14_index, True)
    r = O.ref_list(s.x, s.cell, 10.0, ff.excl_offset, ff.excl_index, ff.p14_offset, ff.p14_index, True)
    try:
        X, Y, Z, pm, pe, p14 = O.orc_positions_prune(o, x1, 8.8)
        Xr, Yr, Zr, kept = O.ref_prune(r, x1, 8.8)
        assert np.array_equal(X, Xr) and np.array_equal(Y, Yr) and np.array_equal(Z, Zr)
        keep = (pm | pe) != 0
        assert np.array_equal(o["cj"][keep], kept["cj"])
        assert np.array_equal(o["code"][keep], kept["code"])
        assert np.array_equal(pm[keep], kept["mask"])
        assert np.array_equal(pe[keep], kept["ex"])
        assert np.array_equal(p14[keep], kept["ex14"])
    finally:
        O.orc_list_free(o)
        O.ref_list_free(r)


def _listed_pairs(o, x, cell, rc, which="mask"):
    """(i, j, n1, n2, n3) keys of the listed lanes with r < rc: atoms and lattice vector n with d = x_j + n·cell − x_i"""
    perm, ioff, qint = o["perm"], o["ioff"].reshape(-1, 3), o["qint"].reshape(-1, 3)
    A = cell[3:].reshape(3, 3)
    keys = []
    ncl = o["first"].size - 1
    ci = np.repeat(np.arange(ncl), np.diff(o["first"]))
    bits = o[which].astype(np.int64)
    for a in range(4):
        for l in range(4):
            sel = (bits >> (4 * a + l)) & 1 == 1
            si, sj = 4 * ci[sel] + a, 4 * o["cj"][sel].astype(np.int64) + l
            ai, aj = perm[si].astype(np.int64), perm[sj].astype(np.int64)
            n = ioff[sj] - ioff[si] + qint[o["code"][sel]]
            d = x[aj] + n @ A - x[ai]
            r2 = (d * d).sum(1)
            ok = r2 < rc * rc
            keys.append(np.column_stack([ai[ok], aj[ok], n[ok]]))
    k = np.concatenate(keys)
    return _canonical(k)


def _canonical(k):
    """an unordered pair with its image: (i, j, n) and (j, i, −n) are the same; for i == j keep the larger of ±n"""
    k = k.copy()
    sw = k[:, 0] > k[:, 1]
    k[sw] = np.column_stack([k[sw, 1], k[sw, 0], -k[sw, 2:]])
    same = k[:, 0] == k[:, 1]
    neg = same & ((k[:, 2] < 0) | ((k[:, 2] == 0) & ((k[:, 3] < 0) | ((k[:, 3] == 0) & (k[:, 4] < 0)))))
    k[neg, 2:] = -k[neg, 2:]
    return k


def _brute_pairs(x, cell, rc, excl_off, excl_idx):
    """every unordered pair (and self image) within rc over periodic images; excluded pairs split off at the nearest
    image (field.cpp's nearest-image rule, d² ≤ d_min² + 1e-6)"""
    A = cell[3:].reshape(3, 3)
    Ainv = np.linalg.inv(A)
    n = len(x)
    w = 1.0 / np.linalg.norm(Ainv, axis=0)   # distances between opposite faces
    R = np.ceil(rc / w).astype(int) + 1
    imgs = np.array([(a, b, c) for a in range(-R[0], R[0] + 1) for b in range(-R[1], R[1] + 1) for c in range(-R[2], R[2] + 1)])
    excl = set()
    for i in range(n):
        for j in excl_idx[excl_off[i]:excl_off[i + 1]]:
            excl.add((min(i, j), max(i, j)))
    keep, exk = [], []
    for i in range(n):
        j = np.arange(i, n)
        d0 = x[j] - x[i]
        for img in imgs:
            d = d0 + img @ A
            r2 = (d * d).sum(1)
            sel = r2 < rc * rc
            if not sel.any():
                continue
            for jj, rr in zip(j[sel], r2[sel]):
                if jj == i and (img == 0).all():
                    continue
                key = (i, int(jj), *img)
                if (i, int(jj)) in excl:
                    f = (x[jj] - x[i]) @ Ainv
                    m = f - np.round(f)
                    dm = m @ A
                    if rr <= dm @ dm + 1e-6:
                        exk.append(key)
                        continue
                keep.append(key)
    return _canonical(np.array(keep)), _canonical(np.array(exk)) if exk else np.zeros((0, 5), int)


def _rows(k):
    return set(map(tuple, k.tolist()))


@pytest.mark.parametrize("name", ["triclinic", "ortho"])
def test_list_holds_every_pair_within_the_cutoff_once(name):
    if name == "triclinic":
        s = H.triclinic_case(n=300)
    else:
        s = H.triclinic_case(n=300)
        s.cell = np.array([0, 0, 0, 21.0, 0, 0, 0, 20.0, 0, 0, 0, 22.0])
        s.x = np.mod(s.x, [21.0, 20.0, 22.0])
    ff = s.ff
    rc, rl = 6.0, 8.0
    o = O.orc_list(s.x, s.cell, rl, ff.excl_offset, ff.excl_index, None, None, True)
    try:
        lm = _listed_pairs(o, s.x, s.cell, rc, "mask")
        le = _listed_pairs(o, s.x, s.cell, rc, "ex")
    finally:
        O.orc_list_free(o)
    bk, be = _brute_pairs(s.x, s.cell, rc, ff.excl_offset, ff.excl_index)
    assert len(_rows(lm)) == len(lm), "a pair is listed twice"
    assert _rows(lm) == _rows(bk)
    assert _rows(le) == _rows(be)
