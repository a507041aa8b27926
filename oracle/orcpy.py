"""Python bindings of the CPU oracle (liborc.so) and of the reference build (_ref/libcapsref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.
The product (caps_amd/) never imports this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_orc = None
_ref = None

f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


def _opt(a, dtype):
    """NULL for None, else a contiguous array of dtype"""
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=dtype).ctypes.data_as(C.c_void_p)


def orc():
    global _orc
    if _orc is None:
        path = os.path.join(HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle orc`")
        L = C.CDLL(path)
        L.orc_list_build.restype = C.c_void_p
        L.orc_list_build.argtypes = [f64p, C.c_int32, f64p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
        L.orc_list_free.argtypes = [C.c_void_p]
        L.orc_list_sizes.argtypes = [C.c_void_p, i64p]
        L.orc_list_get.argtypes = [C.c_void_p] + [C.c_void_p] * 16
        L.orc_list_positions.argtypes = [C.c_void_p, f64p, f32p, f32p, f32p]
        L.orc_list_prune.argtypes = [C.c_void_p, f32p, f32p, f32p, C.c_double, u16p, u16p, u16p]
        L.orc_pairs_direct.restype = C.c_int
        L.orc_pairs_direct.argtypes = [f64p, C.c_int32, f64p, i32p, C.c_int32, f64p, f64p, C.c_int32, f64p, C.c_double, C.c_int32,
                                       C.c_double, C.c_double, C.c_double, C.c_int32, C.c_double, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_double, C.c_int32, f64p, f64p, f64p]
        L.orc_bonded.restype = C.c_int
        L.orc_bonded.argtypes = [f64p, C.c_int32, f64p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                 C.c_int32, C.c_void_p, C.c_void_p, f64p, f64p, f64p]
        L.orc_tail.argtypes = [C.c_int32, f64p, f64p, f64p, C.c_int32, C.c_double, C.c_double, f64p]
        L.orc_ewald_beta.restype = C.c_double
        L.orc_ewald_beta.argtypes = [C.c_double, C.c_double]
        L.orc_good_size.restype = C.c_int32
        L.orc_good_size.argtypes = [C.c_int32]
        L.orc_pme_grid.argtypes = [f64p, C.c_double, C.c_int32, i32p]
        L.orc_bspline.argtypes = [C.c_double, C.c_int32, f64p, f64p]
        L.orc_bspline_moduli.argtypes = [C.c_int32, C.c_int32, f64p]
        L.orc_fft1d.argtypes = [C.c_int32, C.c_int32, f64p, f64p, f64p, f64p]
        L.orc_fft3d.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, f64p, f64p]
        L.orc_pme.restype = C.c_double
        L.orc_pme.argtypes = [f64p, f64p, C.c_int32, f64p, i32p, C.c_int32, C.c_double, C.c_double, f64p, f64p]
        L.orc_ewald_recip.restype = C.c_double
        L.orc_ewald_recip.argtypes = [f64p, f64p, C.c_int32, f64p, i32p, C.c_double, C.c_double, f64p, f64p]
        L.orc_pme_green.argtypes = [C.c_int32, C.c_int32, C.c_int32, i32p, f64p, C.c_int32, C.c_double, C.c_double,
                                    C.POINTER(C.c_double), f64p]
        L.orc_rng_seed.argtypes = [C.c_void_p, C.c_uint64]
        L.orc_rng_next.restype = C.c_uint64
        L.orc_rng_next.argtypes = [C.c_void_p]
        L.orc_rng_uniform01.restype = C.c_double
        L.orc_rng_uniform01.argtypes = [C.c_void_p]
        L.orc_rng_normal.restype = C.c_double
        L.orc_rng_normal.argtypes = [C.c_void_p]
        L.orc_rng_gamma.restype = C.c_double
        L.orc_rng_gamma.argtypes = [C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.orc_kinetic_energy.restype = C.c_double
        L.orc_kinetic_energy.argtypes = [f64p, f64p, C.c_int32]
        L.orc_initial_velocities.argtypes = [C.c_void_p, f64p, C.c_int32, C.c_double, f64p]
        L.orc_bussi_alpha.restype = C.c_double
        L.orc_bussi_alpha.argtypes = [C.c_double] * 6
        L.orc_kremer_grest.restype = C.c_double
        L.orc_kremer_grest.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64, f64p]
        L.orc_mlp_forward.argtypes = [f64p] * 7 + [C.c_int32, f64p]
        L.orc_mlp_backward.restype = C.c_double
        L.orc_mlp_backward.argtypes = [f64p] * 8 + [C.c_int32] + [f64p] * 6
        L.orc_descriptors.argtypes = [f64p, f64p, C.c_int32, f64p, f64p]
        L.orc_last_error.restype = C.c_char_p
        _orc = L
    return _orc


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libcapsref.so"))


def ref():
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libcapsref.so")
        if not os.path.exists(path):
            raise RuntimeError("reference build absent: run `make -C oracle ref` where the reference sources are")
        L = C.CDLL(path)
        L.ref_last_error.restype = C.c_char_p
        L.ref_ewald_beta.restype = C.c_double
        L.ref_ewald_beta.argtypes = [C.c_double, C.c_double]
        L.ref_good_size.restype = C.c_int
        L.ref_good_size.argtypes = [C.c_int]
        L.ref_pme_grid.argtypes = [f64p, C.c_double, C.c_double, C.c_int, i32p]
        L.ref_fft3d.argtypes = [f64p, f64p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ref_pme.restype = C.c_double
        L.ref_pme.argtypes = [f64p, f64p, C.c_int, f64p, i32p, C.c_int, C.c_double, C.c_int, f64p, f64p]
        L.ref_ewald_recip.restype = C.c_double
        L.ref_ewald_recip.argtypes = [f64p, f64p, C.c_int, f64p, i32p, C.c_double, f64p, f64p]
        L.ref_list_build.restype = C.c_void_p
        L.ref_list_build.argtypes = [f64p, C.c_int, f64p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        L.ref_list_free.argtypes = [C.c_void_p]
        L.ref_list_sizes.argtypes = [C.c_void_p, i64p]
        L.ref_list_get.argtypes = [C.c_void_p, u32p, u32p, i64p, u32p, u16p, u16p, u16p, u16p, f64p, f64p, i32p, i32p, f64p]
        L.ref_list_prune.restype = C.c_int64
        L.ref_list_prune.argtypes = [C.c_void_p, f64p, C.c_double, f32p, f32p, f32p, u32p, u16p, u16p, u16p, u16p, C.c_int64]
        L.ref_eval_create.restype = C.c_void_p
        L.ref_eval_create.argtypes = [f64p, C.c_int, f64p, i32p, C.c_int, f64p, f64p, f64p, C.c_int, C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_double, C.c_int, C.c_void_p, C.c_int]
        L.ref_eval_free.argtypes = [C.c_void_p]
        L.ref_eval_compute.restype = C.c_int
        L.ref_eval_compute.argtypes = [C.c_void_p, f64p, C.c_int, C.c_int, f64p, f64p, f64p]
        L.ref_eval_threads.restype = C.c_int
        L.ref_eval_threads.argtypes = [C.c_void_p]
        L.ref_rng_normal.argtypes = [C.c_uint64, C.c_int, f64p]
        L.ref_rng_bussi.argtypes = [C.c_uint64, C.c_double, C.c_int, f64p]
        L.ref_rng_uniform.argtypes = [C.c_uint64, C.c_double, C.c_double, C.c_int, f64p]
        _ref = L
    return _ref


# ---- convenience wrappers -----------------------------------------------------------------------------------------
LIST_FIELDS = ("perm", "iperm", "first", "cj", "code", "mask", "ex", "ex14", "org", "off", "ioff", "qint", "qvec")


def orc_list(x, cell, rlist, excl_off, excl_idx, p14_off=None, p14_idx=None, keep_excluded=True):
    """the oracle's list as a dict of arrays"""
    L = orc()
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    n = x.size // 3
    h = L.orc_list_build(x, n, np.ascontiguousarray(cell, np.float64), rlist, _opt(excl_off, np.int32), _opt(excl_idx, np.int32),
                         _opt(p14_off, np.int32), _opt(p14_idx, np.int32), int(keep_excluded))
    if not h:
        raise RuntimeError(L.orc_last_error().decode())
    s = np.zeros(7, np.int64)
    L.orc_list_sizes(h, s)
    n, ncl, ns, ne, nc = (int(v) for v in s[:5])
    out = dict(perm=np.zeros(ns, np.uint32), iperm=np.zeros(n, np.uint32), first=np.zeros(ncl + 1, np.int64),
               cj=np.zeros(ne, np.uint32), code=np.zeros(ne, np.uint16), mask=np.zeros(ne, np.uint16), ex=np.zeros(ne, np.uint16),
               ex14=np.zeros(ne, np.uint16), org=np.zeros(3 * ncl), off=np.zeros(3 * ns), ioff=np.zeros(3 * ns, np.int32),
               qint=np.zeros(3 * nc, np.int32), qvec=np.zeros(3 * nc), bx=np.zeros(ns, np.float32), by=np.zeros(ns, np.float32),
               bz=np.zeros(ns, np.float32))
    keys = ("perm", "iperm", "first", "cj", "code", "mask", "ex", "ex14", "org", "off", "ioff", "qint", "qvec", "bx", "by", "bz")
    L.orc_list_get(h, *[out[k].ctypes.data_as(C.c_void_p) for k in keys])
    out["nu"], out["nv"] = int(s[5]), int(s[6])
    out["_h"] = h
    return out


def orc_list_free(lst):
    orc().orc_list_free(lst.pop("_h"))


def orc_positions_prune(lst, x, radius):
    L = orc()
    ns = lst["perm"].size
    X, Y, Z = (np.zeros(ns, np.float32) for _ in range(3))
    L.orc_list_positions(lst["_h"], np.ascontiguousarray(x, np.float64).reshape(-1), X, Y, Z)
    ne = lst["cj"].size
    pm, pe, p14 = (np.zeros(ne, np.uint16) for _ in range(3))
    L.orc_list_prune(lst["_h"], X, Y, Z, radius, pm, pe, p14)
    return X, Y, Z, pm, pe, p14


def ref_list(x, cell, rlist, excl_off, excl_idx, p14_off=None, p14_idx=None, keep_excluded=True, threads=1):
    R = ref()
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    n = x.size // 3
    h = R.ref_list_build(x, n, np.ascontiguousarray(cell, np.float64), rlist, _opt(excl_off, np.int32), _opt(excl_idx, np.int32),
                         _opt(p14_off, np.int32), _opt(p14_idx, np.int32), int(keep_excluded), threads)
    if not h:
        raise RuntimeError(R.ref_last_error().decode())
    s = np.zeros(5, np.int64)
    R.ref_list_sizes(h, s)
    n, ncl, ns, ne, nc = (int(v) for v in s)
    out = dict(perm=np.zeros(ns, np.uint32), iperm=np.zeros(n, np.uint32), first=np.zeros(ncl + 1, np.int64),
               cj=np.zeros(ne, np.uint32), code=np.zeros(ne, np.uint16), mask=np.zeros(ne, np.uint16), ex=np.zeros(ne, np.uint16),
               ex14=np.zeros(ne, np.uint16), org=np.zeros(3 * ncl), off=np.zeros(3 * ns), ioff=np.zeros(3 * ns, np.int32),
               qint=np.zeros(3 * nc, np.int32), qvec=np.zeros(3 * nc))
    R.ref_list_get(h, *[out[k] for k in LIST_FIELDS])
    out["_h"] = h
    return out


def ref_prune(lst, x, radius):
    R = ref()
    ns = lst["perm"].size
    X, Y, Z = (np.zeros(ns, np.float32) for _ in range(3))
    cap = lst["cj"].size
    cj = np.zeros(cap, np.uint32)
    code, mask, ex, ex14 = (np.zeros(cap, np.uint16) for _ in range(4))
    k = R.ref_list_prune(lst["_h"], np.ascontiguousarray(x, np.float64).reshape(-1), radius, X, Y, Z, cj, code, mask, ex, ex14, cap)
    return X, Y, Z, dict(cj=cj[:k], code=code[:k], mask=mask[:k], ex=ex[:k], ex14=ex14[:k])


def ref_list_free(lst):
    ref().ref_list_free(lst.pop("_h"))


class Rng:
    """the oracle's std::mt19937_64 with caps::Normal / caps::Gamma"""

    class _S(C.Structure):
        _fields_ = [("mt", C.c_uint64 * 312), ("mti", C.c_int32), ("have", C.c_int32), ("spare", C.c_double)]

    def __init__(self, seed):
        self.s = Rng._S()
        orc().orc_rng_seed(C.byref(self.s), seed)

    def next(self):
        return orc().orc_rng_next(C.byref(self.s))

    def uniform01(self):
        return orc().orc_rng_uniform01(C.byref(self.s))

    def normal(self):
        return orc().orc_rng_normal(C.byref(self.s))

    def gamma(self, shape, scale):
        h, sp = C.c_int32(0), C.c_double(0.0)
        return orc().orc_rng_gamma(C.byref(self.s), shape, scale, C.byref(h), C.byref(sp))

    def ptr(self):
        return C.byref(self.s)
