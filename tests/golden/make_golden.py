"""Writes the golden vectors under tests/golden/ from the reference's own code compiled into oracle/_ref (make -C oracle
ref, which compiles the sources where they lie under /root/reference). Inputs and the reference's outputs only.

    python tests/golden/make_golden.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(HERE))

import helpers as H  # noqa: E402
import orcpy as O  # noqa: E402


def main():
    assert O.ref_available(), "build oracle/_ref first (make -C oracle ref)"
    s = H.triclinic_case(n=300)
    ff = s.ff
    common = dict(x=s.x, cell=s.cell, type=ff.type, charge=ff.charge, eps=ff.eps, sigma=ff.sigma, excl_offset=ff.excl_offset,
                  excl_index=ff.excl_index, p14_offset=ff.p14_offset, p14_index=ff.p14_index)
    # the cluster list at r_list = 10 Å and its prune at 8.8 Å after a displacement
    r = O.ref_list(s.x, s.cell, 10.0, ff.excl_offset, ff.excl_index, ff.p14_offset, ff.p14_index, True)
    x1 = s.x + np.random.default_rng(3).normal(0, 0.15, s.x.shape)
    X, Y, Z, kept = O.ref_prune(r, x1, 8.8)
    lst = {k: r[k] for k in O.LIST_FIELDS}
    O.ref_list_free(r)
    np.savez_compressed(os.path.join(HERE, "list_triclinic300.npz"), rlist=10.0, prune_radius=8.8, x_moved=x1, X=X, Y=Y, Z=Z,
                        **{"p_" + k: v for k, v in kept.items()}, **lst, **common)
    # the cluster kernel (12-6, Ewald real space κ = 0.32, r_c 8 Å, skin 2 Å, excluded-pair corrections in the kernel)
    (ev, ec), f, w = H.ref_pairs(s, 8.0, 2.0, 1, 0.32)
    np.savez_compressed(os.path.join(HERE, "kernel_triclinic300.npz"), rc=8.0, skin=2.0, kappa=0.32, evdw=ev, ecoul=ec, f=f, w=w, **common)
    # the Fortran PME (order 5, 1 Å spacing, β from r_c 9 Å and 1e-5)
    beta = O.ref().ref_ewald_beta(9.0, 1e-5)
    grid = np.zeros(3, np.int32)
    O.ref().ref_pme_grid(np.ascontiguousarray(s.cell, np.float64), beta, 1.0, 5, grid)
    fp, vp = np.zeros(s.x.size), np.zeros(6)
    ep = O.ref().ref_pme(s.x.reshape(-1).copy(), np.ascontiguousarray(ff.charge), ff.n, np.ascontiguousarray(s.cell, np.float64), grid, 5,
                         beta, 1, fp, vp)
    np.savez_compressed(os.path.join(HERE, "pme_triclinic300.npz"), beta=beta, grid=grid, order=5, energy=ep, f=fp.reshape(-1, 3), vir=vp,
                        **common)
    # the random streams of core/include/caps/rng.hpp
    nrm = np.zeros(2000)
    O.ref().ref_rng_normal(123, 2000, nrm)
    bus = np.zeros(2 * 100)
    O.ref().ref_rng_bussi(77, 3.0 * 512 - 3, 100, bus)
    np.savez_compressed(os.path.join(HERE, "rng.npz"), normal_seed=123, normal=nrm, bussi_seed=77, bussi_ndof=3.0 * 512 - 3, bussi=bus)
    # the Kremer-Grest builder's bead positions (kremer_grest.cpp:15-83): flexible and with bending stiffness
    kg = {}
    for tag, args in (("flexible", (20, 50, 0.85, 0.0, 0.97, 1)), ("stiff", (10, 30, 0.85, 1.5, 0.97, 3))):
        raw = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "kg_golden")] + [str(a) for a in args], capture_output=True,
                             check=True).stdout
        kg[tag + "_args"] = np.array(args, np.float64)
        kg[tag + "_x"] = np.frombuffer(raw, np.float64).reshape(-1, 3)
    np.savez_compressed(os.path.join(HERE, "kg.npz"), **kg)
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
