"""Shared test inputs: small systems in the shapes the reference's own tests use, and the oracle calls on them."""
import numpy as np

import caps_amd as ca
import orcpy as O

KC = 332.06371


def csr(parts):
    off = np.zeros(len(parts) + 1, np.int32)
    idx = []
    for i, p in enumerate(parts):
        off[i + 1] = off[i] + len(p)
        idx += sorted(p)
    return off, np.array(idx if idx else [0], np.int32)


def triclinic_case(n=1100, seed=7, nt=3):
    """the shape of the reference's tests/nbnxm_check.cpp:33-58: charged LJ particles on a jittered lattice in a
    triclinic cell, pairs (2i, 2i+1) excluded, every other one a 1-4 pair"""
    rng = np.random.default_rng(seed)
    cell = np.array([0, 0, 0, 24, 0, 0, 3, 22, 0, -2, 4, 23], float)
    m = 11
    i = np.arange(n)
    f = np.stack([(i % m + 0.3 * rng.random(n)) / m, (i // m % m + 0.3 * rng.random(n)) / m, (i // (m * m) + 0.3 * rng.random(n)) / 10.0], 1)
    x = f @ cell[3:].reshape(3, 3) + cell[:3]
    ex = [[] for _ in range(n)]
    for a in range(0, n - 1, 2):
        ex[a].append(a + 1), ex[a + 1].append(a)
    p14 = [[] for _ in range(n)]
    for a in range(0, n - 1, 4):
        p14[a].append(a + 1), p14[a + 1].append(a)
    eo, ei = csr(ex)
    po, pi = csr(p14)
    types = (i % nt).astype(np.int32)
    q = np.where(i % 2 == 1, 0.35, -0.35) * (1 + 0.1 * types)
    eps1 = np.array([0.07, 0.03, 0.11])[:nt]
    sig1 = np.array([3.5, 2.5, 3.2])[:nt]
    eps = np.sqrt(np.outer(eps1, eps1))
    sig = np.sqrt(np.outer(sig1, sig1))
    ff = ca.ForceField(type=types, charge=q, mass=np.full(n, 12.0), eps=eps, sigma=sig, excl_offset=eo, excl_index=ei,
                       p14_offset=po, p14_index=pi, coul14=0.5)
    return ca.System(x=x, cell=cell, ff=ff)


def lj_tables(ff):
    """c12, c6 per type pair (12-6: 4εσ¹², 4εσ⁶; 9-6: 2εσ⁹, 3εσ⁶) as Evaluator::cluster_coefficients makes them"""
    s6 = ff.sigma ** 6
    if ff.lj96:
        return 2 * ff.eps * s6 * np.sqrt(s6), 3 * ff.eps * s6
    return 4 * ff.eps * s6 * s6, 4 * ff.eps * s6


def direct_pairs(system, rc, coul, kappa=0.0, shift=False, with_excl=True, dsf=(0.0, 0.0)):
    """the oracle's FP64 direct pair sum: (vdw, coulomb), forces (N, 3), w[6]"""
    ff = system.ff
    c12, c6 = lj_tables(ff)
    x = np.ascontiguousarray(system.x, np.float64).reshape(-1)
    e, f, w = np.zeros(2), np.zeros(x.size), np.zeros(6)
    O.orc().orc_pairs_direct(x, ff.n, np.ascontiguousarray(system.cell, np.float64), np.ascontiguousarray(ff.type, np.int32), ff.ntypes,
                             np.ascontiguousarray(c12.reshape(-1)), np.ascontiguousarray(c6.reshape(-1)), 2 if ff.lj96 else 1,
                             np.ascontiguousarray(ff.charge, np.float64), KC, coul, kappa, dsf[0], dsf[1], int(shift), rc,
                             O._opt(ff.excl_offset, np.int32), O._opt(ff.excl_index, np.int32), O._opt(ff.p14_offset, np.int32),
                             O._opt(ff.p14_index, np.int32), ff.coul14, int(with_excl), e, f, w)
    return e, f.reshape(-1, 3), w


def oracle_bonded(system):
    ff = system.ff
    x = np.ascontiguousarray(system.x, np.float64).reshape(-1)
    e, f, w = np.zeros(3), np.zeros(x.size), np.zeros(6)
    rc = O.orc().orc_bonded(x, ff.n, np.ascontiguousarray(system.cell, np.float64), len(ff.bond_atoms), O._opt(ff.bond_atoms, np.int32),
                            O._opt(ff.bond_param, np.float64), len(ff.angle_atoms), O._opt(ff.angle_atoms, np.int32),
                            O._opt(ff.angle_param, np.float64), len(ff.fene_atoms), O._opt(ff.fene_atoms, np.int32),
                            O._opt(ff.fene_param, np.float64), e, f, w)
    if rc != 0:
        raise ca.FieldError("FENE bond past R0")
    return e, f.reshape(-1, 3), w


def oracle_pme(system, beta, grid, order):
    ff = system.ff
    x = np.ascontiguousarray(system.x, np.float64).reshape(-1)
    f, vir = np.zeros(x.size), np.zeros(6)
    e = O.orc().orc_pme(x, np.ascontiguousarray(ff.charge, np.float64), ff.n, np.ascontiguousarray(system.cell, np.float64),
                        np.ascontiguousarray(grid, np.int32), order, beta, KC, f, vir)
    return e, f.reshape(-1, 3), vir


def oracle_total(system, opts):
    """the full evaluator in FP64: pairs (direct sum) + bonded + PME + self/background + tail, as EnergyTerms-like dict"""
    ff = system.ff
    rc = ff.cutoff if ff.cutoff > 0 else opts.cutoff
    pme = opts.coulomb and opts.electrostatics == "pme"
    beta = O.orc().orc_ewald_beta(rc, opts.ewald_rtol) if pme else 0.0
    tail = opts.tail and not ff.lj_shift
    coul = (1 if pme else 2) if opts.coulomb else 0
    if coul == 2:
        a = opts.dsf_alpha
        P, A = 0.3275911, (0.254829592, -0.284496736, 1.421413741, -1.453152027, 1.061405429)
        t = 1 / (1 + P * a * rc)
        erfc_rc = t * (A[0] + t * (A[1] + t * (A[2] + t * (A[3] + t * A[4])))) * np.exp(-(a * rc) ** 2)
        dsf = (erfc_rc / rc, erfc_rc / rc ** 2 + 2 * a / np.sqrt(np.pi) * np.exp(-a * a * rc * rc) / rc)
        kappa = a
    else:
        dsf, kappa = (0.0, 0.0), beta
    (ev, ec), f, w = direct_pairs(system, rc, coul, kappa, shift=not tail, with_excl=True, dsf=dsf)
    eb, fb, wb = oracle_bonded(system)
    out = dict(vdw=ev, coulomb=ec, bond=eb[0], angle=eb[1])
    f = f + fb
    w = w + wb
    q = ff.charge
    V = abs(np.linalg.det(system.cell[3:].reshape(3, 3)))
    if pme:
        grid = np.zeros(3, np.int32)
        O.orc().orc_pme_grid(np.ascontiguousarray(system.cell, np.float64), opts.pme_spacing, opts.pme_order, grid)
        er, fr, vr = oracle_pme(system, beta, grid, opts.pme_order)
        out["coulomb"] += er - KC * beta / np.sqrt(np.pi) * (q * q).sum()
        f = f + fr
        w = w + vr
        qt = q.sum()
        if abs(qt) > 1e-10:
            eb_ = -KC * np.pi * qt * qt / (2 * V * beta * beta)
            out["coulomb"] += eb_
            w[:3] += eb_
    elif coul == 2:
        a = opts.dsf_alpha
        out["coulomb"] -= (dsf[0] * rc / (2 * rc) + dsf[1] * rc / 2 + a / np.sqrt(np.pi)) * (q * q).sum() * KC
    if tail:
        cnt = np.bincount(ff.type, minlength=ff.ntypes).astype(np.float64)
        tt = np.zeros(2)
        O.orc().orc_tail(ff.ntypes, np.ascontiguousarray(ff.eps.reshape(-1)), np.ascontiguousarray((ff.sigma ** 6).reshape(-1)), cnt,
                         2 if ff.lj96 else 1, rc, V, tt)
        out["vdw"] += tt[0]
        w[:3] += tt[1] / 3
    out["w"] = w
    out["virial"] = w[:3].sum()
    return out, f


def melt_case(chains=16, seed=11):
    """the all-atom melt (128-atom chains, 1.0 g/cm³) of the 32,768- and 262,144-atom workloads, at a small size"""
    return ca.build_melt(chains, seed=seed)


def ref_pairs(system, rc, skin, coul, kappa=0.0, shift=False, dsf=(0.0, 0.0), prune=True, threads=4):
    """the reference's cluster path (oracle/_ref: List::build, positions, prune, kernel, pair and shift virial):
    (vdw, coulomb), forces (N, 3), w[6]"""
    ff = system.ff
    R = O.ref()
    c12, c6 = lj_tables(ff)
    f12, f6 = np.sqrt(np.diag(c12)).copy(), np.sqrt(np.diag(c6)).copy()
    x = np.ascontiguousarray(system.x, np.float64).reshape(-1)
    h = R.ref_eval_create(x, ff.n, np.ascontiguousarray(system.cell, np.float64), np.ascontiguousarray(ff.type, np.int32), ff.ntypes,
                          f12, f6, np.ascontiguousarray(ff.charge, np.float64), 2 if ff.lj96 else 1, coul, rc, skin, kappa, dsf[0],
                          dsf[1], int(shift), O._opt(ff.excl_offset, np.int32), O._opt(ff.excl_index, np.int32),
                          O._opt(ff.p14_offset, np.int32), O._opt(ff.p14_index, np.int32), ff.coul14, 0, None, threads)
    if not h:
        raise RuntimeError(R.ref_last_error().decode())
    try:
        f, e, w = np.zeros(x.size), np.zeros(2), np.zeros(6)
        if R.ref_eval_compute(h, x, 1, int(prune), f, e, w) != 0:
            raise RuntimeError(R.ref_last_error().decode())
    finally:
        R.ref_eval_free(h)
    return e, f.reshape(-1, 3), w


def dsf_constants(alpha, rc):
    """the DSF shift constants e0 = erfc(αrc)/rc, f0 = erfc(αrc)/rc² + 2α/√π e^{−α²rc²}/rc with A&S 7.1.26 erfc, as the
    evaluator computes them (field.cpp:585-590)"""
    P, A = 0.3275911, (0.254829592, -0.284496736, 1.421413741, -1.453152027, 1.061405429)
    t = 1 / (1 + P * alpha * rc)
    erfc_rc = t * (A[0] + t * (A[1] + t * (A[2] + t * (A[3] + t * A[4])))) * np.exp(-(alpha * rc) ** 2)
    return erfc_rc / rc, erfc_rc / rc ** 2 + 2 * alpha / np.sqrt(np.pi) * np.exp(-alpha * alpha * rc * rc) / rc
