"""Test infrastructure only: a numpy restatement of caps::minimise (core/src/relax.cpp:65-282) — steepest descent,
Polak–Ribière CG with its parabola step, L-BFGS (m = 10, two-loop recursion) and FIRE — driven by any energy/force
function, as the checker of the GPU minimiser. Pinned by the closed form of its test problems (bonded chains whose
minimum energy is exactly zero, a quadratic bowl whose minimum is known) in tests/test_oracle_relax.py."""
from collections import deque

import numpy as np

STOPS = ("iterations", "force", "energy", "line search")


def _max3(v):
    return float(np.sqrt((v.reshape(-1, 3) ** 2).sum(1).max()))


def minimise(evaluate, x, method="lbfgs", ftol=0.5, etol=1e-8, max_iterations=5000):
    """evaluate(x) -> (energy, forces); x (3N,) float64 is not modified. Returns (x, energy, fmax, iterations,
    evaluations, stopped_by)."""
    x = np.array(x, np.float64).reshape(-1).copy()
    nev = [0]

    def ev(p):
        nev[0] += 1
        e, f = evaluate(p)
        return e, np.array(f, np.float64).reshape(-1)

    e, f = ev(x)
    fmax = _max3(f)
    it = 0
    stop = "force" if fmax < ftol else "iterations"
    if method == "fire":
        nmin, finc, fdec, astart, falpha, dtmax, maxstep = 5, 1.1, 0.5, 0.1, 0.99, 0.5, 0.1
        dt, alpha, npos = 0.05, astart, 0
        v = np.zeros_like(x)
        while it < max_iterations and stop != "force":
            p = float(f @ v)
            if p > 0:
                vn, fn_ = np.sqrt(v @ v), np.sqrt(f @ f)
                v = (1 - alpha) * v + alpha * vn * f / max(fn_, 1e-300)
                npos += 1
                if npos > nmin:
                    dt, alpha = min(dt * finc, dtmax), alpha * falpha
            else:
                v[:] = 0.0
                dt *= fdec
                alpha, npos = astart, 0
            v = v + dt * f
            d = dt * v
            md = _max3(d)
            if md > maxstep:
                d *= maxstep / md
            x = x + d
            e, f = ev(x)
            fmax = _max3(f)
            if fmax < ftol:
                stop = "force"
            it += 1
        return x, e, fmax, it, nev[0], stop
    m = 10
    S, Y, rho = deque(), deque(), deque()
    fprev, dprev = None, np.zeros_like(x)
    cg_restarted = False
    step = 0.1
    step_max = 0.3 if method == "sd" else 0.5
    while it < max_iterations and stop != "force":
        if method == "sd":
            d = f.copy()
        elif method == "cg":
            if fprev is None:
                d = f.copy()
            else:
                den = float(fprev @ fprev)
                beta = max(0.0, float(f @ (f - fprev)) / den) if den > 0 else 0.0
                d = f + beta * dprev
        else:
            q = -f.copy()
            al = [0.0] * len(S)
            for j in range(len(S) - 1, -1, -1):
                al[j] = rho[j] * float(S[j] @ q)
                q -= al[j] * Y[j]
            gamma = 1.0 / 300.0
            if S:
                gamma = float(S[-1] @ Y[-1]) / float(Y[-1] @ Y[-1])
            q *= gamma
            for j in range(len(S)):
                b = rho[j] * float(Y[j] @ q)
                q += S[j] * (al[j] - b)
            d = -q
        slope = float(f @ d)
        if slope <= 0:
            d = f.copy()
            slope = float(f @ f)
            S.clear(), Y.clear(), rho.clear()
        md = _max3(d)
        a = min(1.0, step_max / md) if (method == "lbfgs" and S) else step / md
        ok, en, xn, fn = False, 0.0, None, None
        for _ in range(30):
            xn = x + a * d
            en, fn = ev(xn)
            if en <= e - 1e-4 * a * slope:
                ok = True
                break
            a *= 0.5
        if ok and method == "cg":
            curv = en - e + a * slope
            if curv > 0:
                aq = 0.5 * slope * a * a / curv
                if 0 < aq < 4 * a and abs(aq - a) > 1e-3 * a and aq * md <= step_max:
                    xq = x + aq * d
                    eq, fq = ev(xq)
                    if eq < en:
                        xn, fn, en, a = xq, fq, eq, aq
        if not ok:
            if S or fprev is not None:
                S.clear(), Y.clear(), rho.clear()
                fprev = None
                step = 0.02
                it += 1
                continue
            stop = "line search"
            break
        step = min(max(a * md * 1.5, 1e-4), step_max)
        if method == "lbfgs":
            s, y = xn - x, f - fn
            sy = float(s @ y)
            if sy > 1e-10:
                S.append(s), Y.append(y), rho.append(1.0 / sy)
                if len(S) > m:
                    S.popleft(), Y.popleft(), rho.popleft()
        fprev, dprev = f, d
        de = e - en
        x, f, e = xn, fn, en
        fmax = _max3(f)
        if fmax < ftol:
            stop = "force"
        elif 0 <= de < etol * max(1.0, abs(e)):
            if method == "cg" and not cg_restarted:
                cg_restarted = True
                fprev = None
                step = 0.1
                it += 1
                continue
            stop = "energy"
            it += 1
            break
        else:
            cg_restarted = False
        it += 1
    return x, e, fmax, it, nev[0], stop
