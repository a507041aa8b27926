"""Benchmark of the force-and-step path on one MI355X: the 262,144-atom all-atom melt (2048 chains of 128 atoms,
1.0 g/cm³; PME order 5 on 128³, r_c 12 Å, skin 2 Å, Ewald tolerance 1e-5; Langevin 300 K, τ 100 fs, dt 1 fs).

Prints one JSON line: ns/day of the timed steps (inputs resident in HBM; the time is the device loop's own clock,
DynamicsReport.seconds, around the steps alone), the roofline of the dominant kernel measured with HIP events on the
library's stream in an untimed pass, the force error against the FP64 oracle, the NVE energy drift, and the
reference's CPU path timed on this host's cores over the same work.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ns/day, 256k-atom PME melt: plain vs learned-RESPA at equal force RMSE"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
CHAINS, SEED = 2048, 1
RC, SKIN, RTOL, ORDER, SPACING = 12.0, 2.0, 1e-5, 5, 1.0
DT, TEMP, TAU = 1.0, 300.0, 100.0
KB = 0.0019872041            # kcal/(mol K) (core/src/dynamics.cpp:22)
KACC = 4.184e-4              # (kcal/mol/A)/(g/mol) -> A/fs^2 (core/src/dynamics.cpp:23)
CPU_SHARE = 16               # host threads of one GPU's share of the box


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def ns_per_day(seconds, steps, dt=DT):
    return dt * 1e-6 * 86400.0 * steps / seconds


def pme_bytes(K, n):
    """algorithmic HBM bytes of one PME evaluation (FP32 mesh): spread write 4K³, the forward and inverse real/complex
    passes 2·(4K³ + 8·K²(K/2+1)), the convolution 3·8·K²(K/2+1), the gather read 4K³, 28 B per atom"""
    half = 8.0 * K * K * (K // 2 + 1)
    return 4.0 * K ** 3 + 2 * (4.0 * K ** 3 + half) + 3 * half + 4.0 * K ** 3 + 28.0 * n


def kernel_bytes(name, n, K):
    """algorithmic bytes per launch of the HBM-bound kernels timed here (n atoms, K³ mesh)"""
    k3, half = float(K) ** 3, 8.0 * K * K * (K // 2 + 1)
    return {
        "k_splines": 24.0 * n + 8.0 * n + 16.0 * n + 16.0 * n + 4.0 * n,   # x, q in; weights, base points, tile counts out
        "k_spread": 4.0 * k3 + 32.0 * n,                                  # the mesh written once; atoms' weights read
        "k_fft_x_r2c": 4.0 * k3 + half, "k_fft_x_c2r": half + 4.0 * k3,
        "k_fft_y": 2.0 * half, "k_fft_z_conv": 2.0 * half + 4.0 * K * K * (K // 2 + 1),
        "k_gather": 4.0 * k3 + 32.0 * n + 12.0 * n,
        "k_baoab_open": 8.0 * (3 + 3 + 3 + 1 + 1) * n + 8.0 * 6 * n + 8.0 * 6 * n,
        "k_kick": 8.0 * (3 + 3 + 1) * n + 8.0 * 3 * n,
        "k_bonded": 8.0 * 3 * n * 3 + 8.0 * 3 * n,
        "k_sum_forces": 4.0 * 3 * n + 8.0 * 3 * n * 2,
    }.get(name)


def pairs_in_cutoff(system, rc):
    """P_rc: unordered non-excluded atom pairs closer than rc (the pair kernel's algorithmic unit), counted with a
    periodic k-d tree in the cubic box"""
    import numpy as np
    from scipy.spatial import cKDTree
    L = system.cell[3]
    x = np.mod(system.x, L)
    tree = cKDTree(x, boxsize=L)
    total = int(tree.count_neighbors(tree, rc) - len(x)) // 2
    ff = system.ff
    i = np.repeat(np.arange(ff.n), np.diff(ff.excl_offset))
    j = ff.excl_index[: ff.excl_offset[-1]]
    d = x[j] - x[i]
    d -= L * np.round(d / L)
    excl = int(((d * d).sum(1) < rc * rc).sum()) // 2
    return total - excl


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(system, budget_s=15.0):
    """The same partial step on this host's cores: the reference's own threaded PME (oracle/_ref, kspace.cpp:80-175 with
    the Fortran mesh code), the bonded terms (the oracle's C restatement of field.cpp:1980-2003) and a Langevin BAOAB
    step (numpy restatement of dynamics.cpp:613-627) — median step time over a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import caps_amd as ca
    import helpers as H
    import orcpy as O
    if not O.ref_available():
        return None
    ff = system.ff
    R = O.ref()
    x = np.ascontiguousarray(system.x, np.float64).reshape(-1).copy()
    v = np.zeros_like(x)
    q = np.ascontiguousarray(ff.charge, np.float64)
    cell = np.ascontiguousarray(system.cell, np.float64)
    beta = O.orc().orc_ewald_beta(RC, RTOL)
    grid = np.zeros(3, np.int32)
    O.orc().orc_pme_grid(cell, SPACING, ORDER, grid)
    threads = max(1, min(len(os.sched_getaffinity(0)), CPU_SHARE))
    m = np.repeat(ff.mass, 3)
    c1 = np.exp(-DT / TAU)
    c2 = np.sqrt(1 - c1 * c1)
    sd = np.sqrt(KB * TEMP * KACC / m)
    rng = np.random.default_rng(0)
    f = np.zeros_like(x)

    def forces(xx):
        f[:] = 0.0
        vir = np.zeros(6)
        R.ref_pme(xx, q, ff.n, cell, grid, ORDER, beta, threads, f, vir)
        s = ca.System(x=xx.reshape(-1, 3), cell=cell, ff=ff)
        _, fb, _ = H.oracle_bonded(s)
        f[:] += fb.reshape(-1)

    forces(x)
    times = []
    t_all = time.perf_counter()
    while time.perf_counter() - t_all < budget_s and len(times) < 30:
        t = time.perf_counter()
        v += 0.5 * DT * KACC * f / m
        x += 0.5 * DT * v
        v[:] = c1 * v + c2 * sd * rng.standard_normal(v.size)
        x += 0.5 * DT * v
        forces(x)
        v += 0.5 * DT * KACC * f / m
        times.append(time.perf_counter() - t)
    per = sorted(times)[len(times) // 2]
    return {"value": ns_per_day(per, 1), "unit": "ns/day", "cores": threads, "kind": "reference",
            "cpu": "%s (%d logical CPUs visible; %d threads used: one GPU's share of the host)" % (cpu_model(), os.cpu_count() or 0, threads),
            "sample": "%d Langevin steps of the 262,144-atom melt (the same partial step as the GPU leg: PME 128^3 order 5 by the "
                      "reference's threaded k-space code, bonded terms by the oracle's C restatement on one thread, the BAOAB "
                      "update in numpy), median %.3f s per step" % (len(times), per),
            "atom_steps_per_s": ff.n / per}


def force_accuracy(ca, system, opts):
    """per-atom force error of the GPU evaluation against the FP64 oracle (bonded terms + PME reciprocal part with its
    self term) at the given positions"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import helpers as H
    import orcpy as O
    ev = ca.Evaluator(system.ff, opts)
    t, f = ev.compute(system.x, system.cell)
    ev.close()
    beta = O.orc().orc_ewald_beta(RC, RTOL)
    grid = np.zeros(3, np.int32)
    O.orc().orc_pme_grid(np.ascontiguousarray(system.cell, np.float64), SPACING, ORDER, grid)
    er, fr, _ = H.oracle_pme(system, beta, grid, ORDER)
    eb, fb, _ = H.oracle_bonded(system)
    fo = fr + fb
    d = f - fo
    q = system.ff.charge
    ecoul = er - 332.06371 * beta / np.sqrt(np.pi) * float((q * q).sum())
    return {"force_rmse_abs": float(np.sqrt((d * d).sum(1).mean())), "unit": "kcal/mol/A",
            "force_rmse_rel": float(np.sqrt((d * d).sum() / (fo * fo).sum())),
            "force_max_abs": float(np.sqrt((d * d).sum(1)).max()),
            "coulomb_rel": abs(t.coulomb - ecoul) / abs(ecoul), "bond_rel": abs(t.bond - eb[0]) / abs(eb[0]),
            "what": "GPU (FP32 mesh, FP64 bonded and sums) against the FP64 oracle (oracle/orc_pme.c, orc_forces.c) at the "
                    "benchmark's positions after warm-up; bonded terms + PME reciprocal part"}


def ewald_beta(rc, rtol):
    """erfc(beta rc) = rtol by bisection on [0, 20/rc] (core/src/kspace.cpp:35-43)"""
    from scipy.special import erfc
    lo, hi = 0.0, 20.0 / rc
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        if erfc(mid * rc) > rtol:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def local_floor(ca, system, opts, k=4, radius=8.0, sample=32768, seed=SEED):
    """The error floor of any predictor that sees only the atoms within `radius` of each atom: over one k-step interval
    of the benchmark's Langevin dynamics, the increment of PME's reciprocal force (the GPU's, both ends) against the
    exact increment of the reciprocal-space pair field of the atoms within the radius (Ewald's erf(beta r)/r over the
    minimum image, summed in FP64 with numpy), on a random sample of atoms. Both relative to the whole force at the
    interval's end (bonded + PME), as the learned run judges its own predictions; the stale force's error (no increment)
    beside it."""
    import copy
    import numpy as np
    from scipy.spatial import cKDTree
    from scipy.special import erf
    rec = copy.copy(opts)
    rec.parts = 4
    ev_rec, ev_all = ca.Evaluator(system.ff, rec), ca.Evaluator(system.ff, opts)
    s = ca.System(x=system.x.copy(), cell=system.cell.copy(), ff=system.ff, v=None if system.v is None else system.v.copy())
    x0 = s.x.copy()
    fr0 = ev_rec.compute(x0, s.cell)[1]
    ca.run_dynamics(s, ca.DynamicsOptions(dt=DT, steps=k, temperature=TEMP, thermostat="langevin", tau_t=TAU, seed=seed + 7,
                                          new_velocities=s.v is None, thermo_every=10 ** 9, energy=opts), evaluator=ev_all)
    x1 = s.x.copy()
    fr1 = ev_rec.compute(x1, s.cell)[1]
    F1 = ev_all.compute(x1, s.cell)[1]
    ev_rec.close(), ev_all.close()
    h = s.cell[3:].reshape(3, 3)
    if np.count_nonzero(h - np.diag(np.diag(h))):
        return None                      # the minimum image below is for a rectangular cell
    L = np.diag(h)
    beta, q = ewald_beta(RC, RTOL), s.ff.charge
    n = s.ff.n
    idx = np.sort(np.random.default_rng(seed).choice(n, size=min(sample, n), replace=False))
    tree = cKDTree(np.mod(x0 - s.cell[:3], L), boxsize=L)
    nb = tree.query_ball_point(np.mod(x0[idx] - s.cell[:3], L), radius)
    cnt = np.array([len(v) for v in nb])
    i = np.repeat(idx, cnt)
    j = np.concatenate([np.asarray(v, np.int64) for v in nb])
    keep = i != j
    i, j, row = i[keep], j[keep], np.repeat(np.arange(idx.size), cnt)[keep]

    def field(x):
        d = x[i] - x[j]
        d -= L * np.round(d / L)
        r = np.sqrt((d * d).sum(1))
        g = (erf(beta * r) / r ** 2 - 2.0 * beta / np.sqrt(np.pi) * np.exp(-beta * beta * r * r) / r) / r * q[j]
        return np.stack([np.bincount(row, weights=g * d[:, m], minlength=idx.size) for m in range(3)], 1)

    dloc = 332.06371 * q[idx, None] * (field(x1) - field(x0))
    dtrue = (fr1 - fr0)[idx]
    den = (F1[idx] ** 2).sum()
    return {"k": k, "radius_A": radius, "atoms_sampled": int(idx.size), "pairs": int(i.size),
            "force_rmse_stale": float(np.sqrt((dtrue ** 2).sum() / den)),
            "force_rmse_local_floor": float(np.sqrt(((dtrue - dloc) ** 2).sum() / den)),
            "what": "over one k-step interval: the increment of the reciprocal force left unexplained by the exact field "
                    "change of every atom within the radius (what a predictor seeing only those atoms can reach at best), "
                    "relative to the whole force; the stale force's error beside it"}


def nve_drift(ca, system, opts, steps):
    """energy drift of NVE velocity Verlet from the benchmark state, kT/ns/atom (the reference's T5 pass criterion is
    |drift| < 2 kT/ns/atom at 300 K, core/src/bench.cpp:386-390)"""
    import numpy as np
    ev = ca.Evaluator(system.ff, opts)
    s = ca.System(x=system.x.copy(), cell=system.cell.copy(), ff=system.ff, v=None if system.v is None else system.v.copy())
    every = max(1, steps // 50)
    r = ca.run_dynamics(s, ca.DynamicsOptions(dt=DT, steps=steps, temperature=TEMP, thermostat="none", seed=SEED,
                                              new_velocities=s.v is None, thermo_every=every, energy=opts), evaluator=ev)
    ev.close()
    t = np.array([row["time_ps"] for row in r.thermo]) * 1e-3      # ns
    e = np.array([row["total"] for row in r.thermo])
    slope = float(np.polyfit(t, e, 1)[0])                          # kcal/mol/ns
    return {"steps": steps, "dt_fs": DT, "drift_kT_per_ns_per_atom": slope / (KB * TEMP) / system.ff.n,
            "fluctuation_kT_per_atom": float(e.std() / (KB * TEMP) / system.ff.n),
            "what": "NVE from the benchmark state, bonded terms + PME reciprocal part, total energy fitted linearly"}


def replica_batch(ca, torch, steps, count=64, seed=SEED):
    """Many melts of the all-atom chemistry stepped together: `count` replicas with sizes drawn uniformly from 3,000 to
    40,000 atoms (whole 128-atom chains), each its own cubic box and PME grid (order 4, 1.0 A), Langevin on a 500 -> 200 K
    ladder; PME reciprocal part + bonded terms (the pair path is held back). Timed four ways over the same steps: one
    replica after another through run_dynamics (as a cooling scan runs them), the batch with direct launches, the batch
    as grouped launches (each kernel once for all replicas, 16 steps per graph), and the batch with its tuned shape."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = rng.integers(3000, 40001, count)
    chains = [max(1, int(round(int(m) / 128))) for m in sizes]
    temps = np.linspace(500.0, 200.0, count)
    opts = ca.EnergyOptions(cutoff=RC, skin=SKIN, coulomb=True, electrostatics="pme", ewald_rtol=RTOL, pme_order=4, pme_spacing=SPACING,
                            tail=False, parts=5)
    log("building %d replicas" % count)
    systems = [ca.build_melt(c, seed=seed + 100 + k) for k, c in enumerate(chains)]
    atoms = sum(s.ff.n for s in systems)
    evs = [ca.Evaluator(s.ff, opts) for s in systems]
    out = {"workload": "%d all-atom melts of 3,000-40,000 atoms (seeded), PME order 4 at 1.0 A, Langevin 500 -> 200 K ladder, dt 1 fs; "
                       "PME reciprocal part + bonded terms" % count, "replicas": count, "atoms": atoms, "steps": steps,
           "box_A": [min(s.cell[3] for s in systems), max(s.cell[3] for s in systems)]}

    def sequential(nsteps, new_v):
        total = 0.0
        for k, (s, ev) in enumerate(zip(systems, evs)):
            d = ca.DynamicsOptions(dt=DT, steps=nsteps, temperature=float(temps[k]), thermostat="langevin", tau_t=TAU, seed=seed + k,
                                   thermo_every=10 ** 9, new_velocities=new_v, energy=opts)
            total += ca.run_dynamics(s, d, evaluator=ev).seconds
        return total
    sequential(5, True)
    secs = sequential(steps, False)
    out["one_after_another"] = {"seconds": secs, "atom_steps_per_s": atoms * steps / secs, "ns_per_day": ns_per_day(secs, steps)}
    b = ca.Batch(evs)
    for mode in ("direct", "grouped", "tuned"):
        o = ca.BatchOptions(dt=DT, steps=steps, tau_t=TAU, seed=seed, new_velocities=False, mode=mode, graph_steps=16)
        r = b.run(systems, temps, o)
        out[mode] = {"seconds": r.seconds, "atom_steps_per_s": r.atom_steps_per_s, "ns_per_day": r.ns_per_day, "mode": r.mode,
                     "branches": r.branches, "graph_steps": r.graph_steps, "tune_seconds": r.tune_seconds,
                     "speedup_vs_one_after_another": secs / r.seconds if r.seconds else None}
    b.close()
    for ev in evs:
        ev.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="leave out the CPU baseline")
    ap.add_argument("--learned-train", type=int, default=2000, help="PME frames the learned r-RESPA network trains on (-1: skip)")
    ap.add_argument("--replicas", type=int, default=64, help="melts in the replica-batch leg (0: skip)")
    ap.add_argument("--nve-steps", type=int, default=2000, help="NVE steps of the drift measurement (0: skip)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d asked)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import caps_amd as ca

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    log("building the %d-atom melt" % (CHAINS * 128))
    system = ca.build_melt(CHAINS, seed=SEED)
    n = system.ff.n
    grid = 128
    full = ca.EnergyOptions(cutoff=RC, skin=SKIN, coulomb=True, electrostatics="pme", ewald_rtol=RTOL, pme_order=ORDER,
                            pme_spacing=SPACING, tail=True)
    cfg = {"workload": "262,144-atom all-atom melt (2048 x 128-atom chains, 1.0 g/cm3), PME order 5 on 128^3, r_c 12 A, "
                       "Langevin 300 K, dt 1 fs", "atoms": n, "pme_grid": [grid] * 3, "cutoff_A": RC}
    out = {"metric": METRIC, "value": None, "unit": "ns/day", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 pairs/mesh, f64 sums/integration", "data": "synthetic (seeded builder, random charges)", "config": cfg}
    dyn = ca.DynamicsOptions(dt=DT, temperature=TEMP, thermostat="langevin", tau_t=TAU, seed=SEED, new_velocities=True,
                             thermo_every=10 ** 9, energy=full)
    # the full step: cluster pair list and kernel, PME, bonded terms, Langevin
    ev = ca.Evaluator(system.ff, full)
    try:
        dyn.steps = 1
        ca.run_dynamics(system, dyn, evaluator=ev)
        measured = "full"
    except ca.DeviceError as e:
        out["status"] = "the full step did not run: %s" % e
        measured = "without_pairs"
    if measured == "without_pairs":
        # the parts that run: PME's reciprocal part, the bonded terms and the Langevin step (not the metric's step)
        ev.close()
        part = ca.EnergyOptions(cutoff=RC, skin=SKIN, coulomb=True, electrostatics="pme", ewald_rtol=RTOL, pme_order=ORDER,
                                pme_spacing=SPACING, tail=False, parts=5)
        ev = ca.Evaluator(system.ff, part)
        dyn.energy = part
    log("warm-up (%d steps)" % a.warmup)
    dyn.steps = a.warmup
    ca.run_dynamics(system, dyn, evaluator=ev)
    dyn.new_velocities = False
    # the mesh transforms' tile shape and the gather's workgroups for this grid, picked by timing
    cfg["transform_tiles"] = list(ev.tune())
    cfg["gather_shape"] = "runs of 64 sorted atoms" if ev.gather_shape() == 1 else "tiles"
    # an untimed profiled pass: every kernel's launch time (two HIP events per launch on the library's stream)
    log("profiled pass")
    ev.profile(True)
    dyn.steps = min(a.steps, 40)
    ca.run_dynamics(system, dyn, evaluator=ev)
    kernels = ev.profile_kernels()
    stages = ev.profile_read()
    ev.profile(False)
    name = max(kernels.items(), key=lambda kv: kv[1][0])[0]
    # the dominant kernel alone: events around its launches only, in a second untimed pass of the same length as the
    # timed one
    ev.profile_only(name)
    ev.profile(True)
    dyn.steps = a.steps
    ca.run_dynamics(system, dyn, evaluator=ev)
    timed = ev.profile_kernels()
    ev.profile(False)
    ev.profile_only(None)
    # the timed steps: nothing recorded but the kernels themselves; the device loop's clock around the steps alone
    log("timed steps (%d)" % a.steps)
    rep = ca.run_dynamics(system, dyn, evaluator=ev)
    secs = rep.seconds
    ms_step = 1e3 * secs / a.steps
    nsday = ns_per_day(secs, a.steps)
    if measured == "full":
        out["value"], out["ms_per_step"] = nsday, ms_step
    else:
        out["without_pairs"] = {"ms_per_step": ms_step, "ns_per_day": nsday, "atom_steps_per_s": n * a.steps / secs,
                                "what": "PME reciprocal part + bonded terms + Langevin step, no pair list or pair kernel"}
    out["stages_ms_per_eval"] = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in stages.items()}
    out["kernels_ms_per_launch"] = {k: v[0] / v[1] for k, v in sorted(kernels.items(), key=lambda kv: -kv[1][0])}
    # the dominant kernel and its roofline
    tot, cnt = timed[name]
    avg_s = tot / cnt * 1e-3
    if name == "k_pairs":
        # FP32 vector ALU: 60 FLOP per in-cutoff pair on force-only steps (nbnxm.cpp:460-533), the full list walking each
        # pair from both sides counts it once
        prc = pairs_in_cutoff(system, RC)
        ach = 60.0 * prc / avg_s / 1e12
        out["roofline"] = {"kernel": name, "bound": "valu", "unit": "TFLOP/s", "peak": FP32_PEAK_TFLOPS, "achieved": ach,
                           "frac": ach / FP32_PEAK_TFLOPS, "traffic": None, "avg_launch_ms": tot / cnt, "launches_timed": cnt,
                           "pairs_in_cutoff": prc, "algorithmic_flop_per_launch": 60.0 * prc}
    else:
        b = kernel_bytes(name, n, grid)
        ach = b / avg_s / 1e9 if b else None
        out["roofline"] = {"kernel": name, "bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS, "achieved": ach,
                           "frac": ach / HBM_PEAK_GBS if ach else None, "traffic": None, "avg_launch_ms": tot / cnt,
                           "launches_timed": cnt, "algorithmic_bytes_per_launch": b}
        # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE measurement of the same kernel and workload, when there
        # is one (profiles/<kernel>_traffic_262k_melt.json, written by caps_amd/traffic.py)
        tpath = os.path.join(ROOT, "profiles", "%s_traffic_262k_melt.json" % name)
        if os.path.exists(tpath):
            tr = json.load(open(tpath))
            out["roofline"]["traffic"] = tr["traffic_bytes"]
            out["roofline"]["traffic_source"] = os.path.relpath(tpath, ROOT)
    # PME per evaluation: its kernels' event-timed totals over the pass's evaluations (one spread each), so a kernel that
    # runs in some evaluations only (the splines, done by the integrator's opening launch in the dynamics) counts as often
    # as it ran
    evals = kernels["k_spread"][1] if "k_spread" in kernels else 0
    pme_ms = sum(v[0] for k, v in kernels.items() if k.startswith(("k_splines", "k_scan_tiles", "k_tile", "k_spread", "k_fft", "k_gather"))) / evals if evals else 0.0
    out["pme"] = {"ms_per_eval": pme_ms, "algorithmic_bytes": pme_bytes(grid, n),
                  "achieved_GBs": pme_bytes(grid, n) / (pme_ms * 1e-3) / 1e9 if pme_ms else None,
                  "what": "the PME kernels' times per evaluation from two HIP events around each launch in the profiled pass "
                          "(the events add 1-3 us to each launch); the per-atom bookkeeping runs inside the integrator's opening "
                          "launch in the dynamics"}
    out["atom_steps_per_s"] = n * a.steps / secs
    # learned r-RESPA on the same step: full PME every 4 steps, the network's predicted increment in between, after
    # training on PME frames; judged at every full step against PME and against the stale force
    if a.learned_train >= 0:
        log("learned r-RESPA (%d training frames)" % a.learned_train)
        lo = ca.DynamicsOptions(dt=DT, temperature=TEMP, thermostat="langevin", tau_t=TAU, seed=SEED, new_velocities=False,
                                thermo_every=4, energy=dyn.energy, learned_k=4, learned_tol=1e-3, learned_train=a.learned_train)
        lo.steps = a.learned_train + 4
        ca.run_dynamics(system, lo, evaluator=ev)
        lo.steps, lo.learned_train = a.steps, 0
        out["learned_respa"] = {"k": 4, "training_frames": a.learned_train,
                                "rmse_definition": "|F_slow predicted - F_slow PME| / |F| summed over all atoms at each judging full "
                                                   "step, F the whole force this step evaluates (bonded + PME; the pair path is held "
                                                   "back), the prediction made k steps earlier"}
        for tol in (1e-3, 5e-3):
            lo.learned_tol = tol
            rep = ca.run_dynamics(system, lo, evaluator=ev)
            judged = [r for r in rep.thermo if r["learned_rmse"] > 0]
            leg = {"tolerance": tol, "ns_per_day": ns_per_day(rep.seconds, a.steps), "ms_per_step": 1e3 * rep.seconds / a.steps,
                   "learned_steps": rep.learned_steps, "full_steps": rep.full_steps, "judgments": rep.judgments,
                   "fallbacks": rep.fallbacks,
                   "force_rmse_learned": float(np.mean([r["learned_rmse"] for r in judged])) if judged else None,
                   "force_rmse_stale": float(np.mean([r["learned_rmse_stale"] for r in judged])) if judged else None,
                   "speedup_vs_plain": secs / rep.seconds}
            out["learned_respa"]["at_tolerance_%g" % tol] = leg
        log("the error floor of an 8 A local predictor")
        out["learned_respa"]["floor_8A"] = local_floor(ca, system, dyn.energy)
    # r-RESPA (dynamics.cpp:113-123, 643-656) on the same system and terms: Bussi 300 K, the slow forces (here PME's
    # reciprocal part) kicking over a 2 fs outer step, the bonded terms over two 1 fs inner steps
    ro = ca.DynamicsOptions(dt=2.0, temperature=TEMP, thermostat="bussi", tau_t=TAU, seed=SEED, new_velocities=False,
                            thermo_every=10 ** 9, energy=dyn.energy, respa=2, steps=10)
    ca.run_dynamics(system, ro, evaluator=ev)
    ro.steps = a.steps
    rep = ca.run_dynamics(system, ro, evaluator=ev)
    out["respa"] = {"ns_per_day": ns_per_day(rep.seconds, a.steps, ro.dt), "ms_per_step": 1e3 * rep.seconds / a.steps, "outer_dt_fs": ro.dt,
                    "inner_steps": 2, "thermostat": "bussi",
                    "what": "outer kicks with the slow forces (PME reciprocal part), bonded terms on the inner steps"}
    ev.close()
    log("force error against the FP64 oracle")
    out["accuracy"] = force_accuracy(ca, system, dyn.energy)
    if a.nve_steps > 0:
        log("NVE drift (%d steps)" % a.nve_steps)
        out["nve"] = nve_drift(ca, system, dyn.energy, a.nve_steps)
    if a.replicas > 0:
        out["replica_batch"] = replica_batch(ca, torch, a.steps, a.replicas)
    if not a.no_cpu:
        log("the reference CPU path")
        cb = cpu_baseline(system)
        out["cpu_baseline"] = cb
        if cb:
            out["gpu_vs_cpu_same_step"] = (n * a.steps / secs) / cb["atom_steps_per_s"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
