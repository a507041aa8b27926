"""Benchmark of the force-and-step path on one MI355X: the 262,144-atom all-atom melt (2048 chains of 128 atoms,
1.0 g/cm³; PME order 5 on 128³, r_c 12 Å, skin 2 Å, Ewald tolerance 1e-5; Langevin 300 K, τ 100 fs, dt 1 fs).

Prints one JSON line: ns/day of the timed steps (inputs resident in HBM), the roofline of the dominant kernel measured
live with HIP events on the library's stream, and the reference CPU path timed on this host's cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
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


def pme_bytes(K, n):
    """algorithmic HBM bytes of one PME evaluation (FP32 mesh): spread write 4K³, the forward and inverse real/complex
    passes 2·(4K³ + 8·K²(K/2+1)), the convolution 3·8·K²(K/2+1), the gather read 4K³, 28 B per atom"""
    half = 8.0 * K * K * (K // 2 + 1)
    return 4.0 * K ** 3 + 2 * (4.0 * K ** 3 + half) + 3 * half + 4.0 * K ** 3 + 28.0 * n


def kernel_bytes(name, n, K):
    """algorithmic bytes per launch of the kernels timed here (n atoms, K³ mesh)"""
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


def cpu_baseline(system, budget_s=20.0):
    """the reference's own cluster path (oracle/_ref: list build, prune, AVX2 pair kernel, threaded PME) on this host's
    cores: ns/day from the time of one force evaluation (the list built once outside the timing, as it is amortised)"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import helpers as H
    import orcpy as O
    if not O.ref_available():
        return None
    ff = system.ff
    R = O.ref()
    c12, c6 = H.lj_tables(ff)
    f12, f6 = np.sqrt(np.diag(c12)).copy(), np.sqrt(np.diag(c6)).copy()
    x = np.ascontiguousarray(system.x, np.float64).reshape(-1)
    beta = O.orc().orc_ewald_beta(RC, RTOL)
    grid = np.zeros(3, np.int32)
    O.orc().orc_pme_grid(np.ascontiguousarray(system.cell, np.float64), SPACING, ORDER, grid)
    threads = min(os.cpu_count() or 1, 16)
    h = R.ref_eval_create(x, ff.n, np.ascontiguousarray(system.cell, np.float64), np.ascontiguousarray(ff.type, np.int32), ff.ntypes, f12,
                          f6, np.ascontiguousarray(ff.charge, np.float64), 1, 1, RC, SKIN, beta, 0.0, 0.0, 0, O._opt(ff.excl_offset, np.int32),
                          O._opt(ff.excl_index, np.int32), None, None, 1.0, ORDER, O._opt(grid, np.int32), threads)
    if not h:
        return None
    try:
        used = R.ref_eval_threads(h)
        f, e, w = np.zeros(x.size), np.zeros(2), np.zeros(6)
        times = []
        t_all = time.perf_counter()
        while time.perf_counter() - t_all < budget_s and len(times) < 20:
            t = time.perf_counter()
            R.ref_eval_compute(h, x, 0, 1, f, e, w)
            times.append(time.perf_counter() - t)
    finally:
        R.ref_eval_free(h)
    per = sorted(times)[len(times) // 2]
    return {"value": DT * 1e-6 * 86400.0 / per, "unit": "ns/day", "cores": used, "kind": "reference",
            "sample": "%d force evaluations (prune, pair kernel, PME 128^3 order 5) of the 262,144-atom melt, median %.3f s; "
                      "list built once outside the timing; bonded terms and the integrator left out" % (len(times), per)}


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
    print("bench: building %d replicas" % count, file=sys.stderr, flush=True)
    systems = [ca.build_melt(c, seed=seed + 100 + k) for k, c in enumerate(chains)]
    atoms = sum(s.ff.n for s in systems)
    evs = [ca.Evaluator(s.ff, opts) for s in systems]
    out = {"workload": "%d all-atom melts of 3,000-40,000 atoms (seeded), PME order 4 at 1.0 A, Langevin 500 -> 200 K ladder, dt 1 fs; "
                       "PME reciprocal part + bonded terms" % count, "replicas": count, "atoms": atoms, "steps": steps,
           "box_A": [min(s.cell[3] for s in systems), max(s.cell[3] for s in systems)]}
    # one replica after another (warmed first)
    def sequential(nsteps, new_v):
        total = 0.0
        for k, (s, ev) in enumerate(zip(systems, evs)):
            d = ca.DynamicsOptions(dt=DT, steps=nsteps, temperature=float(temps[k]), thermostat="langevin", tau_t=TAU, seed=seed + k,
                                   thermo_every=10 ** 9, new_velocities=new_v, energy=opts)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ca.run_dynamics(s, d, evaluator=ev)
            torch.cuda.synchronize()
            total += time.perf_counter() - t0
        return total
    sequential(5, True)
    secs = sequential(steps, False)
    out["one_after_another"] = {"seconds": secs, "atom_steps_per_s": atoms * steps / secs, "ns_per_day": steps * DT * 1e-6 * 86400 / secs}
    b = ca.Batch(evs)
    for mode in ("direct", "grouped", "tuned"):
        o = ca.BatchOptions(dt=DT, steps=steps, tau_t=TAU, seed=seed, new_velocities=False, mode=mode, graph_steps=16)
        r = b.run(systems, temps, o)
        out[mode] = {"seconds": r.seconds, "atom_steps_per_s": r.atom_steps_per_s, "ns_per_day": r.ns_per_day, "mode": r.mode,
                     "branches": r.branches, "graph_steps": r.graph_steps, "tune_seconds": r.tune_seconds,
                     "speedup_vs_one_after_another": secs / r.seconds if r.seconds else None}
    b.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="leave out the CPU baseline")
    ap.add_argument("--learned-train", type=int, default=2000, help="PME frames the learned r-RESPA network trains on (-1: skip)")
    ap.add_argument("--replicas", type=int, default=64, help="melts in the replica-batch leg (0: skip)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d asked)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import caps_amd as ca

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    print("bench: building the %d-atom melt" % (CHAINS * 128), file=sys.stderr, flush=True)
    system = ca.build_melt(CHAINS, seed=SEED)
    n = system.ff.n
    grid = 128
    full = ca.EnergyOptions(cutoff=RC, skin=SKIN, coulomb=True, electrostatics="pme", ewald_rtol=RTOL, pme_order=ORDER,
                            pme_spacing=SPACING, tail=True)
    cfg = {"workload": "262,144-atom all-atom melt (2048 x 128-atom chains, 1.0 g/cm3), PME order 5 on 128^3, r_c 12 A, "
                       "Langevin 300 K, dt 1 fs", "atoms": n, "pme_grid": [grid] * 3, "cutoff_A": RC}
    out = {"metric": METRIC, "value": None, "unit": "ns/day", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 pairs/mesh, f64 sums/integration",
           "data": "synthetic (seeded builder, random charges)", "config": cfg}
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
        part = ca.EnergyOptions(cutoff=RC, skin=SKIN, coulomb=True, electrostatics="pme", ewald_rtol=RTOL, pme_order=ORDER,
                                pme_spacing=SPACING, tail=False, parts=5)
        ev = ca.Evaluator(system.ff, part)
        dyn.energy = part
    dyn.steps = a.warmup
    ca.run_dynamics(system, dyn, evaluator=ev)
    dyn.new_velocities = False
    # the mesh transforms' tile shape for this grid, picked by timing (lines per workgroup along a, and along b and c)
    cfg["transform_tiles"] = list(ev.tune())
    # an untimed profiled pass: every kernel's launch time (two HIP events per launch on the library's stream)
    ev.profile(True)
    dyn.steps = min(a.steps, 40)
    ca.run_dynamics(system, dyn, evaluator=ev)
    kernels = ev.profile_kernels()
    stages = ev.profile_read()
    ev.profile(False)
    name = max(kernels.items(), key=lambda kv: kv[1][0])[0]
    # the timed steps, with events around the dominant kernel's launches only (the roofline's average launch time)
    ev.profile_only(name)
    ev.profile(True)
    dyn.steps = a.steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ca.run_dynamics(system, dyn, evaluator=ev)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    timed = ev.profile_kernels()
    ev.profile(False)
    ev.profile_only(None)
    ms_step = 1e3 * secs / a.steps
    nsday = DT * 1e-6 * 86400.0 / (secs / a.steps)
    if measured == "full":
        out["value"], out["ms_per_step"] = nsday, ms_step
    else:
        out["without_pairs"] = {"ms_per_step": ms_step, "ns_per_day": nsday,
                                "what": "PME reciprocal part + bonded terms + Langevin step, no pair list or pair kernel"}
    out["stages_ms_per_eval"] = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in stages.items()}
    out["kernels_ms_per_launch"] = {k: v[0] / v[1] for k, v in sorted(kernels.items(), key=lambda kv: -kv[1][0])}
    # the dominant kernel and its roofline, its launches timed inside the timed region
    tot, cnt = timed[name]
    avg_s = tot / cnt * 1e-3
    if name == "k_pairs":
        out["roofline"] = {"kernel": name, "bound": "mfma", "unit": "TFLOP/s", "peak": FP32_PEAK_TFLOPS, "achieved": None, "frac": None,
                           "traffic": None}
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
    pme_ms = sum(v[0] / v[1] for k, v in kernels.items() if k.startswith(("k_splines", "k_tile", "k_spread", "k_fft", "k_gather")))
    out["pme"] = {"ms_per_eval": pme_ms, "algorithmic_bytes": pme_bytes(grid, n),
                  "achieved_GBs": pme_bytes(grid, n) / (pme_ms * 1e-3) / 1e9 if pme_ms else None}
    out["atom_steps_per_s"] = n * a.steps / secs
    # learned r-RESPA on the same step: full PME every 4 steps, the network's predicted increment in between, after
    # training on PME frames; judged at every full step against PME and against the stale force
    if a.learned_train >= 0:
        print("bench: learned r-RESPA (%d training frames)" % a.learned_train, file=sys.stderr, flush=True)
        lo = ca.DynamicsOptions(dt=DT, temperature=TEMP, thermostat="langevin", tau_t=TAU, seed=SEED, new_velocities=False,
                                thermo_every=4, energy=dyn.energy, learned_k=4, learned_tol=1e-3, learned_train=a.learned_train)
        lo.steps = a.learned_train + 4
        ca.run_dynamics(system, lo, evaluator=ev)
        lo.steps, lo.learned_train = a.steps, 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rep = ca.run_dynamics(system, lo, evaluator=ev)
        torch.cuda.synchronize()
        lsecs = time.perf_counter() - t0
        judged = [r for r in rep.thermo if r["learned_rmse"] > 0]
        out["learned_respa"] = {
            "ns_per_day": DT * 1e-6 * 86400.0 / (lsecs / a.steps), "ms_per_step": 1e3 * lsecs / a.steps, "k": 4,
            "training_frames": a.learned_train, "learned_steps": rep.learned_steps, "full_steps": rep.full_steps,
            "fallbacks": rep.fallbacks,
            "force_rmse_learned": float(np.mean([r["learned_rmse"] for r in judged])) if judged else None,
            "force_rmse_stale": float(np.mean([r["learned_rmse_stale"] for r in judged])) if judged else None,
            "rmse_definition": "|F_slow predicted - F_slow PME| / |F total| over all atoms at each full step",
            "speedup_vs_plain": (secs / lsecs)}
        # the same trained network with a tolerance above its judged error, so the intervals between full steps are
        # predicted: the speed of learned r-RESPA when it runs, at the error it reaches
        lo.learned_tol = 5e-3
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rep = ca.run_dynamics(system, lo, evaluator=ev)
        torch.cuda.synchronize()
        esecs = time.perf_counter() - t0
        judged = [r for r in rep.thermo if r["learned_rmse"] > 0]
        out["learned_respa"]["at_tolerance_5e-3"] = {
            "ns_per_day": DT * 1e-6 * 86400.0 / (esecs / a.steps), "ms_per_step": 1e3 * esecs / a.steps,
            "learned_steps": rep.learned_steps, "full_steps": rep.full_steps, "fallbacks": rep.fallbacks,
            "force_rmse_learned": float(np.mean([r["learned_rmse"] for r in judged])) if judged else None,
            "speedup_vs_plain": secs / esecs}
    # r-RESPA (dynamics.cpp:113-123, 643-656) on the same system and terms: Bussi 300 K, the slow forces (here PME's
    # reciprocal part) kicking over a 2 fs outer step, the bonded terms over two 1 fs inner steps
    ro = ca.DynamicsOptions(dt=2.0, temperature=TEMP, thermostat="bussi", tau_t=TAU, seed=SEED, new_velocities=False,
                            thermo_every=10 ** 9, energy=dyn.energy, respa=2, steps=10)
    ca.run_dynamics(system, ro, evaluator=ev)
    ro.steps = a.steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ca.run_dynamics(system, ro, evaluator=ev)
    torch.cuda.synchronize()
    rsecs = time.perf_counter() - t0
    out["respa"] = {"ns_per_day": ro.dt * 1e-6 * 86400.0 / (rsecs / a.steps), "ms_per_step": 1e3 * rsecs / a.steps, "outer_dt_fs": ro.dt,
                    "inner_steps": 2, "thermostat": "bussi",
                    "what": "outer kicks with the slow forces (PME reciprocal part), bonded terms on the inner steps"}
    if a.replicas > 0:
        out["replica_batch"] = replica_batch(ca, torch, a.steps, a.replicas)
    if not a.no_cpu:
        print("bench: the reference CPU path", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(system)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
