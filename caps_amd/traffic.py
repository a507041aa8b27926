"""HBM traffic of one kernel per launch, from rocprofv3's FETCH_SIZE and WRITE_SIZE counters (KiB per dispatch),
calibrated against a plain device-to-device copy of known size in the same run.

    # the program the counters are taken on: a 256 MiB copy, then plain PME evaluations of the 262,144-atom melt with the
    # kernel shapes the tuner picks
    rocprofv3 --pmc FETCH_SIZE -f csv -d OUT/fetch -o t -- python3 -m caps_amd.traffic run
    rocprofv3 --pmc WRITE_SIZE -f csv -d OUT/write -o t -- python3 -m caps_amd.traffic run
    # the per-launch bytes of a kernel over the last five evaluations (after the tuner's candidates), written as JSON
    # (and printed); a kernel launched twice per evaluation (k_fft_xy: forward and inverse) takes a sixth argument, 2
    python3 -m caps_amd.traffic report OUT/fetch/t_counter_collection.csv OUT/write/t_counter_collection.csv k_spread OUT.json

The copy reads and writes COPY_BYTES; each counter's scale is COPY_BYTES over what it reads for that copy, so the
kernel's traffic is its counter reading times that scale (the counters' own unit and any sampling cancel).
"""
import csv
import json
import sys

COPY_BYTES = 256 << 20
EVALS = 5   # the evaluations after tuning whose launches a report averages


def run():
    import torch
    import caps_amd as ca
    a = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    s = ca.build_melt(2048, seed=1)
    op = ca.EnergyOptions(cutoff=12.0, skin=2.0, coulomb=True, electrostatics="pme", ewald_rtol=1e-5, pme_order=5, pme_spacing=1.0,
                          tail=False, parts=5)
    ev = ca.Evaluator(s.ff, op)
    ev.compute(s.x, s.cell)
    ev.tune()   # the kernel shapes a run takes: the gather's workgroups and the transforms' tiles, picked by timing
    for _ in range(EVALS):
        ev.compute(s.x, s.cell)
    torch.cuda.synchronize()
    print("traffic run done", flush=True)


def _per_dispatch(path, counter):
    """{dispatch id: (kernel name, counter value)} of one counter's CSV, in dispatch order"""
    out = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        d = int(r["Dispatch_Id"])
        name, v = out.get(d, (r["Kernel_Name"], 0.0))
        out[d] = (name, v + float(r["Counter_Value"]))
    return dict(sorted(out.items()))


def report(fetch_csv, write_csv, kernel, out_json, per_eval="1"):
    res = {"kernel": kernel, "copy_bytes": COPY_BYTES, "launches_averaged": "the last %d evaluations' (after tuning)" % EVALS}
    for key, path, counter in (("fetch", fetch_csv, "FETCH_SIZE"), ("write", write_csv, "WRITE_SIZE")):
        d = _per_dispatch(path, counter)
        copies = [v for n, v in d.values() if "copyBuffer" in n]
        launches = [v for n, v in d.values() if kernel in n][-EVALS * int(per_eval):]
        if not copies or not launches:
            raise SystemExit("no %s readings for the copy or for %s in %s" % (counter, kernel, path))
        scale = COPY_BYTES / max(copies)
        res[key + "_counter_copy"] = max(copies)
        res[key + "_counter_kernel_mean"] = sum(launches) / len(launches)
        res[key + "_bytes"] = scale * sum(launches) / len(launches)
        res[key + "_launches"] = len(launches)
    res["traffic_bytes"] = res["fetch_bytes"] + res["write_bytes"]
    json.dump(res, open(out_json, "w"), indent=1)
    print(json.dumps(res))


if __name__ == "__main__":
    if len(sys.argv) >= 2 and sys.argv[1] == "run":
        run()
    elif len(sys.argv) in (6, 7) and sys.argv[1] == "report":
        report(*sys.argv[2:])
    else:
        print(__doc__)
        sys.exit(2)
