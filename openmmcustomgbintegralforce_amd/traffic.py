"""HBM traffic per launch of a kernel from rocprofv3 counter runs.

Two steps, each under its own counter run (the counters do not fit one pass):

    rocprofv3 --pmc FETCH_SIZE -f csv -d <dir_f> -o run -- python3 -m openmmcustomgbintegralforce_amd.traffic run
    rocprofv3 --pmc WRITE_SIZE -f csv -d <dir_w> -o run -- python3 -m openmmcustomgbintegralforce_amd.traffic run
    python3 -m openmmcustomgbintegralforce_amd.traffic summarize <dir_f> <dir_w> > profiles/traffic_....json

``run`` evaluates the workload a few times and, in the same process, copies a device
buffer of known size (torch ``copy_``: reads and writes ``COPY_BYTES`` each).  The copy's
counter values calibrate the counters: factor = known bytes / counted bytes, applied to the
kernel's per-dispatch average.  ``summarize`` prints the JSON that bench.py reads
(``hbm_bytes_per_launch`` = fetch + write bytes per launch, calibrated).
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COPY_BYTES = 256 << 20

# the kernels of each pass (the binning pass: the brick or cell binning kernel and, for brick
# binning, the split that moves the points into their cells)
KERNELS = {"volume": ("k_volume",), "force": ("k_force",), "pairs": ("k_pairs",),
           "bin": ("k_point_bin", "k_point_split")}


def cmd_run(args) -> int:
    import numpy as np
    import torch

    sys.path.insert(0, ROOT)
    from openmmcustomgbintegralforce_amd import _native as N
    from openmmcustomgbintegralforce_amd import synthetic as S

    w = {"gbmv2-5000": (N.MODEL_GBMV, N.INTEGRAL_GBMV2, 50, 5000),
         "gbsw-25000": (N.MODEL_GBSW, N.INTEGRAL_GBSW, 38, 25000)}[args.workload]
    p = N.default_params(w[0], w[1])
    p.n_lebedev = w[2]
    s = S.globule(w[3], 0)
    eng = N.Engine(p, s.charge, s.radius)
    eng.execute(s.pos)
    eng.set_positions(s.pos)
    for _ in range(args.evals):
        eng.compute()
    eng.synchronize()
    eng.close()
    a = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"traffic run: {args.workload} x{args.evals}, copy {COPY_BYTES} B x3", flush=True)
    return 0


def _counters(directory: str):
    """{kernel name: [per-dispatch counter values]} from a rocprofv3 counter_collection csv."""
    files = glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {directory}")
    per = {}
    for path in files:
        with open(path) as fh:
            for row in csv.DictReader(fh):
                name = row.get("Kernel_Name", "")
                per.setdefault(name, {}).setdefault(row.get("Dispatch_Id", ""), 0.0)
                per[name][row.get("Dispatch_Id", "")] += float(row.get("Counter_Value", 0.0))
    return {k: list(v.values()) for k, v in per.items()}


def _average(per, match):
    vals = [x for k, v in per.items() if match(k) for x in v]
    if not vals:
        raise SystemExit("kernel not found in the counter data")
    return sum(vals) / len(vals), len(vals)


def cmd_summarize(args) -> int:
    fetch = _counters(args.fetch_dir)
    write = _counters(args.write_dir)
    is_copy = lambda n: ("copy" in n.lower() or "elementwise" in n.lower()) and "gbi" not in n  # noqa: E731
    # per evaluation: the sum over the pass's kernels of each one's per-dispatch average
    f_k = w_k = 0.0
    n_f, n_w, parts = [], [], {}
    for kname in KERNELS[args.kernel]:
        if not any(kname in n for n in fetch):
            continue  # a kernel the workload does not launch (cell binning has no split)
        fa, nf = _average(fetch, lambda n, kname=kname: kname in n)
        wa, nw = _average(write, lambda n, kname=kname: kname in n)
        f_k, w_k = f_k + fa, w_k + wa
        n_f.append(nf)
        n_w.append(nw)
        parts[kname] = {"fetch_kib": fa, "write_kib": wa}
    # the largest copy-like dispatch is the calibration copy
    f_c = max(x for k, v in fetch.items() if is_copy(k) for x in v)
    w_c = max(x for k, v in write.items() if is_copy(k) for x in v)
    kib = 1024.0
    f_factor = COPY_BYTES / (f_c * kib)
    w_factor = COPY_BYTES / (w_c * kib)
    out = {
        "kernel": args.kernel,
        "workload": args.workload,
        "fetch_kib_per_launch": f_k,
        "write_kib_per_launch": w_k,
        "copy_fetch_kib": f_c,
        "copy_write_kib": w_c,
        "fetch_factor": f_factor,
        "write_factor": w_factor,
        "hbm_bytes_per_launch": f_k * kib * f_factor + w_k * kib * w_factor,
        "launches": [n_f, n_w],
        "kernels": parts,
        "commands": [
            f"rocprofv3 --pmc FETCH_SIZE -f csv -d {args.fetch_dir} -o run -- python3 -m "
            f"openmmcustomgbintegralforce_amd.traffic run --workload {args.workload}",
            f"rocprofv3 --pmc WRITE_SIZE -f csv -d {args.write_dir} -o run -- python3 -m "
            f"openmmcustomgbintegralforce_amd.traffic run --workload {args.workload}",
            f"python3 -m openmmcustomgbintegralforce_amd.traffic summarize {args.fetch_dir} {args.write_dir} "
            f"--kernel {args.kernel} --workload {args.workload}",
        ],
    }
    print(json.dumps(out, indent=1))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--workload", default="gbmv2-5000", choices=["gbmv2-5000", "gbsw-25000"])
    r.add_argument("--evals", type=int, default=5)
    sm = sub.add_parser("summarize")
    sm.add_argument("fetch_dir")
    sm.add_argument("write_dir")
    sm.add_argument("--kernel", default="volume", choices=sorted(KERNELS))
    sm.add_argument("--workload", default="gbmv2-5000")
    args = ap.parse_args()
    return cmd_run(args) if args.cmd == "run" else cmd_summarize(args)


if __name__ == "__main__":
    sys.exit(main())
