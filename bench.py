#!/usr/bin/env python3
"""GB energy+force evaluations per second on one MI355X, with force error vs FP64.

A step is one energy+force evaluation of one system (the whole HIP path: lookup grid,
volume integral, Born radii, pair energy, chain-rule forces).  Inputs are resident in
HBM before the timed region starts.  The default workload is the 5,000-atom GBMV
(method II) globule with 24 x 50 quadrature; ``--workload gbsw-25000`` runs the
25,000-atom GBSW all-pairs globule.

Prints one JSON line (rank 0).  Beside the throughput it reports:
* force relative RMS and energy relative error against the FP64 oracle (same input);
* ``roofline`` of the dominant (longest) pass and ``rooflines`` of every modelled pass:
  volume, chain-rule and pair passes as algorithmic FLOPs per launch (interaction counts x
  fixed weights, openmmcustomgbintegralforce_amd/flops.py) against the 157.3 TFLOP/s FP32
  vector peak, the binning pass as algorithmic bytes against the 8 TB/s HBM peak; durations
  from HIP events on the library's stream (the binning pass includes its scan and worklist
  launches); ``traffic`` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summaries;
* ``cpu_baseline``: the oracle in the reference's loop structure (dense, FP64, 1 thread)
  timed on this host over a bounded sample;
* ``extra`` (default workload only; ``--no-extras`` skips it): evals/s of the 25,000-atom
  GBSW globule, and the 64-system mix batched in one context against the same systems as
  per-system launches (sequential on one stream, concurrent on 64 streams, and plain
  gbi_execute calls with positions in and forces out).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "GB energy+force evals/s at 2.5k–25k atoms (1 GPU); force RMS err vs FP64"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    "gbmv2-5000": dict(model="gbmv", integral=2, n_atoms=5000, n_lebedev=50, seed=0,
                       desc="CharmmGBMVForce GBMVIntegralTypeII, 5,000-atom globule, 24x50 quadrature, "
                            "lookup grid 0.15 nm, no cutoff"),
    "gbsw-25000": dict(model="gbsw", integral=0, n_atoms=25000, n_lebedev=38, seed=0,
                       desc="CharmmGBSWForce, 25,000-atom globule, 24x38 quadrature, smoothing 0.03 nm, "
                            "all pairs; exact cell lists on a 0.20 nm grid (same V as the reference's "
                            "0.15 nm lookup grid, which is complete for this force)"),
    "gbsw-2500": dict(model="gbsw", integral=0, n_atoms=2500, n_lebedev=38, seed=0,
                      desc="CharmmGBSWForce, 2,500-atom globule, 24x38 quadrature, all pairs"),
    "adjoint-2000x8": dict(model="adjoint", integral=0, n_atoms=2000, n_lebedev=38, seed=26,
                           desc="CharmmGBSWForce adjoint: 2,000-atom globule, 8 conformations (0.02 nm Gaussian "
                                "perturbations of one base geometry) batched in one context, 24x38 quadrature; "
                                "energy, forces, dE/dR, dE/dq and dE/d(GB constants) per conformation, beside "
                                "the same batch without the adjoint"),
    "gbsw-25000-cutoff": dict(model="gbsw", integral=0, n_atoms=25000, n_lebedev=38, seed=0, method=1, cutoff=1.6,
                              desc="CharmmGBSWForce CutoffNonPeriodic 1.6 nm, 25,000-atom globule, 24x38 "
                                   "quadrature (pairs beyond the cutoff and their 64-atom blocks skipped; the "
                                   "chain rule truncated at the cutoff as the reference does)"),
    "gbmv2-liquid-periodic": dict(model="gbmv", integral=2, n_atoms=0, n_lebedev=50, seed=1, method=2, cutoff=1.6,
                                  box=(4.0, 4.0, 4.0),
                                  desc="CharmmGBMVForce GBMVIntegralTypeII CutoffPeriodic 1.6 nm: atoms filling a "
                                       "4 nm box at 100 atoms/nm^3 (6,400 atoms), 24x50 quadrature, periodic "
                                       "lookup grid 0.15 nm"),
    "gbmv2-5000-expressions": dict(model="gbmv", integral=2, n_atoms=5000, n_lebedev=50, seed=0, expressions=True,
                                   desc="CharmmGBMVForce GBMVIntegralTypeII, 5,000-atom globule, 24x50 quadrature, "
                                        "the canonical GB forms given as general expressions: value and single-particle "
                                        "kernel over FP64 dual numbers, pair kernel over FP32 dual numbers, generated "
                                        "and compiled with hiprtc"),
    "batch64": dict(model="batch", integral=-1, n_atoms=0, n_lebedev=0, seed=0,
                    desc="64 systems in one batched context, N ~ U[300, 6000] (seeded), alternating "
                         "CharmmGBSWForce 24x38 and CharmmGBMVForce GBMVIntegralTypeII 24x50, hipGraph replay; "
                         "beside it the same 64 systems as 64 separate contexts, enqueued back to back on one "
                         "stream (sequential), on one stream each (concurrent) and through gbi_execute one "
                         "after another (plain)"),
}


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="gbmv2-5000", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-baseline-evals", type=int, default=3,
                    help="evaluations of the single-threaded oracle timed for cpu_baseline, median taken (0: skip)")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP64 accuracy check")
    ap.add_argument("--no-baseline", action="store_true",
                    help="batch64: skip the per-system contexts (for profiling the batched launch alone)")
    ap.add_argument("--no-extras", action="store_true",
                    help="gbmv2-5000: skip the extra keys (gbsw-25000 evals/s, batch64 against per-system launches)")
    return ap.parse_args()


def kind_name(w) -> str:
    return "gbsw" if w["model"] == "gbsw" or w["integral"] == 0 else ("gbmv1" if w["integral"] == 1 else "gbmv2")


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note, not built).", file=sys.stderr)
        return 2
    from openmmcustomgbintegralforce_amd import _native as N
    from openmmcustomgbintegralforce_amd import flops as FL
    from openmmcustomgbintegralforce_amd import synthetic as S

    w = WORKLOADS[args.workload]
    if w["model"] == "batch":
        return run_batch(args, w)
    if w["model"] == "adjoint":
        return run_adjoint(args, w)
    kind = kind_name(w)
    box = w.get("box")
    sysm = S.liquid(box, w["seed"]) if box else S.globule(w["n_atoms"], w["seed"])
    n_atoms = sysm.n
    if w["model"] == "gbsw":
        p = N.default_params(N.MODEL_GBSW)
    else:
        p = N.default_params(N.MODEL_GBMV, w["integral"])
    p.n_lebedev = w["n_lebedev"]
    p.nonbonded_method = w.get("method", N.NO_CUTOFF)
    p.cutoff = w.get("cutoff", 1.0)
    if w.get("expressions"):
        G = dict(alpha0=p.alpha0, alpha1=p.alpha1, eta=p.eta, slope=p.slope, shift=p.shift, P=p.prefactor)
        eng = N.Engine.custom(p, ["q", "radius"], np.stack([sysm.charge, sysm.radius], 1), 1, ["I4", "I7"],
                              [("B", "slope*1/(alpha0*(1/eta-I4)+alpha1*(1/(4*eta^4)-I7)^(1/4))+shift")],
                              [("0.5*P*q^2/B", N.TERM_SINGLE),
                               ("P*q1*q2/sqrt(r^2+B1*B2*exp(-r^2/(4*B1*B2)))", N.TERM_PAIR_NO_EXCLUSIONS)], G)
    else:
        eng = N.Engine(p, sysm.charge, sysm.radius)
    if box:
        eng.set_periodic_box(box)

    # first evaluation sizes the device buffers; result kept for the accuracy check
    e_gpu, f_gpu = eng.execute(sysm.pos)

    # ---- timed region: positions resident in HBM, K evaluations ----
    eng.set_positions(sysm.pos)
    for _ in range(args.warmup):
        eng.compute()
    eng.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.compute()
    eng.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = args.steps / elapsed
    # per-pass times (HIP events between the passes, on the evaluation stream) from a second,
    # untimed run of the same K evaluations: the events cost ~5 % of an evaluation
    eng.profile_enable(True)
    for _ in range(args.steps):
        eng.compute()
    eng.synchronize()
    passes, n_prof = eng.profile_read()
    eng.profile_enable(False)
    stats = eng.stats()

    # ---- accuracy vs the FP64 oracle (same input) and algorithmic counts ----
    acc = {}
    n_frc = None
    if not args.no_accuracy:
        import oracle

        op = oracle.params_from_abi(p, nthreads=min(16, os.cpu_count() or 1), box=box)
        ref = oracle.compute(op, sysm.pos, sysm.charge, sysm.radius, want_stats=(kind == "gbsw"))
        acc = {"force_rel_rms": float(np.linalg.norm(f_gpu - ref.forces) / np.linalg.norm(ref.forces)),
               "energy_rel_err": float(abs(e_gpu - ref.energy) / abs(ref.energy)),
               "energy_kj_mol": e_gpu}
        if kind == "gbsw":
            n_frc = ref.stats["n_frc_pairs"]
            n_vol_exact = ref.stats["n_vol_exact"]
    if kind != "gbsw":
        # the chain-rule pass visits the list atoms of the points with a chain-rule term only
        # (saturated points have none): those pairs are its work
        n_frc = stats["n_frc_pairs"]
        n_vol_alg = stats["n_vol_pairs"]
    else:
        n_vol_alg = n_vol_exact if n_frc is not None else None

    # with a cutoff the pair pass skips blocks: its work is not the all-pairs count
    # (and the generated pair kernel's dual-number work is not the hand-written pass's count)
    kernel_flops = {} if p.nonbonded_method != N.NO_CUTOFF or w.get("expressions") else {
        "pairs": FL.pair_flops(n_atoms)}
    if n_vol_alg is not None:
        kernel_flops["volume"] = FL.volume_flops(kind, n_vol_alg, stats["n_points"])
    if n_frc is not None:
        kernel_flops["force"] = FL.force_flops(kind, n_frc)
    # every modelled pass against its roof; "roofline" is the one that takes the longest
    rooflines = []
    for k in ("volume", "force", "pairs"):
        if k not in kernel_flops or passes.get(k, 0.0) <= 0:
            continue
        ach = kernel_flops[k] / (passes[k] * 1e-3) / 1e12
        traffic = None
        tpath = os.path.join(ROOT, "profiles", f"traffic_{k}_{args.workload}.json")
        if os.path.exists(tpath):
            with open(tpath) as fh:
                traffic = json.load(fh).get("hbm_bytes_per_launch")
        # compute-bound against the published FP32 peak (the "mfma" class of the line's roofline
        # vocabulary); the per-pair math has no matrix shape, so it runs on the FP32 vector ALU
        rooflines.append({"bound": "mfma", "engine": "fp32 vector ALU (VALU), not matrix cores", "kernel": k,
                          "achieved": ach,
                          "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": ach / PEAK_FP32_TFLOPS,
                          "traffic": traffic, "alg_flops_per_launch": kernel_flops[k], "avg_ms": passes[k]})
    if passes.get("bin", 0.0) > 0:
        # binning moves bytes: per quadrature point its cell slot (a 4 B packed point, or a 16 B
        # record on grids binned by cells) or a 4 B volume (decided directly), plus the atom
        # positions; it is bound by its atomics' latency
        nq = stats["n_quad"] * n_atoms
        bbytes = FL.bin_bytes(stats["n_points"], nq, n_atoms, stats["slot_bytes"])
        ach = bbytes / (passes["bin"] * 1e-3) / 1e9
        traffic = None
        tpath = os.path.join(ROOT, "profiles", f"traffic_bin_{args.workload}.json")
        if os.path.exists(tpath):
            with open(tpath) as fh:
                traffic = json.load(fh).get("hbm_bytes_per_launch")
        rooflines.append({"bound": "hbm", "kernel": "bin", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": ach / PEAK_HBM_GBS, "traffic": traffic, "alg_bytes_per_launch": bbytes,
                          "avg_ms": passes["bin"]})
    roof = max(rooflines, key=lambda r: r["avg_ms"]) if rooflines else None

    # the whole evaluation against the FP32 peak: algorithmic FLOPs per evaluation (every
    # modelled pass plus the per-point and per-atom work) times evaluations per second
    eval_fp32 = None
    if n_vol_alg is not None and n_frc is not None and "pairs" in kernel_flops:
        f_alg = (kernel_flops["pairs"] + FL.volume_flops(kind, n_vol_alg, stats["n_quad"] * n_atoms)
                 + FL.force_flops(kind, n_frc) + FL.C_ATOM * n_atoms)
        eval_fp32 = {"alg_flops_per_eval": f_alg, "achieved_tflops": f_alg * value / 1e12,
                     "peak_tflops": PEAK_FP32_TFLOPS, "frac": f_alg * value / 1e12 / PEAK_FP32_TFLOPS,
                     "counts": "pairs N(N-1)/2, volume pairs of the lists, chain-rule pairs of the points "
                               "with a chain-rule term, every quadrature point, every atom"}

    # ---- CPU baseline: the oracle in the reference's loop structure, 1 thread ----
    cpu = None
    if args.cpu_baseline_evals > 0:
        import oracle

        op = oracle.params_from_abi(p, nthreads=1, dense=1, box=box)
        times = []
        for _ in range(args.cpu_baseline_evals):
            tc = time.perf_counter()
            oracle.compute(op, sysm.pos, sysm.charge, sysm.radius)
            times.append(time.perf_counter() - tc)
        dt = float(np.median(times))
        cpu = {"value": 1.0 / dt, "unit": "evals/s", "cores": 1, "kind": "port",
               "sample": f"median of {args.cpu_baseline_evals} full evaluation(s) of the same {n_atoms}-atom "
                         "system: FP64 oracle in the Reference platform's loop order (q-outer integral loop, "
                         "dense dI/dx arrays, N^2 chain rule), one thread",
               "seconds": dt, "seconds_each": times}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": (f"synthetic: seeded {n_atoms}-atom liquid filling a {box} nm box (synthetic.liquid seed {w['seed']})"
                 if box else
                 f"synthetic: seeded {n_atoms}-atom globule at 100 atoms/nm^3 (synthetic.globule seed {w['seed']})"),
        "config": {"workload": args.workload, "description": w["desc"], "n_atoms": n_atoms,
                   "quadrature": f"24x{w['n_lebedev']}", "precision": "FP32 pair/quadrature math, FP64 "
                   "integral and energy sums, int64 fixed-point forces"},
        **acc,
        "roofline": roof,
        "rooflines": rooflines,
        "eval_fp32": eval_fp32,
        "cpu_baseline": cpu,
        "passes_ms": passes,
        "profiled_launches": n_prof,
        "counts": {"n_points": stats["n_points"], "n_vol_pairs": stats["n_vol_pairs"],
                   "n_frc_pairs": n_frc, "n_active_points": stats["n_active"],
                   "n_chain_pairs_visited": stats["n_frc_pairs"], "n_cells": stats["n_cells"]},
    }
    eng.close()
    if args.workload == "gbmv2-5000" and not args.no_extras:
        # the other headline configurations, timed in the same run: the 25,000-atom GBSW
        # globule, and the 64-system mix batched against its per-system launches
        bres = measure_batch(max(4, args.steps // 2), 2)
        out["extra"] = {"gbsw-25000": measure_single("gbsw-25000", args.steps, args.warmup),
                        "batch64": {**batch_summary(bres), "n_atoms": int(np.sum(bres["sizes"])),
                                    "ms_per_round": 1e3 * bres["batch"]}}
    print(json.dumps(out), flush=True)
    return 0


def run_adjoint(args, w) -> int:
    """One step = energy, forces and the adjoint of all 8 conformations (one batched context,
    graph replay); value = conformation evaluations per second."""
    from openmmcustomgbintegralforce_amd import _native as N
    from openmmcustomgbintegralforce_amd import synthetic as S

    base = S.globule(w["n_atoms"], w["seed"])
    confs = S.conformations(base, 8, seed=w["seed"] + 1)
    p = N.default_params(N.MODEL_GBSW)
    p.n_lebedev = w["n_lebedev"]
    n = base.n
    pos = np.concatenate([c.pos + 40.0 * k for k, c in enumerate(confs)])
    eng = N.Engine.batch([p] * 8, [c.charge for c in confs], [c.radius for c in confs])
    out = eng.execute_adjoint(pos)
    eng.set_graph(True)
    eng.set_positions(pos)

    def timed(flags):
        for _ in range(args.warmup):
            eng.compute(flags)
        eng.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            eng.compute(flags)
        eng.synchronize()
        return time.perf_counter() - t0

    t_adj = timed(N.FORCES | N.ENERGY | N.ADJOINT)
    t_plain = timed(N.FORCES | N.ENERGY)
    eng.close()

    acc = {}
    if not args.no_accuracy:
        import oracle

        # size-independent identity on every conformation: E is homogeneous of degree 2 in q
        e = out["energy"]
        q_res = max(abs(float(np.dot(confs[k].charge, out["dEdq"][k * n:(k + 1) * n])) - 2.0 * e[k]) / abs(e[k])
                    for k in range(8))
        op = oracle.params_from_abi(p, nthreads=min(16, os.cpu_count() or 1))
        c0 = confs[0]
        ref = oracle.compute(op, c0.pos, c0.charge, c0.radius)
        # central differences of the FP64 oracle energy: two constants and four radii of conformation 0
        fd_err = 0.0
        for name, field in (("alpha0", "alpha0"), ("rscale", "rscale")):
            v = getattr(op, field)
            h = 1e-5 * abs(v)
            setattr(op, field, v + h)
            ep = oracle.energy(op, c0.pos, c0.charge, c0.radius)
            setattr(op, field, v - h)
            em = oracle.energy(op, c0.pos, c0.charge, c0.radius)
            setattr(op, field, v)
            fd = (ep - em) / (2 * h)
            fd_err = max(fd_err, abs(out["dEdtheta"][name][0] - fd) / abs(fd))
        atoms = np.random.default_rng(3).choice(n, 4, replace=False)
        fd_r = []
        for j in atoms:
            R = c0.radius.copy()
            R[j] += 1e-4
            ep = oracle.energy(op, c0.pos, c0.charge, R)
            R[j] -= 2e-4
            em = oracle.energy(op, c0.pos, c0.charge, R)
            fd_r.append((ep - em) / 2e-4)
        fd_r = np.array(fd_r)
        fd_err = max(fd_err, float(np.max(np.abs(out["dEdR"][atoms] - fd_r)) / np.max(np.abs(fd_r))))
        acc = {"force_rel_rms": float(np.linalg.norm(out["forces"][:n] - ref.forces) / np.linalg.norm(ref.forces)),
               "energy_rel_err": float(abs(e[0] - ref.energy) / abs(ref.energy)),
               "adjoint_fd_max_rel_err": fd_err, "adjoint_fd_checked": "dE/dalpha0, dE/drscale, dE/dR of 4 atoms",
               "charge_identity_max_rel_residual": q_res}
    value = 8 * args.steps / t_adj
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_adj / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": f"synthetic: seeded {n}-atom globule (synthetic.globule seed {w['seed']}) and 8 conformations "
                f"(synthetic.conformations seed {w['seed'] + 1})",
        "config": {"workload": args.workload, "description": w["desc"], "n_atoms": n, "conformations": 8,
                   "quadrature": "24x38"},
        **acc,
        "plain_evals_per_s": 8 * args.steps / t_plain,
        "adjoint_overhead": t_adj / t_plain,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)
    return 0


def measure_batch(steps, warmup, baselines=True, accuracy=False, seed=0):
    """The 64-system mix, batched (one context, graph replay) and, with `baselines`, the same
    systems as 64 contexts: enqueued back to back on one stream (sequential: device-resident
    positions, one synchronisation per round), on one stream each (concurrent), and evaluated
    one after another through gbi_execute (plain: positions in, energy and forces out per
    system).  Returns seconds per round of 64 system evaluations and the accuracy check."""
    from openmmcustomgbintegralforce_amd import _native as N
    from openmmcustomgbintegralforce_amd import synthetic as S

    sizes = S.batch_sizes(64, 300, 6000, seed=seed)
    systems = [S.globule(int(n), 1000 + k) for k, n in enumerate(sizes)]
    params = []
    for k in range(64):
        if k % 2 == 0:
            p = N.default_params(N.MODEL_GBSW)
        else:
            p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
            p.n_lebedev = 50
        params.append(p)
    shifted = [s.pos + 40.0 * k for k, s in enumerate(systems)]
    pos = np.concatenate(shifted)

    def timed(step_fn, sync_fn, n_steps=steps):
        for _ in range(warmup):
            step_fn()
        sync_fn()
        t0 = time.perf_counter()
        for _ in range(n_steps):
            step_fn()
        sync_fn()
        return (time.perf_counter() - t0) / n_steps

    res = {"sizes": sizes}
    eng = N.Engine.batch(params, [s.charge for s in systems], [s.radius for s in systems])
    e_b, f_b = eng.execute(pos)
    eng.set_graph(True)
    eng.set_positions(pos)
    eng.compute()
    eng.synchronize()
    res["batch"] = timed(eng.compute, eng.synchronize)
    eng.close()
    if baselines:
        engines = [N.Engine(p, s.charge, s.radius) for p, s in zip(params, systems)]
        for e1, x in zip(engines, shifted):
            e1.execute(x)
            e1.set_positions(x)

        def seq_step():
            for e1 in engines:
                e1.compute()

        def seq_sync():
            for e1 in engines:
                e1.synchronize()

        res["concurrent"] = timed(seq_step, seq_sync)
        shared = engines[0].device_view().stream
        for e1 in engines[1:]:
            e1.set_stream(shared)
        res["sequential"] = timed(seq_step, seq_sync)
        for e1 in engines:
            e1.set_stream(None)

        def plain_step():
            for e1, x in zip(engines, shifted):
                e1.execute(x)

        res["plain"] = timed(plain_step, lambda: None, max(1, steps // 4))
        for e1 in engines:
            e1.close()
    if accuracy:
        import oracle

        offs = np.concatenate([[0], np.cumsum(sizes)])
        worst_f = worst_e = 0.0
        checked = []
        for k in (0, 1, 2, 3):
            ref = oracle.compute(oracle.params_from_abi(params[k], nthreads=min(16, os.cpu_count() or 1)),
                                 systems[k].pos, systems[k].charge, systems[k].radius)
            f = f_b[offs[k]:offs[k + 1]]
            worst_f = max(worst_f, float(np.linalg.norm(f - ref.forces) / np.linalg.norm(ref.forces)))
            worst_e = max(worst_e, float(abs(e_b[k] - ref.energy) / abs(ref.energy)))
            checked.append(int(sizes[k]))
        res["accuracy"] = {"force_rel_rms": worst_f, "energy_rel_err": worst_e, "accuracy_systems": checked}
    return res


def batch_summary(res) -> dict:
    out = {"batch_evals_per_s": 64 / res["batch"]}
    if "sequential" in res:
        out.update({"sequential_evals_per_s": 64 / res["sequential"],
                    "concurrent_evals_per_s": 64 / res["concurrent"],
                    "plain_evals_per_s": 64 / res["plain"],
                    "batch_speedup": res["sequential"] / res["batch"],
                    "batch_speedup_vs_concurrent": res["concurrent"] / res["batch"],
                    "batch_speedup_vs_plain": res["plain"] / res["batch"]})
    return out


def run_batch(args, w) -> int:
    """One step = one evaluation of all 64 systems; value = system evaluations per second."""
    res = measure_batch(args.steps, args.warmup, baselines=not args.no_baseline, accuracy=not args.no_accuracy,
                        seed=w["seed"])
    sizes = res["sizes"]
    out = {
        "metric": METRIC,
        "value": 64 / res["batch"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * res["batch"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: 64 seeded globules at 100 atoms/nm^3, sizes synthetic.batch_sizes(64, 300, 6000, seed 0)",
        "config": {"workload": args.workload, "description": w["desc"], "n_atoms": int(np.sum(sizes)),
                   "systems": 64},
        **res.get("accuracy", {}),
        **batch_summary(res),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def measure_single(workload, steps, warmup) -> dict:
    """Evaluations per second of one more single-system workload (positions resident)."""
    from openmmcustomgbintegralforce_amd import _native as N
    from openmmcustomgbintegralforce_amd import synthetic as S

    w = WORKLOADS[workload]
    sysm = S.globule(w["n_atoms"], w["seed"])
    p = N.default_params(N.MODEL_GBSW) if w["model"] == "gbsw" else N.default_params(N.MODEL_GBMV, w["integral"])
    p.n_lebedev = w["n_lebedev"]
    eng = N.Engine(p, sysm.charge, sysm.radius)
    eng.execute(sysm.pos)
    eng.set_positions(sysm.pos)
    for _ in range(warmup):
        eng.compute()
    eng.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.compute()
    eng.synchronize()
    dt = (time.perf_counter() - t0) / steps
    eng.close()
    return {"evals_per_s": 1.0 / dt, "ms_per_eval": 1e3 * dt, "n_atoms": sysm.n, "steps": steps}


if __name__ == "__main__":
    sys.exit(main())
