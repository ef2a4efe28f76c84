#!/usr/bin/env python3
"""Benchmark: converged voxel-increments/s of the EVP-FFT increment loop on one MI355X.

Workload: periodic Voronoi polycrystal, 128^3 voxels, 1000 grains, uniform
random orientations, FCC Cu, x3 tension at 1e-3 1/s with stress-free lateral
faces.  A "step" is one loading increment iterated to the deck's Err (2e-5):
all outer fixed-point iterations (FFT -> Green operator -> inverse FFT ->
per-voxel Newton solve -> error norms) plus the end-of-increment hardening.

The increment is TimeStep 0.1 s (0.01% strain).  At 0.1% per increment the
elastic predictor of the first increment puts the slip-rate exponent near 150
and the reference's plain Newton iteration overflows; 0.1 s is the TimeStep of
the reference's own deck.

Both precisions run the same increments:
  fp64  -- every pass in FP64
  mixed -- FFT and Green passes in FP32 (mean stress removed before the cast),
           Newton solve, hardening, DisGrad accumulation and residuals in FP64.
`value` is the mixed-precision throughput; the FP64 figure, the relative
macroscopic-stress difference between the two, per-kernel times, roofline
fractions and a CPU baseline (the FP64 CPU oracle on a bounded sample of the
same workload) are reported beside it.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 (vector and matrix), TFLOP/s
F_NR = 2300.0                  # flops per voxel per Newton iteration (algorithmic count)
# HBM bytes per voxel per launch of the Newton kernel (after the first of an increment), from
# rocprofv3 FETCH_SIZE (x2.0) and WRITE_SIZE (x1.0) runs calibrated on a 1 GiB device copy
# (profiles/evp_fft_128cubed_pmc.md); mixed precision reads the increment's Newton matrix and
# per-system constants as FP32
NR_TRAFFIC_BYTES_PER_VOXEL = {"fp64": 817.2, "mixed": 609.4}
METRIC = "converged voxel-increments/s at 128³, plain FP64 vs mixed-precision at equal stress error"


def fft_green_bytes_per_voxel(real_bytes: int, n2: int, c_out: int = 9) -> float:
    """Algorithmic HBM bytes per voxel of one FFT + Green + inverse FFT + DisGrad update pass:
    FP64 stress in (6 x 8 B), DisGrad read+write (c_out x 16 B), and the half-spectrum
    passes: 6 comps x 4 touches forward, c_out comps x 4 touches inverse, each 2*s*h bytes."""
    h = (n2 // 2 + 1) / n2
    return 48.0 + 16.0 * c_out + (24.0 + 4.0 * c_out) * 2.0 * real_bytes * h


def run_precision(deck: str, precision: int, steps: int, warmup: int, timing: int, keep_fields: bool = False,
                  warm_start: int = 0, blob=None) -> dict:
    """One solver run: `warmup` untimed increments, then `steps` timed ones (wall clock).
    timing=0: no events, every batch of outer iterations replayed from its HIP graph (the
    headline run); timing=2: HIP events around each Newton-kernel launch on the solver's stream
    (the roofline run); timing=1: events around every kernel (the per-kernel breakdown run).
    Batches with events are enqueued launch by launch."""
    import srx_amd
    s = srx_amd.Solver(deck, precision=precision, reserved=[timing], warm_start=warm_start)
    if blob is not None:
        s.mlp_set(blob)
    for _ in range(warmup):
        s.step()
    kt0 = s.kernel_times()
    reps = []
    t0 = time.perf_counter()
    for _ in range(steps):
        reps.append(s.step())
    t1 = time.perf_counter()
    kt1 = s.kernel_times()
    kt = {k: (kt1[k][0] - kt0.get(k, (0, 0))[0], kt1[k][1] - kt0.get(k, (0, 0))[1]) for k in kt1}
    sig = np.array(reps[-1]["sig_avg"])
    fields = {}
    if keep_fields:   # outside the timed region
        s9 = s.field("sig").reshape(-1, 3, 3)
        dev = s9 - np.trace(s9, axis1=1, axis2=2)[:, None, None] * np.eye(3) / 3
        fields = {"vm": np.sqrt(1.5 * (dev * dev).sum(axis=(1, 2))), "rho_s": s.field("rho_s")}
    s.close()
    return {"seconds": t1 - t0, "reports": reps, "kernels": kt, "sig_avg": sig, "fields": fields,
            "iters": sum(r["iters"] for r in reps), "nr_total": sum(r["nr_iters_total"] for r in reps),
            "nr_wave": sum(r["nr_wave_iters"] for r in reps)}


def cpu_baseline(n: int, grains: int, seed: int, increments: int) -> dict | None:
    """The FP64 CPU oracle on a bounded sample: the first `increments` increments of a
    `n`^3, `grains`-grain Voronoi RVE with the same material, loading and Err, each iterated to
    convergence and all of them timed; the rate is voxels x increments / seconds."""
    try:
        from oracle import oracle
        from srx_amd import inputs
        oracle.build()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        work = tempfile.mkdtemp(prefix="srx_cpu_")
        deck = inputs.make_case(work, n=(n, n, n), grains=grains, seed=seed, nsteps=increments, timestep=0.1)
        o = oracle.Oracle(deck, nthreads=threads)
        nvox = n ** 3
        secs, iters, nr = [], 0, 0
        for k in range(increments):
            t0 = time.perf_counter()
            r = o.step()
            secs.append(time.perf_counter() - t0)
            if not r["converged"]:
                raise RuntimeError("oracle increment %d did not converge" % k)
            iters += r["iters"]
            nr += r["nr_iters_total"]
            print("bench: cpu oracle increment %d: %d outer iterations, %.1f s" % (k, r["iters"], secs[-1]),
                  file=sys.stderr, flush=True)
        o.close()
    except Exception as exc:  # the baseline is reported, never required
        return {"error": str(exc)}
    total = sum(secs)
    return {"value": nvox * increments / total, "unit": "voxel-increments/s", "cores": threads, "kind": "port",
            "sample": "CPU oracle (C, OpenMP, FP64) run from the start on a %d^3, %d-grain Voronoi RVE of the same "
                      "material, loading and Err: increments 0-%d, each iterated to convergence and all timed "
                      "(%.1f s, %d outer iterations, %.2f Newton iterations per voxel per outer iteration); "
                      "rate = voxels x increments / seconds" % (n, grains, increments - 1, total, iters,
                                                               nr / nvox / max(iters, 1)),
            "seconds": total, "outer_iters": iters}


def ensemble_bench(args) -> int:
    """RVE ensemble: members of uneven size, each with its own grain count, texture and strain rate
    (0.01% strain per increment), run by srx_run_ensemble `concurrency` at a time and once one at
    a time.  Prints one JSON line; not the headline metric."""
    import srx_amd
    from srx_amd import inputs
    rng = np.random.default_rng(args.seed)
    work = args.workdir or tempfile.mkdtemp(prefix="srx_ens_")
    decks, nvox = [], []
    for spec in args.ensemble.split(","):
        count, n = (int(x) for x in spec.split("x"))
        for _ in range(count):
            k = len(decks)
            grains = int(rng.integers(20, 1001))
            rate = float(10 ** rng.uniform(-4, -2))
            tex = "random" if rng.random() < 0.5 else "cube"
            d = os.path.join(work, "rve%02d_%d" % (k, n))
            deck = os.path.join(d, "deck.in")
            if not os.path.exists(deck):
                inputs.make_case(d, n=(n, n, n), grains=grains, seed=args.seed + k, nsteps=args.steps,
                                 timestep=1e-4 / rate, rate=rate, texture=tex)
            decks.append(deck)
            nvox.append(n ** 3)
    print("bench: ensemble of %d RVEs, %d voxels, in %s" % (len(decks), sum(nvox), work), file=sys.stderr, flush=True)
    prec = srx_amd.PREC_MIXED if args.precisions.split(",")[-1] == "mixed" else srx_amd.PREC_FP64
    out = {}
    # 0: all members in lockstep, their per-voxel kernels in one launch each; then `concurrency`
    # members at a time on their own streams; then one at a time
    for conc in (0, -2, -4, args.concurrency, 1):
        t0 = time.perf_counter()
        res = srx_amd.run_ensemble(decks, concurrency=conc, precision=prec, nsteps_override=args.steps)
        wall = time.perf_counter() - t0
        work_vi = sum(r["voxels"] * r["increments"] for r in res)
        run = res[0]["run_phase_seconds"]
        out[conc] = {"run_phase_seconds": run, "value": work_vi / run, "voxel_increments": work_vi,
                     "wall_seconds_incl_setup": wall, "value_incl_setup": work_vi / wall,
                     "setup_seconds_sum": sum(r["setup_seconds"] for r in res),
                     "run_seconds_sum": sum(r["run_seconds"] for r in res)}
        print("bench: %s: increment loop %.2f s (%.3g voxel-increments/s), %.2f s wall with setup" % (
            "lockstep, %d group(s)" % max(1, -conc) if conc <= 0 else "concurrency %d" % conc, run, work_vi / run, wall),
            file=sys.stderr, flush=True)
    best = max(out, key=lambda k: out[k]["value"])
    line = {"metric": "converged voxel-increments/s, RVE ensemble (increment loop, setup excluded)",
            "value": out[best]["value"], "unit": "voxel-increments/s", "n_gpus": 1,
            "steps": args.steps, "higher_is_better": True, "data": "synthetic",
            "dtype": "f64+f32" if prec == srx_amd.PREC_MIXED else "f64",
            "config": {"workload": "RVE ensemble %s (count x edge), grains U{20..1000}, random or cube texture, "
                                   "rates log-U[1e-4, 1e-2]/s, 0.01%% strain per increment" % args.ensemble,
                       "concurrency": best,
                       "mode": ("lockstep, %d group(s)" % max(1, -best)) if best <= 0 else "streams"},
            "lockstep_1_group": out[0], "lockstep_2_groups": out[-2], "lockstep_4_groups": out[-4],
            "sequential": out[1], "concurrent": out[args.concurrency],
            "concurrency_speedup": out[args.concurrency]["value"] / out[1]["value"],
            "lockstep_speedup": max(out[0]["value"], out[-2]["value"], out[-4]["value"]) / out[1]["value"]}
    print(json.dumps(line))
    return 0


def learned_bench(args) -> int:
    """Learned Newton start: the network is trained on the GPU on states recorded from a 64^3,
    125-grain run (extrapolated start), then the bench workload runs with the plain (reference),
    extrapolated and learned starts.  Prints one JSON line; not the headline metric."""
    import srx_amd
    from srx_amd import inputs
    work = args.workdir or tempfile.mkdtemp(prefix="srx_mlp_")
    tdeck = os.path.join(work, "train64", "deck.in")
    if not os.path.exists(tdeck):
        inputs.make_case(os.path.dirname(tdeck), n=(64, 64, 64), grains=125, seed=args.seed + 1,
                         nsteps=args.train_increments + 1, timestep=0.1)
    t0 = time.perf_counter()
    s = srx_amd.Solver(tdeck, precision=srx_amd.PREC_MIXED, warm_start=1)
    s.mlp_record(4, 4000000)
    for _ in range(args.train_increments):
        s.step()
    x, t = s.mlp_samples()
    s.close()
    t1 = time.perf_counter()
    blob, hist = srx_amd.mlp_train(x, t, epochs=20)
    t2 = time.perf_counter()
    print("bench: %d training samples recorded in %.1f s, trained in %.1f s (loss %.4f -> %.4f)" % (
        len(x), t1 - t0, t2 - t1, hist[0], hist[-1]), file=sys.stderr, flush=True)
    n = (args.n, args.n, args.n)
    nvox = n[0] * n[1] * n[2]
    deck = os.path.join(work, "eval%d" % args.n, "deck.in")
    if not os.path.exists(deck):
        inputs.make_case(os.path.dirname(deck), n=n, grains=args.grains, seed=args.seed,
                         nsteps=args.warmup + args.steps + 1, timestep=0.1)
    out = {}
    for name, ws in (("plain", 0), ("extrapolated", 1), ("learned", 2)):
        r = run_precision(deck, srx_amd.PREC_MIXED, args.steps, args.warmup, timing=2, warm_start=ws,
                          blob=blob if ws == 2 else None)
        out[name] = {"seconds": r["seconds"], "value": args.steps * nvox / r["seconds"],
                     "outer_iters": r["iters"], "nr_iters_per_voxel_per_iter": r["nr_total"] / nvox / r["iters"],
                     "nr_wave_iters_per_iter": r["nr_wave"] / r["iters"], "nr_kernel_ms": r["kernels"]["nr"][0],
                     "sig_avg": list(r["sig_avg"])}
        print("bench: %s start: %.3f s, %.3f Newton iterations/voxel/outer iteration, Newton kernel %.1f ms" % (
            name, r["seconds"], out[name]["nr_iters_per_voxel_per_iter"], r["kernels"]["nr"][0]),
            file=sys.stderr, flush=True)
    P, L = out["plain"], out["learned"]
    ref = np.array(P["sig_avg"])
    for v in out.values():
        v["rel_macro_stress_diff_vs_plain"] = float(np.abs(np.array(v.pop("sig_avg")) - ref).max() /
                                                    np.abs(ref).max())
    line = {"metric": "converged voxel-increments/s with the learned Newton start at %d^3" % args.n,
            "value": L["value"], "unit": "voxel-increments/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "higher_is_better": True, "dtype": "f64+f32+bf16", "data": "synthetic",
            "config": {"workload": "EVP-FFT increment loop, %d^3 periodic Voronoi, %d grains, mixed precision, "
                                   "TimeStep 0.1 s, Err 2e-5; network 24-64-64-6 bf16 trained on %d states of "
                                   "a 64^3 125-grain run" % (args.n, args.grains, len(x))},
            "train": {"samples": len(x), "record_seconds": t1 - t0, "train_seconds": t2 - t1, "epochs": 20,
                      "loss_first_epoch": float(hist[0]), "loss_last_epoch": float(hist[-1])},
            "plain": P, "extrapolated": out["extrapolated"], "learned": L,
            "speedup_vs_plain": L["value"] / P["value"],
            "speedup_vs_extrapolated": L["value"] / out["extrapolated"]["value"],
            "nr_iterations_saved_vs_plain": 1.0 - L["nr_iters_per_voxel_per_iter"] / P["nr_iters_per_voxel_per_iter"],
            "nr_wave_iterations_saved_vs_plain": 1.0 - L["nr_wave_iters_per_iter"] / P["nr_wave_iters_per_iter"]}
    print(json.dumps(line))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed increments")
    ap.add_argument("--warmup", type=int, default=5,
                    help="untimed increments before the timed ones (the texture starts to rotate at increment 3)")
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--grains", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--precisions", default="fp64,mixed")
    ap.add_argument("--cpu-n", type=int, default=64, help="edge of the CPU-baseline RVE")
    ap.add_argument("--cpu-grains", type=int, default=125)
    ap.add_argument("--cpu-increments", type=int, default=4, help="converged increments the CPU baseline times")
    ap.add_argument("--from-start", type=int, default=50,
                    help="also run this many increments from istep 0 in each precision (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workdir", default=None)
    ap.add_argument("--ensemble", default=None,
                    help="run an RVE ensemble instead, e.g. 16x32,16x64,12x96,4x128 (count x edge)")
    ap.add_argument("--concurrency", type=int, default=8, help="ensemble members run at once")
    ap.add_argument("--learned", action="store_true",
                    help="train the learned Newton start and compare plain, extrapolated and learned starts")
    ap.add_argument("--train-increments", type=int, default=4, help="increments recorded for training")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the whole grid is resident on one MI355X; "
              "ensembles would shard by RVE); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        return 2

    if args.ensemble:
        return ensemble_bench(args)
    if args.learned:
        return learned_bench(args)
    from srx_amd import inputs
    import srx_amd

    work = args.workdir or tempfile.mkdtemp(prefix="srx_bench_")
    n = (args.n, args.n, args.n)
    nvox = n[0] * n[1] * n[2]
    deck = os.path.join(work, "deck.in")
    nsteps_deck = max(args.warmup + args.steps, args.from_start) + 1
    if not os.path.exists(deck):
        inputs.make_case(work, n=n, grains=args.grains, seed=args.seed, nsteps=nsteps_deck, timestep=0.1)
    print("bench: inputs in %s" % work, file=sys.stderr, flush=True)

    res, nre, brk, fs = {}, {}, {}, {}
    for name in args.precisions.split(","):
        prec = srx_amd.PREC_MIXED if name == "mixed" else srx_amd.PREC_FP64
        print("bench: %s run (%d warmup + %d timed increments, graph replay)" % (name, args.warmup, args.steps),
              file=sys.stderr, flush=True)
        res[name] = run_precision(deck, prec, args.steps, args.warmup, timing=0, keep_fields=True)
        r = res[name]
        print("bench: %s %.3f s, %d outer iterations" % (name, r["seconds"], r["iters"]), file=sys.stderr, flush=True)
        nre[name] = run_precision(deck, prec, args.steps, args.warmup, timing=2)
        e = nre[name]
        print("bench: %s Newton-event run %.3f s, %d outer iterations, Newton kernel %.1f ms in %d launches" % (
            name, e["seconds"], e["iters"], e["kernels"]["nr"][0], e["kernels"]["nr"][1]), file=sys.stderr, flush=True)
        brk[name] = run_precision(deck, prec, args.steps, args.warmup, timing=1)
        b = brk[name]
        print("bench: %s breakdown run %.3f s, kernels %s" % (
            name, b["seconds"], {k: round(v[0], 1) for k, v in b["kernels"].items()}), file=sys.stderr, flush=True)
        if args.from_start > 0:
            fs[name] = run_precision(deck, prec, args.from_start, 0, timing=0, keep_fields=True)
            print("bench: %s %d increments from the start: %.3f s, %d outer iterations" % (
                name, args.from_start, fs[name]["seconds"], fs[name]["iters"]), file=sys.stderr, flush=True)

    head = "mixed" if "mixed" in res else "fp64"
    H = res[head]
    value = args.steps * nvox / H["seconds"]
    out = {"metric": METRIC, "value": value, "unit": "voxel-increments/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1000.0 * H["seconds"] / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None,
           "dtype": "f64+f32" if head == "mixed" else "f64", "data": "synthetic",
           "config": {"workload": "EVP-FFT increment loop, %d^3 periodic Voronoi, %d grains, FCC Cu, "
                                  "x3 tension 1e-3/s, TimeStep 0.1 s, Err 2e-5, increments %d-%d" % (
                                      args.n, args.grains, args.warmup, args.warmup + args.steps - 1),
                      "n": args.n, "grains": args.grains, "precision": head, "timestep": 0.1, "err": 2e-5,
                      "execution": "outer-iteration batches replayed from HIP graphs, no events"}}

    def stress_errors(a, b):
        fa, fb = a["fields"], b["fields"]
        return {"rel_macro_stress_error": float(np.linalg.norm(a["sig_avg"] - b["sig_avg"]) / np.linalg.norm(b["sig_avg"])),
                "max_voxel_von_mises_error_rel": float(np.abs(fa["vm"] - fb["vm"]).max() / fb["vm"].max()),
                "max_slip_system_rho_s_error_rel": float(np.abs(fa["rho_s"] - fb["rho_s"]).max()
                                                         / np.abs(fb["rho_s"]).max())}

    per = {}
    for name, r in res.items():
        kt = brk[name]["kernels"]
        e = nre[name]
        fg_names = ("fwd_z", "fwd_y", "green_x", "inv_y", "inv_z")
        fg_ms = sum(kt[k][0] for k in fg_names if k in kt)
        b_iters = max(brk[name]["iters"], 1)
        nr_ms = e["kernels"].get("nr", (0, 0))[0]
        rb = 4 if name == "mixed" else 8
        b_fg = fft_green_bytes_per_voxel(rb, n[2]) * nvox
        per[name] = {
            "value": args.steps * nvox / r["seconds"], "seconds": r["seconds"],
            "newton_event_run_seconds": e["seconds"], "breakdown_run_seconds": brk[name]["seconds"],
            "outer_iters_per_increment": r["iters"] / args.steps,
            "nr_iters_per_voxel_per_iter": r["nr_total"] / nvox / max(r["iters"], 1),
            "nr_wavefront_divergence": 64.0 * r["nr_wave"] / max(r["nr_total"], 1),
            "kernel_ms": {k: v[0] for k, v in kt.items()},
            "fft_green_gbs": b_fg * b_iters / (fg_ms * 1e-3) / 1e9 if fg_ms else None,
            "fft_green_frac_hbm": (b_fg * b_iters / (fg_ms * 1e-3) / 1e9) / HBM_PEAK_GBS if fg_ms else None,
            "nr_tflops": F_NR * e["nr_total"] / (nr_ms * 1e-3) / 1e12 if nr_ms else None,
            "fft_green_bytes_per_voxel": fft_green_bytes_per_voxel(rb, n[2]),
        }
        if name in fs:
            f = fs[name]
            per[name]["from_start"] = {"increments": args.from_start, "seconds": f["seconds"],
                                       "value": args.from_start * nvox / f["seconds"],
                                       "outer_iters_per_increment": f["iters"] / args.from_start}
    out["fp64"] = per.get("fp64")
    out["mixed"] = per.get("mixed")
    if "fp64" in res and "mixed" in res:
        out["mixed_vs_fp64"] = {"speedup": per["mixed"]["value"] / per["fp64"]["value"],
                                **stress_errors(res["mixed"], res["fp64"])}
        if fs:
            out["mixed_vs_fp64_from_start"] = {
                "increments": args.from_start,
                "speedup": per["mixed"]["from_start"]["value"] / per["fp64"]["from_start"]["value"],
                **stress_errors(fs["mixed"], fs["fp64"])}
    # roofline of the dominant kernel (largest share in the breakdown run); the Newton kernel is
    # timed with HIP events on the solver's stream in the Newton-event run over the same increments
    bk = brk[head]["kernels"]
    dom = max(bk, key=lambda k: bk[k][0])
    if dom == "nr":
        E = nre[head]
        kt = E["kernels"]
        ach = F_NR * E["nr_total"] / (kt["nr"][0] * 1e-3) / 1e12
        out["roofline"] = {"kernel": "nr", "bound": "fp64_valu", "peak_kind": "FP64 vector ALU (no MFMA in this kernel)",
                           "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                           "frac": ach / FP64_PEAK_TFLOPS, "traffic": NR_TRAFFIC_BYTES_PER_VOXEL[head] * nvox,
                           "traffic_unit": "HBM bytes per launch (rocprofv3 PMC, profiles/)",
                           "algorithmic": "%.0f flop per voxel per Newton iteration x %.3f Newton iterations per voxel "
                                          "per launch" % (F_NR, E["nr_total"] / nvox / max(E["iters"], 1))}
    else:
        kt = bk
        rb = 4 if head == "mixed" else 8
        h = (n[2] // 2 + 1) / n[2]
        pass_bytes = {"fwd_z": (48 + 6 * 2 * rb * h), "fwd_y": 6 * 2 * 2 * rb * h, "green_x": 15 * 2 * rb * h,
                      "inv_y": 9 * 2 * 2 * rb * h, "inv_z": 9 * 2 * rb * h + 144}
        bpv = pass_bytes.get(dom, 0.0)
        ach = bpv * nvox * kt[dom][1] / (kt[dom][0] * 1e-3) / 1e9
        out["roofline"] = {"kernel": dom, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": ach / HBM_PEAK_GBS, "traffic": None,
                           "algorithmic": "%.1f B per voxel per launch" % bpv}
    try:   # the achievable HBM rate of a plain streaming copy, beside the FFT passes' rate
        hs = srx_amd.hbm_stream(1 << 30, 5)
        out["hbm_copy_ceiling"] = hs
        for name in per:
            if per[name].get("fft_green_gbs"):
                per[name]["fft_green_frac_of_copy_ceiling"] = per[name]["fft_green_gbs"] / hs["best_gbs"]
    except Exception as exc:
        out["hbm_copy_ceiling"] = {"error": str(exc)}
    out["roofline"]["avg_launch_ms"] = kt[dom][0] / max(kt[dom][1], 1)
    out["roofline"]["launches"] = kt[dom][1]
    if not args.no_cpu_baseline:
        print("bench: cpu baseline sample", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_grains, args.seed, args.cpu_increments)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
