#!/usr/bin/env python3
"""Throughput of the device legalizer on the 1M-cell mixed-height workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload mixed_1m|plain_20k]
                    [--no-cpu-baseline] [--cpu-sample-cells N] [--cpu-threads T]

One step = one whole legalization of the design on the GPU: lg_solve (usage
scatter, rho / lambda0 init, ADMM negotiation loop with partitioning, ordering,
primal and dual updates, final refine) followed by lg_postprocess (rotation),
starting from the preplaced positions that lg_create left in HBM.  Host import
and preplace run once before the timed region and are reported beside it.

Prints one JSON line: cells legalized per second over the K timed steps, the
roofline of the dominant kernel (the partition kernel, timed per launch with
HIP events on the solve's stream) and a CPU baseline from the oracle on a
bounded sample of the same recipe.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

WORKLOADS = {
    "mixed_1m": "1M mixed-height cells (86/10/3/1 % of 1/2/3/4-row cells, even heights under "
                "rail parity), 1600 rows x 12000 sites, 85 % utilisation, 3 fences, 200 macros, seed 0",
    "plain_20k": "20k single-row cells, 250 rows x 490 sites, widths 1-8 sites, 75 % utilisation, "
                 "4 Gaussian hotspots, 10 macros, seed 0",
    "density_1m": "electrostatic density: 1M cells, area LogNormal(ln 4, 0.5), half uniform / half in 8 "
                  "Gaussian clusters, square die at 70 % fill, 1024 x 1024 bins, split-BF16 Poisson, seed 0",
}
MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16


def traffic_from_profiles(workload: str):
    """HBM bytes per partition-kernel launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), or None."""
    path = os.path.join(HERE, "profiles", f"k_partition_{workload}_traffic.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"^bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(cells: int, threads: int):
    """The oracle legalizes a `cells`-cell instance of the mixed-height recipe
    end to end (import, preplace, ADMM, post) on `threads` OpenMP threads."""
    from admm_legalization_amd import synth
    from admm_legalization_amd.design import default_params
    from oracle import ref
    d = synth.mixed(cells, seed=1)
    p = default_params()
    p.num_cpu_threads = threads
    t = time.perf_counter()
    _, _, _, st = ref.run_design(d, p)
    el = time.perf_counter() - t
    if not st["converged"]:
        raise RuntimeError("oracle sample did not converge")
    return {"value": cells / el, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C++ restatement, OpenMP over partitions) legalizing a {cells}-cell "
                      f"instance of the mixed-height recipe ({d.meta['x_size']} sites x "
                      f"{d.meta['y_size']} rows, same height mix / utilisation / 3 fences, seed 1) "
                      f"end to end: {st['iterations']} iterations in {el:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="mixed_1m", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-cells", type=int, default=10000)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 256 MiB device-to-device three times (a kernel of known "
                         "HBM traffic for scaling PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the legalizer runs on a single device "
              "(no multi-GPU path is built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from admm_legalization_amd import lg, synth
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    lg.lib()  # raises if the HIP library is not built
    if args.workload == "density_1m":
        return bench_density(args)
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    t = time.perf_counter()
    design = synth.mixed_1m() if args.workload == "mixed_1m" else synth.plain_20k()
    t_gen = time.perf_counter() - t
    t = time.perf_counter()
    grid, cells, params = lg.device_inputs(design)
    t_pre = time.perf_counter() - t
    N = int(len(cells["w_grid"]))
    P = int(grid["x_size"]) * int(grid["y_size"])
    print(f"# workload {args.workload}: N={N} P={P} gen {t_gen:.2f}s host import+preplace {t_pre:.2f}s",
          file=sys.stderr, flush=True)

    torch.cuda.synchronize()
    t = time.perf_counter()
    dev = lg.DeviceLegalizer(grid, cells, params)
    torch.cuda.synchronize()
    t_upload = time.perf_counter() - t

    for w in range(args.warmup):
        dev.solve()
        dev.postprocess()
        print(f"# warmup {w} done", file=sys.stderr, flush=True)

    dev.profile_partition(True)
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        dev.solve()
        dev.postprocess()
        lat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms, launches = dev.profile_read()
    dev.profile_partition(False)

    vio = dev.check()
    st = dev.stats()
    if not st["converged"] or vio["total_ovf"] != 0 or vio["num_row_orient_vio"] != 0:
        print(f"# legalization not legal: {st} {vio}", file=sys.stderr)
        sys.exit(1)
    t_copy = time.perf_counter()
    dev.positions()
    t_copy = time.perf_counter() - t_copy

    # dominant kernel: the partition kernel's LDS class with the most time
    c = int(np.argmax(ms))
    avg_s = ms[c] / max(launches[c], 1) / 1e3
    bytes_per_launch = 12.0 * P + 24.0 * N
    achieved = bytes_per_launch / avg_s / 1e9
    value = N * args.steps / el
    out = {
        "metric": "cells legalized/s + p99 request latency, served vs plain, at ≤0.5% disp gap",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {WORKLOADS[args.workload]}",
            "mode": "plain (one request at a time)",
            "cells": N, "pixels": P, "partitions": int(dev_partitions(grid, params)),
            "admm_iterations": int(st["iterations"]),
            "total_disp_dbu": int(st["total_disp_dbu"]), "max_disp_dbu": int(st["max_disp_dbu"]),
            "request_latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                                   "p99": float(np.percentile(lat, 99) * 1e3)},
            "host_import_preplace_s": t_pre, "h2d_create_s": t_upload, "d2h_positions_s": t_copy,
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_partition<0,{c}> (ADMM primal + dual update, one wave per partition)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(args.workload),
            "avg_launch_us": avg_s * 1e6,
            "launches": int(launches[c]),
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "cpu_baseline": None,
    }
    dev.close()
    if not args.no_cpu_baseline:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample_cells, args.cpu_threads)
    print(json.dumps(out), flush=True)


def density_cpu_baseline(pl, g):
    """The FP64 oracle evaluates the same 1M-cell, 1024 x 1024 density once
    (scatter, overflow, closed-form Poisson, gather) with numpy on the host."""
    from oracle import ref_density as R
    geom = dict(num_bin_x=g.num_bin_x, num_bin_y=g.num_bin_y, bin_w=g.bin_w, bin_h=g.bin_h,
                bin_lx=g.bin_lx, bin_ux=g.bin_ux, bin_ly=g.bin_ly, bin_uy=g.bin_uy, die=g.die)
    t = time.perf_counter()
    mov, fil = R.scatter(geom, pl["cx"], pl["cy"], g.cell_dw, g.cell_dh, g.cell_scale)
    fixed = g.fixed_area.astype(np.float64).reshape(mov.shape)
    scaled = g.scaled_bin_area.astype(np.float64).reshape(mov.shape)
    dens, _ = R.overflow(mov, fil, fixed, scaled, g.sum_movable_area)
    _, ex, ey = R.poisson(dens)
    R.gather(geom, pl["cx"], pl["cy"], g.cell_dw, g.cell_dh, g.cell_scale, ex, ey)
    el = time.perf_counter() - t
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": 1.0 / el, "unit": "evaluations/s", "cores": threads, "kind": "port",
            "sample": f"one full evaluation by the FP64 numpy oracle (BLAS matrix products for the "
                      f"transforms) on the same 1M-cell, 1024x1024-bin input: {el:.2f} s"}


def bench_density(args):
    import torch
    from admm_legalization_amd import density as D
    from admm_legalization_amd import synth
    pl = synth.density_placement(1_000_000, seed=0)
    g = D.prepare(pl["die"], pl["dx"], pl["dy"])
    n_cells = len(pl["dx"])
    B = g.num_bin_x * g.num_bin_y
    print(f"# density_1m: N={n_cells} bins {g.num_bin_x}x{g.num_bin_y}", file=sys.stderr, flush=True)
    dg = D.DensityGradient(g, precision=D.PREC_SPLIT_BF16)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        cx = torch.tensor(pl["cx"], device="cuda")
        cy = torch.tensor(pl["cy"], device="cuda")
        gx, gy = torch.empty_like(cx), torch.empty_like(cx)
        for _ in range(args.warmup):
            dg.computeGrad(gx, gy, cx, cy)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record(s)
        for _ in range(args.steps):
            dg.computeGrad(gx, gy, cx, cy)
        e1.record(s)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        ms_eval = e0.elapsed_time(e1) / args.steps
        ovf = dg.overflow()
        # the Poisson solve alone, per precision, on the evaluation's density map
        rho = dg.getDeviceBinDensity()
        ps = D.PoissonSolver(g.num_bin_x, g.num_bin_y)
        psi, ex, ey = (torch.empty_like(rho) for _ in range(3))
        solve_ms = {}
        for name, prec in (("bf16", D.PREC_BF16), ("split_bf16", D.PREC_SPLIT_BF16), ("f32", D.PREC_F32)):
            for _ in range(2):
                ps.solvePoisson(rho, psi, ex, ey, precision=prec)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(s)
            reps = max(args.steps, 5)
            for _ in range(reps):
                ps.solvePoisson(rho, psi, ex, ey, precision=prec)
            b.record(s)
            torch.cuda.synchronize()
            solve_ms[name] = a.elapsed_time(b) / reps
        ps.close()
    n = g.num_bin_x
    flops = 16.0 * n ** 3  # four 2-D transforms as two n^3 GEMMs each
    achieved = flops / (solve_ms["split_bf16"] / 1e3) / 1e12
    out = {
        "metric": "density evaluations/s (scatter + overflow + Poisson + gather)",
        "value": 1e3 / ms_eval, "unit": "evaluations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3, "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "f32 (split-bf16 MFMA transforms)",
        "data": "synthetic",
        "config": {"workload": f"density_1m: {WORKLOADS['density_1m']}", "cells": n_cells,
                   "bins": [g.num_bin_x, g.num_bin_y], "overflow": ovf, "event_ms_per_eval": ms_eval,
                   "cells_per_s": n_cells * 1e3 / ms_eval, "poisson_ms": solve_ms,
                   "parallelism": "1 GPU"},
        "roofline": {"bound": "mfma", "kernel": "Poisson solve (split-BF16 cosine/sine GEMMs)",
                     "achieved": achieved, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / MFMA_BF16_PEAK_TFLOPS, "traffic": None,
                     "algorithmic_flops_per_solve": flops,
                     "note": "split-BF16 issues three MFMA products per algorithmic product"},
        "cpu_baseline": None,
    }
    dg.close()
    if not args.no_cpu_baseline:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = density_cpu_baseline(pl, g)
    print(json.dumps(out), flush=True)


def dev_partitions(grid, params):
    xs, ys = int(grid["x_size"]), int(grid["y_size"])
    xh, yh = int(params.x_hint), int(params.y_hint)
    return -(-xs // xh) * -(-ys // yh)


if __name__ == "__main__":
    main()
