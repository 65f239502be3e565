#!/usr/bin/env python3
"""Throughput of the device legalizer.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload served_stream|mixed_1m|plain_20k|density_1m|density_2m_4k|wl_2m]
                    [--requests R] [--plain-stride S] [--slots S] [--iters-per-sync I] [--load L]
                    [--no-e2e] [--no-cpu-baseline] [--cpu-threads T]

served_stream (default): the seeded 2000-request stream of legalizations
(ECO windows and sub-designs, 500-200k cells, mixed heights, 60-90 %
utilisation, 66.9M cells in all), prepared on the host first.
  * served (the reported value): every request queued at once to one
    continuous-batching request server; cells/s over the whole stream.  A
    step is one slice of R / K requests in arrival order; all slices are in
    flight in the same server, so ms_per_step = time / K.
  * plain: every S-th request (default 4: 500 requests, 16.8M cells) alone
    through lg_create .. lg_copy_out, one after another; each result must be
    bit-identical to its served result.
  * latency: the whole stream again with Poisson arrivals at --load of the
    served capacity (served p50 / p99), and a FIFO replay of the plain
    service times at the same fraction of plain capacity; then the plain
    leg's requests through a server with --latency-slots resident slots, at
    --load of its capacity on them.
  * the 1M-cell mixed-height design end to end through lg_run_design.
  * cpu_baseline: the oracle on every 400th request of the stream.
Every served result must be strictly legal, else the bench exits with an
error.

mixed_1m / plain_20k: one step = one whole legalization of the design on the
GPU (lg_solve + lg_postprocess from the preplaced positions lg_create left in
HBM); host import and preplace run once before the timed region.

Prints one JSON line: cells legalized per second, the roofline of the
dominant kernel (the partition kernel, timed with HIP events on its stream)
and a CPU baseline from the oracle.
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
    "served_stream": "stream of legalization requests (ECO windows below 20k cells, sub-designs with "
                     "3 fences from 20k) with cell counts log-uniform 500-200k, heights 86/10/3/1 % of "
                     "1/2/3/4 rows (even heights under rail parity), utilisation U(0.6, 0.9), "
                     "seed 0; served = continuous batching in one server (128 slots, 200 GB "
                     "device-memory admission budget) vs plain = one request at a time",
    "mixed_1m": "1M mixed-height cells (86/10/3/1 % of 1/2/3/4-row cells, even heights under "
                "rail parity), 1600 rows x 12000 sites, 85 % utilisation, 3 fences, 200 macros, seed 0",
    "plain_20k": "20k single-row cells, 250 rows x 490 sites, widths 1-8 sites, 75 % utilisation, "
                 "4 Gaussian hotspots, 10 macros, seed 0",
    "wl_2m": "weighted-average wirelength gradient (pin update fused): 2M cells (area LogNormal(ln 4, 0.5), "
             "positions half uniform / half clustered), 2.1M nets with degrees power-law on [2, 1000] "
             "(mean 3.8), pins at uniform offsets inside their cells, 0.5 % IO pins, seed 0",
    "density_1m": "electrostatic density: 1M cells, area LogNormal(ln 4, 0.5), half uniform / half in 8 "
                  "Gaussian clusters, square die at 70 % fill, 1024 x 1024 bins, split-BF16 Poisson, seed 0",
    "e2e_2m": "end to end: 2M mixed-height cells (86/10/3/1 % of 1/2/3/4 rows) at 90 % utilisation with 300 "
              "fixed macros and 2.1M nets (power-law degrees 2-1000, mean 3.8), placed from scratch by the "
              "global placer (Nesterov, WA wirelength + density on 4096 x 4096 bins, BF16 transforms) and "
              "legalized, seed 0",
    "density_2m_4k": "electrostatic density of the 2M-cell end-to-end design: 2M cells, area LogNormal(ln 4, "
                     "0.5), half uniform / half in 8 Gaussian clusters, 70 % fill, 4096 x 4096 bins, BF16 "
                     "Poisson, seed 0",
}
MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16


def traffic_from_profiles(workload: str, kernel: str = "k_partition"):
    """HBM bytes per partition-kernel launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), or None."""
    path = os.path.join(HERE, "profiles", f"{kernel}_{workload}_traffic.md")
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
    ap.add_argument("--workload", default="served_stream", choices=sorted(WORKLOADS))
    ap.add_argument("--requests", type=int, default=2000,
                    help="served_stream: the first N requests of the seeded 2000-request stream")
    ap.add_argument("--plain-stride", type=int, default=4,
                    help="served_stream: plain leg over every Nth request of the stream")
    ap.add_argument("--no-e2e", action="store_true",
                    help="served_stream: skip the 1M-cell end-to-end legalization")
    ap.add_argument("--slots", type=int, default=128, help="served_stream: resident requests")
    ap.add_argument("--iters-per-sync", type=int, default=16, help="served_stream: batch length")
    ap.add_argument("--load", type=float, default=0.7,
                    help="served_stream: Poisson arrival rate as a fraction of the measured capacity")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--short-cells", type=int, default=None,
                    help="served_stream: requests with fewer cells iterate in the server's short lane "
                         "(0: one lane; default: the server's)")
    ap.add_argument("--max-overtakes", type=int, default=None,
                    help="served_stream: smallest-first admission with this bound on how many later "
                         "requests may pass a queued one (0: arrival order; default: the server's)")
    ap.add_argument("--latency-slots", type=int, default=16,
                    help="served_stream: resident slots of the latency-oriented server run on the plain "
                         "leg's requests (0: skip)")
    ap.add_argument("--served-only", action="store_true",
                    help="served_stream: only the timed served run (for profiling); prints its figures")
    ap.add_argument("--latency-only", action="store_true",
                    help="served_stream: the timed served run and the Poisson-arrival latency run; "
                         "prints their figures (no plain leg)")
    ap.add_argument("--dump", default=None, help="served_stream: write per-request sizes, plain service "
                                                  "times and iterations to this JSON file")
    ap.add_argument("--graphs", action="store_true",
                    help="replay the ADMM iterations as HIP graphs instead of launching every kernel")
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
    if args.workload in ("density_1m", "density_2m_4k"):
        return bench_density(args)
    if args.workload == "wl_2m":
        return bench_wl(args)
    if args.workload == "e2e_2m":
        return bench_e2e(args)
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.workload == "served_stream":
        return bench_served(args)

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
    dev.use_graphs(args.graphs)
    torch.cuda.synchronize()
    t_upload = time.perf_counter() - t

    for w in range(args.warmup):
        dev.solve()
        dev.postprocess()
        print(f"# warmup {w} done", file=sys.stderr, flush=True)

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
    # partition-kernel timing from one more solve, outside the timed region
    # (HIP events around each iteration's partition launches; eager launches)
    dev.profile_partition(True)
    dev.solve()
    dev.postprocess()
    torch.cuda.synchronize()
    ms, launches = dev.profile_read()
    dev.profile_partition(False)

    vio = dev.check()
    st = dev.stats()
    from admm_legalization_amd.design import strictly_legal
    if not st["converged"] or not strictly_legal(vio):
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
            "hip_graphs": args.graphs,
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
            "kernel": ("k_partition<0,0,4> (ADMM primal + fused write-back/dual update, four waves per partition)"
                       if c == 0 else "k_partition<0,1,1> (ADMM primal, LDS sorts, one wave per partition)"),
            "note": "algorithmic bytes = 12 per pixel + 24 per cell per iteration; the dual update of "
                    "partitions without overflow runs beside this kernel on a second stream",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(args.workload),
            "hbm_measured_GBs": (traffic_from_profiles(args.workload) / avg_s / 1e9
                                 if traffic_from_profiles(args.workload) else None),
            "avg_launch_us": avg_s * 1e6,
            "launches": int(launches[c]),
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "cpu_baseline": None,
    }
    dev.close()
    # end to end as the reference's scripts time it (preplace included), in
    # a process whose GPU runtime is already up: host import + preplace,
    # lg_create + upload, solve, post-process, positions back
    torch.cuda.synchronize()
    t = time.perf_counter()
    g2, c2, p2 = lg.device_inputs(design)
    t_e2e_pre = time.perf_counter() - t
    dev2 = lg.DeviceLegalizer(g2, c2, p2)
    dev2.use_graphs(args.graphs)
    t_e2e_create = time.perf_counter() - t - t_e2e_pre
    dev2.solve()
    dev2.postprocess()
    dev2.positions()
    t_e2e = time.perf_counter() - t
    dev2.close()
    out["config"].update({
        "end_to_end_s": t_e2e, "end_to_end_cells_per_s": N / t_e2e,
        "end_to_end_parts_s": {"import_preplace": t_e2e_pre, "create_upload": t_e2e_create,
                               "solve_post_copy": t_e2e - t_e2e_pre - t_e2e_create},
        "end_to_end_note": "one more legalization after the timed ones, with the timed handle closed: host "
                           "import + preplace, lg_create (upload), solve, post-process, positions to the host; "
                           "h2d_create_s is the first create of the process and includes the GPU runtime's "
                           "one-time setup"})
    if not args.no_cpu_baseline:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample_cells, args.cpu_threads)
    print(json.dumps(out), flush=True)

def make_request(n, util, seed):
    from admm_legalization_amd import lg, synth
    return lg.device_inputs(synth.stream_request(n, util, seed))


def percentiles(xs):
    xs = np.asarray(xs, np.float64)
    return {"p50": float(np.percentile(xs, 50) * 1e3), "p99": float(np.percentile(xs, 99) * 1e3),
            "max": float(xs.max() * 1e3)}


def fifo_replay(arrivals, service):
    """Latencies of a single FIFO server (one request at a time) that takes
    service[i] seconds for request i arriving at arrivals[i]."""
    t, out = 0.0, []
    for a, s in zip(arrivals, service):
        t = max(t, a) + s
        out.append(t - a)
    return out


def host_cpu():
    """Model name and logical CPU count of this host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def served_cpu_baseline(spec, threads, stride=400):
    """The oracle legalizes a stratified sample of the stream (every
    `stride`-th request in arrival order: import, preplace, ADMM, post)
    serially, on `threads` OpenMP threads."""
    from admm_legalization_amd import synth
    from admm_legalization_amd.design import default_params
    from oracle import ref
    p = default_params()
    p.num_cpu_threads = threads
    cells, el, used = 0, 0.0, []
    for i in range(0, len(spec), stride):
        n, u, sd = spec[i]
        d = synth.stream_request(n, u, sd)
        t = time.perf_counter()
        _, _, _, st = ref.run_design(d, p)
        el += time.perf_counter() - t
        if not st["converged"]:
            raise RuntimeError("oracle sample did not converge")
        cells += n
        used.append(n)
    model, ncpu = host_cpu()
    return {"value": cells / el, "unit": "cells/s", "cores": threads, "kind": "port",
            "host_cpu": f"{model}, {ncpu} logical CPUs visible",
            "sample": f"oracle (C++ restatement, OpenMP over partitions and pixels) legalizing every "
                      f"{stride}th request of the stream ({len(used)} requests, {cells} cells: {used}) "
                      f"serially end to end in {el:.2f} s"}


def mixed_1m_end_to_end(repeats=3):
    """The 1M-cell mixed-height design legalized end to end through the
    drop-in entry point lg_run_design, as the reference's scripts time a
    legalization (run_ispd24.py:41-48: preplace + ADMM + post-process):
    import, pixel grid, preplace, upload, solve, post-process, positions to
    the host.  Median of `repeats` runs after one untimed run."""
    from admm_legalization_amd import lg, synth
    from admm_legalization_amd.design import strictly_legal
    d = synth.mixed_1m(0)
    lg.run_design(d)  # untimed: first-use costs of the runtime
    runs = []
    for _ in range(repeats):
        t = time.perf_counter()
        lx, ly, fl, st = lg.run_design(d)
        runs.append((time.perf_counter() - t, st))
    if not all(st["converged"] == 1 and strictly_legal(st["vio"]) for _, st in runs):
        print("# mixed_1m end to end: not legal", file=sys.stderr)
        sys.exit(1)
    runs.sort(key=lambda r: r[0])
    el, st = runs[len(runs) // 2]
    n = int(st["num_cells"])
    return {"cells": n, "wall_s": el, "cells_per_s": n / el, "runs_s": [r[0] for r in runs],
            "iterations": int(st["iterations"]),
            "parts_s": {"import": st["t_import_s"], "preplace": st["t_preplace_s"],
                        "upload_admm": st["t_admm_s"], "post": st["t_post_s"], "total": st["t_total_s"]},
            "under_1s": el < 1.0,
            "note": "lg_run_design wall time (median of 3 after one untimed run); inputs are the host "
                    "design (the reference's entry point takes its design from host memory), so the "
                    "PCIe upload is inside"}


def e2e_2m_end_to_end(num_cells=2_000_000, bins=4096, trace=False):
    """The end-to-end recipe on one GPU: the 2M-cell mixed-height
    design (90 % utilisation, 300 fixed macros, 2.1M nets of power-law degree
    2-1000, mean 3.8) placed from scratch by the global placer
    (gp_place_design: random initial placement, fillers, Nesterov over WA
    wirelength + density on 4096 x 4096 bins with BF16 transforms), then
    legalized by lg_run_design from the exported positions.  Wall time of
    both calls; the design starts in host memory."""
    from admm_legalization_amd import gp, lg, synth
    from admm_legalization_amd.design import strictly_legal
    t = time.perf_counter()
    d = synth.e2e(num_cells, seed=0)
    t_gen = time.perf_counter() - t
    print(f"# e2e_2m: {d.num_insts} instances, {len(d.net_start) - 1} nets, {int(d.net_start[-1])} pins, "
          f"generated in {t_gen:.1f}s", file=sys.stderr, flush=True)
    p = synth.e2e_placer_params(d, bins)
    t = time.perf_counter()
    lx, ly, gst = gp.place_design(d, p, trace_cap=4000 if trace else 0)
    t_gp = time.perf_counter() - t
    print(f"# e2e_2m global placement: {gst['iterations']} iterations, {gst['evaluations']} evaluations, "
          f"overflow {gst['overflow']:.4f}, {t_gp:.2f}s", file=sys.stderr, flush=True)
    placed = d.copy_positions(lx, ly)
    t = time.perf_counter()
    _, _, _, lst = lg.run_design(placed)
    t_lg = time.perf_counter() - t
    legal = lst["converged"] == 1 and strictly_legal(lst["vio"])
    print(f"# e2e_2m legalization: {lst['iterations']} iterations, {t_lg:.2f}s, legal {legal}", file=sys.stderr,
          flush=True)
    if not legal or gst["converged"] != 1:
        print(f"# e2e_2m failed: placer {gst} legalizer {lst}", file=sys.stderr)
        sys.exit(1)
    n = int(lst["num_cells"])
    out = {"cells": n, "nets": int(len(d.net_start) - 1), "pins": int(d.net_start[-1]), "bins": [bins, bins],
           "wall_s": t_gp + t_lg, "cells_per_s": n / (t_gp + t_lg),
           "global_place": {"wall_s": t_gp, "iterations": int(gst["iterations"]),
                            "evaluations": int(gst["evaluations"]), "fillers": int(gst["num_fillers"]),
                            "overflow": float(gst["overflow"]), "hpwl_initial": float(gst["hpwl_initial_db"]),
                            "hpwl": float(gst["hpwl_db"]), "init_s": float(gst["t_init_s"]),
                            "loop_s": float(gst["t_solve_s"]),
                            "ms_per_evaluation": float(gst["t_solve_s"]) / max(int(gst["evaluations"]), 1) * 1e3},
           "legalize": {"wall_s": t_lg, "iterations": int(lst["iterations"]), "avg_disp_um": lst["avg_disp_um"],
                        "max_disp_um": lst["max_disp_um"], "hpwl_before_um": lst["hpwl_before_um"],
                        "hpwl_after_um": lst["hpwl_after_um"], "delta_hpwl_pct": lst["delta_hpwl_pct"],
                        "strictly_legal": legal},
           "note": "gp_place_design + lg_run_design wall time, one run; the design starts in host memory"}
    if trace:
        out["trace"] = {k: v.tolist() for k, v in gst["trace"].items()}
    return out


def bench_e2e(args):
    """--workload e2e_2m: one JSON line for the end-to-end recipe."""
    import torch
    torch.cuda.synchronize()
    runs = [e2e_2m_end_to_end() for _ in range(max(args.steps, 1))]
    el = sum(r["wall_s"] for r in runs)
    r = runs[-1]
    out = {
        "metric": "cells legalized/s + p99 request latency, served vs plain, at ≤0.5% disp gap",
        "value": r["cells"] * len(runs) / el, "unit": "cells/s", "n_gpus": 1, "steps": len(runs),
        "warmup": 0, "ms_per_step": el / len(runs) * 1e3, "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"e2e_2m: {WORKLOADS['e2e_2m']}", "parallelism": "1 GPU", "e2e_2m": r},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_served(args):
    """Served (continuous batching) vs plain (one request at a time) on the
    2000-request stream."""
    import torch
    from concurrent.futures import ThreadPoolExecutor
    from admm_legalization_amd import lg, synth
    from admm_legalization_amd.design import strictly_legal
    from admm_legalization_amd.server import LegalizationServer
    spec = synth.stream_spec(args.requests)
    R = len(spec)
    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=8) as ex:
        reqs = list(ex.map(lambda r: make_request(*r), spec))
    t_prep = time.perf_counter() - t
    cells = np.array([len(c["w_grid"]) for _, c, _ in reqs], np.int64)
    total = int(cells.sum())
    K = max(args.steps, 1)
    print(f"# served_stream: {R} requests, {total} cells (min {cells.min()}, max {cells.max()}); "
          f"synth + host import/preplace {t_prep:.1f}s; {K} steps of {R / K:.0f} requests",
          file=sys.stderr, flush=True)

    srv_cfg = {}

    def served_run(idx, arrivals=None, prof=False, progress=True, slots=None):
        with LegalizationServer(max_slots=slots or args.slots, iters_per_sync=args.iters_per_sync,
                                use_graphs=args.graphs, short_cells=args.short_cells,
                                max_overtakes=args.max_overtakes) as srv:
            srv_cfg.update(short_cells=srv.config.short_cells, max_overtakes=srv.config.max_overtakes)
            if prof:
                srv.profile(True)
            tk = []
            t0 = time.perf_counter()
            for j, i in enumerate(idx):
                if arrivals is not None:
                    d = t0 + arrivals[j] - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                g, c, p = reqs[i]
                tk.append(srv.submit(g, c, p))
            res = []
            for j, k in enumerate(tk):
                res.append(srv.wait(k))
                if progress and (j + 1) % 250 == 0:
                    print(f"# served: {j + 1}/{len(tk)} collected at {time.perf_counter() - t0:.1f}s",
                          file=sys.stderr, flush=True)
            el = time.perf_counter() - t0
            return res, el, srv.counters(), (srv.profile_read() if prof else None)

    # -- warmup: W of the smallest requests through a server (untimed)
    small = [int(i) for i in np.argsort(cells)[:max(args.warmup, 0)]]
    if small:
        served_run(small, progress=False)
    torch.cuda.synchronize()

    # -- timed: the whole stream queued at t = 0 (capacity); step k is the
    # k-th slice of R / K requests in arrival order, all slices in flight in
    # one server
    res, t_served, ctr0, prof0 = served_run(list(range(R)), prof=True)
    torch.cuda.synchronize()
    served_rate = total / t_served
    print(f"# served (saturated): {t_served:.2f}s {served_rate:.0f} cells/s, peak slots "
          f"{ctr0['peak_slots']}", file=sys.stderr, flush=True)
    legal = [r["stats"]["converged"] == 1 and strictly_legal(r["stats"]["vio"]) for r in res]
    if not all(legal):
        print(f"# served results not legal: {R - sum(legal)} of {R}", file=sys.stderr)
        sys.exit(1)
    if args.served_only:
        it_us = prof0["ms"] / max(prof0["iterations"], 1) * 1e3
        print(json.dumps({"workload": f"served_stream: {R} requests, served leg only", "cells": total,
                          "served_s": t_served, "served_cells_per_s": served_rate,
                          "partition_us_per_batched_iteration": it_us,
                          "batched_iterations": prof0["iterations"],
                          "algorithmic_GBs": prof0["bytes"] / (prof0["ms_union"] / 1e3) / 1e9,
                          "algorithmic_GBs_per_launch": prof0["bytes"] / (prof0["ms"] / 1e3) / 1e9}),
              flush=True)
        return

    def served_latency():
        """The whole stream again, arriving as a Poisson process at `load` of
        the measured capacity: (rate, the unit-mean gaps, latencies, latency
        percentiles per size class, queueing percentiles)."""
        rng = np.random.default_rng(7)
        gaps = rng.exponential(1.0, R)
        lam = args.load * R / t_served
        arr = np.cumsum(gaps) / lam
        arr -= arr[0]
        rp, _, _, _ = served_run(list(range(R)), arrivals=arr)
        lat = [r["latency_s"] for r in rp]
        by_size = {}
        for lo, hi in ((0, 5000), (5000, 20000), (20000, 80000), (80000, 10**9)):
            xs = [lat[i] for i in range(R) if lo <= cells[i] < hi]
            if xs:
                by_size[f"{lo}-{hi if hi < 10**9 else 'max'}"] = percentiles(xs)
        return lam, gaps, lat, by_size, percentiles([r["queued_s"] for r in rp])

    if args.latency_only:
        lam_s, _, lat_served, by_size, queued = served_latency()
        print(json.dumps({"workload": f"served_stream: {R} requests, served capacity and latency",
                          "short_cells": srv_cfg["short_cells"], "max_overtakes": srv_cfg["max_overtakes"],
                          "served_s": t_served, "served_cells_per_s": served_rate,
                          "load": args.load, "arrival_rate_per_s": lam_s,
                          "latency_ms": percentiles(lat_served), "by_cells": by_size,
                          "queued_ms": queued}), flush=True)
        return

    # -- plain: every plain_stride-th request alone (lg_create .. lg_copy_out)
    def plain_one(g, c, p):
        dev = lg.DeviceLegalizer(g, c, p)
        dev.use_graphs(args.graphs)
        dev.solve()
        dev.postprocess()
        lx, ly, rot = dev.positions()
        st = dev.stats()
        dev.close()
        return lx, ly, rot, st

    pidx = list(range(0, R, args.plain_stride))
    plain, service = [], []
    t0 = time.perf_counter()
    for j, i in enumerate(pidx):
        ts = time.perf_counter()
        plain.append(plain_one(*reqs[i]))
        service.append(time.perf_counter() - ts)
        if (j + 1) % 100 == 0:
            print(f"# plain {j + 1}/{len(pidx)} {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    t_plain = time.perf_counter() - t0
    plain_cells = int(cells[pidx].sum())
    plain_rate = plain_cells / t_plain
    print(f"# plain: {len(pidx)} requests {t_plain:.2f}s {plain_rate:.0f} cells/s", file=sys.stderr, flush=True)
    if args.dump:
        json.dump({"cells": cells[pidx].tolist(), "service_s": service,
                   "iterations": [int(st["iterations"]) for _, _, _, st in plain]}, open(args.dump, "w"))
    same, disp_gap = 0, 0.0
    for (lx, ly, rot, st), i in zip(plain, pidx):
        r = res[i]
        same += int(np.array_equal(lx, r["lx"]) and np.array_equal(ly, r["ly"]) and np.array_equal(rot, r["rot"]))
        if st["total_disp_dbu"]:
            disp_gap = max(disp_gap, abs(r["stats"]["total_disp_dbu"] - st["total_disp_dbu"]) / st["total_disp_dbu"])
    if same != len(pidx):
        print(f"# served results differ from plain ({same}/{len(pidx)} identical)", file=sys.stderr)
        sys.exit(1)

    # -- latency under Poisson arrivals at `load` of each system's capacity
    lam_s, gaps, lat_served, by_size, queued = served_latency()
    lam_p = args.load * len(pidx) / t_plain
    gp = gaps[:len(pidx)]
    arr_p = np.cumsum(gp) / lam_p
    arr_p -= arr_p[0]
    lat_plain = fifo_replay(arr_p, service)
    print(f"# poisson: served p99 {np.percentile(lat_served, 99) * 1e3:.1f} ms, plain p99 "
          f"{np.percentile(lat_plain, 99) * 1e3:.1f} ms", file=sys.stderr, flush=True)

    # -- the plain leg's requests through a server with fewer resident slots
    # (latency-oriented): its capacity on those requests, then the same
    # Poisson arrival sequence as the plain replay at `load` of that capacity
    lat_cfg = None
    if args.latency_slots > 0:
        _, t_sub, _, _ = served_run(pidx, slots=args.latency_slots, progress=False)
        lam_l = args.load * len(pidx) / t_sub
        arr_l = np.cumsum(gp) / lam_l
        arr_l -= arr_l[0]
        rl, _, _, _ = served_run(pidx, arrivals=arr_l, slots=args.latency_slots, progress=False)
        lat_l = [r["latency_s"] for r in rl]
        lat_cfg = {"slots": args.latency_slots, "requests": len(pidx),
                   "capacity_cells_per_s": plain_cells / t_sub,
                   "capacity_vs_plain": plain_cells / t_sub / plain_rate,
                   "arrival_rate_per_s": lam_l, "served": percentiles(lat_l),
                   "plain": percentiles(lat_plain),
                   "note": "the plain leg's requests; served with max_slots = slots (the other server "
                           "settings as above) at `load` of its capacity on them, plain = the FIFO replay "
                           "at `load` of plain capacity, both under the same unit Poisson sequence"}
        print(f"# latency config ({args.latency_slots} slots): capacity {plain_cells / t_sub:.0f} cells/s, "
              f"p99 {np.percentile(lat_l, 99) * 1e3:.1f} ms", file=sys.stderr, flush=True)

    e2e = None
    if not args.no_e2e:
        print("# mixed_1m end to end ...", file=sys.stderr, flush=True)
        e2e = mixed_1m_end_to_end()
        print(f"# mixed_1m end to end: {e2e['wall_s']:.3f}s", file=sys.stderr, flush=True)

    # the two lanes' partition launches overlap: the rate is taken over the
    # time during which any of them ran (per launch: over their summed time)
    achieved = (prof0["bytes"] / (prof0["ms_union"] / 1e3) / 1e9
                if prof0 and prof0["ms_union"] > 0 else None)
    per_launch = prof0["bytes"] / (prof0["ms"] / 1e3) / 1e9 if prof0 and prof0["ms"] > 0 else None
    traffic = traffic_from_profiles("served_stream", "kb_partition")
    out = {
        "metric": "cells legalized/s + p99 request latency, served vs plain, at ≤0.5% disp gap",
        "value": served_rate,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": t_served / K * 1e3,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"served_stream: {R} of 2000 requests; {WORKLOADS['served_stream']}",
            "mode": "served (continuous batching); value = capacity with the whole stream queued at t=0; "
                    f"a step is one slice of {R / K:.0f} requests in arrival order (all slices in flight "
                    "in one server)",
            "requests": R, "cells": total,
            "served_cells_per_s": served_rate, "served_s": t_served,
            "strictly_legal": f"{sum(legal)}/{R}",
            "plain": {"requests": len(pidx), "stride": args.plain_stride, "cells": plain_cells,
                      "cells_per_s": plain_rate, "seconds": t_plain,
                      "note": f"every {args.plain_stride}th request of the stream (arrival order, so "
                              "unbiased in size) legalized alone through lg_create .. lg_copy_out"},
            "served_vs_plain": served_rate / plain_rate,
            "identical_results": f"{same}/{len(pidx)} plain-solved requests bit-identical to their served result",
            "disp_gap": disp_gap,
            "latency_ms": {
                "load": args.load,
                "served": percentiles(lat_served),
                "served_by_cells": by_size,
                "served_queued": queued,
                "served_arrival_rate_per_s": lam_s,
                "plain": percentiles(lat_plain),
                "plain_arrival_rate_per_s": lam_p,
                "plain_note": "FIFO replay of the measured per-request plain service times under Poisson "
                              "arrivals at the same fraction of plain capacity",
                "served_saturated": percentiles([r["latency_s"] for r in res]),
                "latency_config": lat_cfg,
            },
            "server": {"slots": args.slots, "iters_per_sync": args.iters_per_sync,
                       "short_cells": srv_cfg["short_cells"], "max_overtakes": srv_cfg["max_overtakes"],
                       "peak_slots": ctr0["peak_slots"], "batches": ctr0["batches"],
                       "batched_iterations": ctr0["iterations"],
                       "pool_reserved_bytes": ctr0["pool_reserved_bytes"],
                       "hip_graphs": args.graphs},
            "host_synth_import_preplace_s": t_prep,
            "mixed_1m_end_to_end": e2e,
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "kb_partition<0,0,4> + <0,1,1> (batched ADMM primal + fused dual update over all "
                      "resident requests)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None,
            "achieved_per_launch": per_launch,
            "partition_ms_union": prof0["ms_union"] if prof0 else None,
            "partition_ms_sum": prof0["ms"] if prof0 else None,
            "traffic": traffic,
            "hbm_measured_GBs": (traffic / (prof0["ms"] / max(prof0["iterations"], 1) / 1e3) / 1e9
                                 if prof0 and traffic else None),
            "avg_iteration_us": prof0["ms"] / max(prof0["iterations"], 1) * 1e3 if prof0 else None,
            "iterations": prof0["iterations"] if prof0 else None,
            "algorithmic_bytes_per_iteration": prof0["bytes"] / max(prof0["iterations"], 1) if prof0 else None,
            "note": "algorithmic bytes = 12 per pixel + 24 per cell for every ADMM iteration a request "
                    "executed; time = HIP events around the partition-kernel launches of each batched "
                    "iteration on its lane's stream, over the timed run; achieved = bytes over the time "
                    "during which any lane's partition launches ran (the two lanes' launches overlap), "
                    "achieved_per_launch = bytes over the summed launch time; avg_iteration_us = one "
                    "lane iteration's partition launches; traffic / hbm_measured_GBs = "
                    "the partition kernel's own HBM bytes (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                    "copy-calibrated) -- it skips partitions without overflow, whose duals the dual "
                    "pass updates",
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = served_cpu_baseline(synth.stream_spec(2000), args.cpu_threads)
    print(json.dumps(out), flush=True)


def bench_wl(args):
    """One WA wirelength gradient evaluation (pin update + gradient) per step."""
    import torch
    from admm_legalization_amd import synth
    from admm_legalization_amd.wirelength import WireLengthGradient
    pl = synth.density_placement(2_000_000, seed=0)
    nl = synth.netlist(pl["dx"], pl["dy"], 2_100_000, die=pl["die"], seed=1)
    n_cells, n_nets, n_pins = len(pl["dx"]), len(nl["net_start"]) - 1, len(nl["pin2cell"])
    print(f"# wl_2m: cells {n_cells} nets {n_nets} pins {n_pins}", file=sys.stderr, flush=True)
    wl = WireLengthGradient(n_cells, nl["net_start"], nl["pin2cell"], nl["pin_offset_x"], nl["pin_offset_y"],
                            nl["pin_x"], nl["pin_y"], nl["net_weight"])
    die_w = pl["die"][2] - pl["die"][0]
    wl.setGammaInv(8.0 / (0.01 * die_w))  # gamma of 1/8 of a 1 % bin width scale
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        cx = torch.tensor(pl["cx"], device="cuda")
        cy = torch.tensor(pl["cy"], device="cuda")
        gx, gy = torch.empty_like(cx), torch.empty_like(cx)
        for _ in range(max(args.warmup, 1)):
            wl.computeGradAt(cx, cy, gx, gy)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = max(args.steps, 20)
        t0 = time.perf_counter()
        e0.record(s)
        for _ in range(reps):
            wl.computeGradAt(cx, cy, gx, gy)
        e1.record(s)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        ms = e0.elapsed_time(e1) / reps
    deg = np.diff(nl["net_start"])
    # algorithmic bytes: per pin 4 (pin2cell) + 8 (offset) + 8 (cell centre) + 8 (pin write)
    # + 8 (term write) + 4 (cell index) + 8 (term read); per net 8; per cell 8 (out) + 4 (CSR)
    bytes_eval = 48.0 * n_pins + 8.0 * n_nets + 12.0 * n_cells
    achieved = bytes_eval / (ms / 1e3) / 1e9
    out = {
        "metric": "WA wirelength gradient evaluations/s (pin update + gradient)",
        "value": 1e3 / ms, "unit": "evaluations/s", "n_gpus": 1, "steps": reps, "warmup": args.warmup,
        "ms_per_step": el / reps * 1e3, "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"wl_2m: {WORKLOADS['wl_2m']}", "cells": n_cells, "nets": n_nets, "pins": n_pins,
                   "nets_above_64_pins": int((deg > 64).sum()), "event_ms_per_eval": ms,
                   "pins_per_s": n_pins * 1e3 / ms, "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "kernel": "k_wl_chunk + k_wl_net_big + k_wl_cell (one evaluation)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes_per_eval": bytes_eval},
        "cpu_baseline": None,
    }
    wl.close()
    if not args.no_cpu_baseline:
        from oracle import ref_wl as W
        k = 200_000  # the oracle on the first 200k nets (numpy, FP64)
        m = int(nl["net_start"][k])
        ns = nl["net_start"][:k + 1]
        t = time.perf_counter()
        px, py = W.update_pins(nl["pin_x"][:m], nl["pin_y"][:m], nl["pin2cell"][:m], nl["pin_offset_x"][:m],
                               nl["pin_offset_y"][:m], pl["cx"], pl["cy"])
        W.wa_grad(ns, px, py, nl["pin2cell"][:m], 1.0, nl["net_weight"][:k], n_cells)
        el = time.perf_counter() - t
        out["cpu_baseline"] = {"value": m / el, "unit": "pins/s", "cores": 1, "kind": "port",
                               "sample": f"FP64 numpy oracle on the first {k} nets ({m} pins) in {el:.2f} s; "
                                         f"GPU rate {n_pins * 1e3 / ms:.3g} pins/s"}
    print(json.dumps(out), flush=True)


def density_cpu_baseline(pl, g):
    """The FP64 oracle evaluates the same density once (scatter, overflow,
    closed-form Poisson, gather) with numpy on the host."""
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
                      f"transforms) on the same {len(pl['cx'])}-cell, {g.num_bin_x}x{g.num_bin_y}-bin "
                      f"input: {el:.2f} s"}


def bench_density(args):
    import torch
    from admm_legalization_amd import density as D
    from admm_legalization_amd import synth
    big = args.workload == "density_2m_4k"
    pl = synth.density_placement(2_000_000 if big else 1_000_000, seed=0)
    g = D.prepare(pl["die"], pl["dx"], pl["dy"], num_bins=(4096, 4096) if big else None)
    n_cells = len(pl["dx"])
    B = g.num_bin_x * g.num_bin_y
    prec_name, prec_main = ("bf16", D.PREC_BF16) if big else ("split_bf16", D.PREC_SPLIT_BF16)
    print(f"# {args.workload}: N={n_cells} bins {g.num_bin_x}x{g.num_bin_y}", file=sys.stderr, flush=True)
    dg = D.DensityGradient(g, precision=prec_main)
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
        solve_ms, solve_tiles = {}, {}
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
            tile, tune_ms = ps.getTile(prec)
            solve_tiles[name] = {"tile": D.TILE_NAMES[tile],
                                 "tuner_ms": [round(m, 4) for m in tune_ms]}
        ps.close()
    n = g.num_bin_x
    flops = 16.0 * n ** 3  # four 2-D transforms as two n^3 GEMMs each
    achieved = flops / (solve_ms[prec_name] / 1e3) / 1e12
    out = {
        "metric": "density evaluations/s (scatter + overflow + Poisson + gather)",
        "value": 1e3 / ms_eval, "unit": "evaluations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3, "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": f"f32 ({prec_name.replace('_', '-')} MFMA transforms)",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {WORKLOADS[args.workload]}", "cells": n_cells,
                   "bins": [g.num_bin_x, g.num_bin_y], "overflow": ovf, "event_ms_per_eval": ms_eval,
                   "cells_per_s": n_cells * 1e3 / ms_eval, "poisson_ms": solve_ms,
                   "poisson_tiles": solve_tiles,
                   "parallelism": "1 GPU"},
        "roofline": {"bound": "mfma",
                     "kernel": f"Poisson solve ({prec_name.replace('_', '-')} cosine/sine GEMMs, k_gemm_bf on "
                               "bf16 hi/lo operands)",
                     "achieved": achieved, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / MFMA_BF16_PEAK_TFLOPS, "traffic": None,
                     "algorithmic_flops_per_solve": flops,
                     "note": ("split-BF16 issues three MFMA products per algorithmic product" if not big else
                              "16 n^3 flops: eight n^3-MAC products (the kernel issues seven: the potential's "
                              "and E_y's last products share one stacked GEMM)")},
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
