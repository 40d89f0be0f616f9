"""Benchmark of the coupled surface time step on one MI355X.

One step = time_step!(OceanOnlyModel, Δt = 20 min) on the 1440×560×10 quarter-degree
workload: three split-RK3 WENO5 tracer stages + interpolation, similarity-theory flux
solve, freeze clamp, net-flux assembly and radiation, with every input resident in HBM.
Prints one JSON line (see the repository's benchmark contract).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "numericalearth.jl_amd"))

METRIC = "cell-updates/s & %HBM peak at 1440×560×10 (FP64 vs FP32); p99 latency batched"
HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "quarter_degree_1440x560x10": (1440, 560, 10),
    "cpu_reference_90x40x10": (90, 40, 10),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="quarter_degree_1440x560x10")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--dt", type=float, default=1200.0)
    ap.add_argument("--cpu-baseline", choices=("on", "off"), default="on")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU work of the baseline sample")
    return ap.parse_args()


def cpu_baseline(case, dt: float, target_s: float) -> dict:
    """The oracle (C restatement of the reference algorithm, OpenMP) on this host's cores."""
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle as O
    threads = min(16, os.cpu_count() or 1)
    O.set_threads(threads)
    om = O.OracleModel(case).reconcile()
    om.time_step(dt)                       # warm-up
    steps, t0 = 0, time.perf_counter()
    while True:
        om.time_step(dt)
        steps += 1
        el = time.perf_counter() - t0
        if el >= target_s or steps >= 50:
            break
    Nx, Ny, Nz = case.shape
    cpu_model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": Nx * Ny * Nz * steps / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} coupled steps of the same {Nx}x{Ny}x{Nz} workload after 1 warm-up step "
                      f"({el:.1f} s, {threads} OpenMP threads, {cpu_model})"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the multi-GPU layout is a design "
                                   "(latitude-band halo exchange), not an implementation", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    from numericalearth_hip.generators import make_case
    from numericalearth_hip.model import OceanOnlyModel, bytes_per_cell_update

    Nx, Ny, Nz = WORKLOADS[args.workload]
    case = make_case(Nx, Ny, Nz, seed=args.seed)
    model = OceanOnlyModel(case, precision=args.precision)
    for _ in range(args.warmup):
        model.time_step(args.dt)
    torch.cuda.synchronize()
    model.ctx.counters(reset=True)
    model.timing_begin(args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        model.time_step(args.dt)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kernel_ms, nrec = model.timing_end()
    counters = model.ctx.counters()

    N3 = Nx * Ny * Nz
    ms_per_step = 1e3 * elapsed / args.steps
    value = N3 * args.steps / elapsed
    s3 = 8 if args.precision == "fp64" else 4
    # RK stage kernels: stage 1 reads u, v, w, T, S and writes T, S (7 elements per cell);
    # stages 2-3 also read T⁰, S⁰ (9). Mean algorithmic bytes per launch:
    stage_bytes = N3 * s3 * (7 + 9 + 9) / 3.0
    stage_ms = (kernel_ms["rk_stage_1"] + kernel_ms["rk_stage_2"] + kernel_ms["rk_stage_3"]) / (3 * nrec)
    achieved = stage_bytes / (stage_ms * 1e-3) / 1e9
    B = bytes_per_cell_update(args.precision, Nz)
    step_gbs = B * N3 / (ms_per_step * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded generator of the 1440x560x10 configuration)",
        "config": {"workload": f"{args.workload} coupled step, dt={args.dt:g}s, precision={args.precision}",
                   "grid": [Nx, Ny, Nz], "seed": args.seed, "land_fraction": round(case.meta["land_fraction"], 4),
                   "bytes_per_cell_update": B, "step_hbm_fraction": step_gbs / HBM_PEAK_GBS},
        "roofline": {"kernel": "tracer_rk_stage (WENO5 tendency + top flux BC + RK stage, mean of 3 stages)",
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "kernels_ms_per_step": {k: v / nrec for k, v in kernel_ms.items()},
        "flux_solve": {"mean_iterations": counters["total_iterations"] / max(1, counters["solved_columns"]),
                       "max_iterations": counters["max_iterations"],
                       "nonconverged_columns": counters["nonconverged_columns"],
                       "nan_columns": counters["nan_columns"]},
    }
    model.close()
    if args.cpu_baseline == "on":
        out["cpu_baseline"] = cpu_baseline(case, args.dt, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
