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
FP64_VALU_PEAK_TFLOPS = 78.6   # MI355X dense FP64 vector peak
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
    ap.add_argument("--timestepper", choices=("SplitRungeKutta3", "QuasiAdamsBashforth2"), default="SplitRungeKutta3",
                    help="the ocean's tracer time stepper (ocean_simulation's timestepper keyword; QAB2 is FP64)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="quarter_degree_1440x560x10")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--dt", type=float, default=1200.0)
    ap.add_argument("--cpu-baseline", choices=("on", "off"), default="on")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU work of the baseline sample")
    ap.add_argument("--mixed-report", action="store_true",
                    help="run the FP64 and mixed forms side by side for --steps steps and report their differences")
    ap.add_argument("--ensemble", type=int, default=0, metavar="N",
                    help="ensemble report: N uneven members stepped as one batch (autotuned, graph) "
                         "against the same members stepped one after another")
    ap.add_argument("--serve", type=str, default="",
                    help="serving report at these request rates (req/s, comma-separated)")
    ap.add_argument("--serve-seconds", type=float, default=4.0, help="arrival window of each serving run")
    ap.add_argument("--serve-steps", type=int, default=100, help="time steps per request")
    ap.add_argument("--serve-budget", type=float, default=200e9, help="HBM admission budget (bytes)")
    ap.add_argument("--serve-cells", type=float, default=32e6, help="work budget: resident cells")
    ap.add_argument("--serve-pool", type=int, default=32, help="distinct pre-generated request inputs")
    ap.add_argument("--serve-plain-max", type=int, default=150, help="requests run on the plain path")
    ap.add_argument("--stream-report", action="store_true",
                    help="step time with the atmosphere and radiation streamed into rings of snapshot slots "
                         "against the same series held whole")
    ap.add_argument("--stream-every", type=int, default=9, help="steps between snapshots (9: 3-hourly at 20 min)")
    ap.add_argument("--stream-window", type=int, default=3, help="snapshot slots per streamed series")
    ap.add_argument("--stream-snapshots", type=int, default=16, help="snapshots in the (cyclical) series")
    ap.add_argument("--extras", choices=("on", "off"), default="on",
                    help="also measure the other precision form, the 64-member ensemble and one serving rate")
    ap.add_argument("--extras-rate", type=float, default=200.0, help="request rate of the serving measurement")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="before the timed steps, launch a kernel of known traffic (5 arrays read, 2 written, "
                         "each the size of one stage array) three times, for memory counters to be read against")
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="only run a device-to-device copy of this many GiB (--steps times) and report its bandwidth")
    return ap.parse_args()


def traffic_probe(model, array_bytes: int, launches: int = 3):
    """ne_stream_probe_f64 with 5 inputs and 2 outputs of `array_bytes` each: a launch moves exactly
    7·array_bytes of HBM (5 read, 2 written), the reference for reading memory counters."""
    import ctypes as C
    import torch
    n = array_bytes // 8
    arrays = [torch.rand(n, dtype=torch.float64, device="cuda") for _ in range(7)]
    ins = (C.c_void_p * 5)(*[a.data_ptr() for a in arrays[:5]])
    outs = (C.c_void_p * 2)(*[a.data_ptr() for a in arrays[5:]])
    for _ in range(launches):
        r = model.lib.ne_stream_probe_f64(model.ctx.handle, ins, 5, outs, 2, n, None)
        if r:
            raise RuntimeError(f"ne_stream_probe_f64 failed with ne_status {r}")
    torch.cuda.synchronize()
    expect = sum(arrays[:5])
    assert torch.equal(arrays[5], expect) and torch.equal(arrays[6], expect + 1.0)
    del arrays


def measured_traffic(workload: str, precision: str):
    """HBM bytes per launch of the RK stage kernel from the committed memory-counter summary: the
    measured read requests at the 128 bytes per request that ne_stream_probe_f64 (a kernel of known
    traffic, profiled in the same runs) establishes, plus the kernel's stores; or None."""
    import re
    f = REPO / "profiles" / "tracer_rk_stage_traffic_1440x560x10.md"
    if workload != "quarter_degree_1440x560x10" or not f.exists():
        return None
    m = re.search(rf"traffic_bytes_per_launch\[{precision}\] = ([0-9.eE+-]+)", f.read_text())
    return float(m.group(1)) if m else None


def measured_valu(workload: str, precision: str):
    """Issued FP64 lane-flops per flux-solve launch and the VALU-busy fraction of the flux-solve
    and stage kernels, from the committed SQ counter summary, or None."""
    import re
    f = REPO / "profiles" / "surface_update_and_stage_valu_1440x560x10.md"
    if workload != "quarter_degree_1440x560x10" or not f.exists():
        return None
    text = f.read_text()
    out = {}
    for key in ("fp64_lane_flops_per_launch", "valu_busy_surface_update", "valu_busy_rk_stage"):
        m = re.search(rf"{key}\[{precision}\] = ([0-9.eE+-]+)", text)
        if m:
            out[key] = float(m.group(1))
    return out or None


def copy_bandwidth(gib: float, reps: int) -> dict:
    """Device-to-device copy (torch's copy kernel) of `gib` GiB: read + write bytes per second."""
    import torch
    n = int(gib * (1 << 30)) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return {"bytes_per_copy": 2 * n * 8, "ms_per_copy": ms, "gbs": 2 * n * 8 / (ms * 1e-3) / 1e9}


def ensemble_measure(n: int, seed: int, precision: str, steps: int, warmup: int, dt: float) -> dict:
    """Throughput of an ensemble stepped as one batch against its members stepped one after another."""
    import torch
    from numericalearth_hip.ensemble import Ensemble, build_members, draw_ensemble

    specs = draw_ensemble(n, seed)
    members = build_members(specs, precision=precision)
    cells = sum(s.cells for s in specs)
    torch.cuda.synchronize()

    def timed(step, k):
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / k

    def sequential():
        for m in members:
            m.time_step(dt)

    t_seq = timed(sequential, steps)
    ens = Ensemble(members)
    default = ens.tuning
    t_default = timed(lambda: ens.time_step(dt), steps)
    chosen = ens.autotune()
    t_tuned = timed(lambda: ens.time_step(dt), steps)
    ens.use_graph(True)
    t_graph = timed(lambda: ens.time_step(dt), steps)
    info = ens.info()
    ens.close()
    for m in members:
        m.close()
    shapes = [(s.Nx, s.Ny, s.Nz) for s in specs]
    return {"report": "ensemble", "members": len(specs), "cells": cells, "precision": precision,
           "steps": steps, "dt": dt,
           "smallest": min(shapes, key=lambda s: s[0] * s[1] * s[2]),
           "largest": max(shapes, key=lambda s: s[0] * s[1] * s[2]),
           "sequential_ms_per_step": 1e3 * t_seq, "sequential_cell_updates_per_s": cells / t_seq,
           "batched_default_ms_per_step": 1e3 * t_default, "batched_tuned_ms_per_step": 1e3 * t_tuned,
           "batched_graph_ms_per_step": 1e3 * t_graph, "batched_graph_cell_updates_per_s": cells / t_graph,
           "speedup_graph_over_sequential": t_seq / t_graph, "default_tuning": default, "autotuned": chosen,
           "work_items": info}


def ensemble_report(args):
    print(json.dumps(ensemble_measure(args.ensemble, args.seed, args.precision, args.steps, args.warmup, args.dt)),
          flush=True)


def serving_measure(args, rates) -> dict:
    """Continuous-batching server against the plain per-request path under Poisson arrivals.

    Requests are 100-step solves whose grids are drawn log-uniform between 1×1×50 and
    720×280×10 (regional, 1/2°), FP64 or mixed 50/50, from a pool of pre-generated inputs
    (the client's data), admitted under a 200 GB HBM budget."""
    import numpy as np
    import torch
    from numericalearth_hip.ensemble import draw_ensemble, member_case, member_params
    from numericalearth_hip.serving import (SURFACE_RESULT_FIELDS, Request, Server, latency_summary,
                                            poisson_arrivals, serve_plain)

    rng = np.random.default_rng(args.seed)
    pool = draw_ensemble(args.serve_pool, args.seed, smallest=(1, 1, 50), largest=(720, 280, 10), resolution=0.5)
    cases = [(member_case(s), member_params(s)) for s in pool]
    out = {"report": "serving", "steps_per_request": args.serve_steps, "dt": args.dt,
           "budget_bytes": args.serve_budget, "cell_budget": args.serve_cells, "duration_s": args.serve_seconds, "pool": len(pool), "rates": {}}

    def make_requests(rate):
        arr = poisson_arrivals(rate, args.serve_seconds, rng)
        reqs = []
        for k, t in enumerate(arr):
            c, p = cases[int(rng.integers(len(cases)))]
            reqs.append(Request(id=k, case=c, params=p, precision="fp64" if rng.uniform() < 0.5 else "mixed",
                                steps=args.serve_steps, dt=args.dt, arrival=float(t)))
        return reqs

    # warm-up: one small request of each precision through both paths
    warm = [Request(id=-1, case=cases[0][0], params=cases[0][1], precision=p, steps=2) for p in ("fp64", "mixed")]
    srv = Server(budget_bytes=args.serve_budget, cell_budget=args.serve_cells,
                     result_fields=SURFACE_RESULT_FIELDS)
    srv.run(warm)
    srv.close()
    serve_plain([Request(id=-1, case=cases[0][0], params=cases[0][1], precision="fp64", steps=2)])
    torch.cuda.synchronize()
    for rate in rates:
        reqs = make_requests(rate)
        srv = Server(budget_bytes=args.serve_budget, cell_budget=args.serve_cells,
                     result_fields=SURFACE_RESULT_FIELDS)
        done = srv.run(reqs)
        srv.close()
        batched = latency_summary(done)
        batched["batch_steps"] = srv.steps_issued
        print(f"# rate {rate}: batched {batched}", file=sys.stderr, flush=True)
        # the plain path on the same arrival stream, cut to the first requests it serves in ~30 s
        plain_in = [Request(id=r.id, case=r.case, params=r.params, precision=r.precision, steps=r.steps,
                            dt=r.dt, arrival=r.arrival) for r in reqs[:args.serve_plain_max]]
        plain = latency_summary(serve_plain(plain_in, result_fields=SURFACE_RESULT_FIELDS))
        print(f"# rate {rate}: plain {plain}", file=sys.stderr, flush=True)
        out["rates"][str(rate)] = {"offered_requests": len(reqs), "batched": batched, "plain": plain,
                                   "throughput_ratio": batched["cell_updates_per_s"] / plain["cell_updates_per_s"]}
    return out


def serving_report(args):
    print(json.dumps(serving_measure(args, [float(x) for x in args.serve.split(",")])), flush=True)


def form_measure(workload: str, precision: str, seed: int, steps: int, warmup: int, dt: float,
                 timestepper: str = "SplitRungeKutta3") -> dict:
    """ms per step of one precision form (and timestepper) on the workload (inputs resident in HBM)."""
    import torch
    from numericalearth_hip.generators import make_case
    from numericalearth_hip.model import OceanOnlyModel

    Nx, Ny, Nz = WORKLOADS[workload]
    m = OceanOnlyModel(make_case(Nx, Ny, Nz, seed=seed), precision=precision, timestepper=timestepper)
    for _ in range(warmup):
        m.time_step(dt)
    m.timing_begin(steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        m.time_step(dt)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kernel_ms, nrec = m.timing_end()
    m.close()
    return {"precision": precision, "timestepper": timestepper, "steps": steps, "ms_per_step": 1e3 * el / steps,
            "cell_updates_per_s": Nx * Ny * Nz * steps / el,
            "kernels_ms_per_step": {k: v / nrec for k, v in kernel_ms.items()}}


def mixed_report(args):
    """SST, surface heat flux and heat-budget differences of the mixed form against the FP64 form
    (which matches the FP64 oracle to 1e-12 per step) after args.steps steps."""
    import numpy as np
    import torch
    from numericalearth_hip.generators import make_case
    from numericalearth_hip.model import OceanOnlyModel

    Nx, Ny, Nz = WORKLOADS[args.workload]
    case = make_case(Nx, Ny, Nz, seed=args.seed)
    m64 = OceanOnlyModel(case, precision="fp64")
    m32 = OceanOnlyModel(case, precision="mixed")
    rho_c = m64.params.ocean_reference_density * m64.params.ocean_heat_capacity
    H0 = m64.heat_content()
    t64 = t32 = 0.0
    for n in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m64.time_step(args.dt)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        m32.time_step(args.dt)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t64 += t1 - t0
        t32 += t2 - t1
        if (n + 1) % 100 == 0:
            print(f"# step {n + 1}", file=sys.stderr, flush=True)
    wet = case.grid.kmin < Nz
    sst64 = m64.T[-1].cpu().numpy()
    sst32 = m32.T[-1].cpu().numpy().astype(np.float64)
    out = {"report": "mixed_vs_fp64", "workload": args.workload, "steps": args.steps, "dt": args.dt}
    diff = np.abs(sst32 - sst64)[wet]
    out["sst_rel_err"] = float(diff.max() / np.abs(sst64[wet]).max())        # max |ΔSST| / max |SST|
    out["sst_max_pointwise_rel_err"] = float(np.max(diff / np.abs(sst64[wet])))
    out["sst_max_abs_err_K"] = float(diff.max())
    out["sst_rms_err_K"] = float(np.sqrt(np.mean(diff ** 2)))
    for k in ("sensible_heat", "latent_heat", "J_T"):
        src = m64.net if k == "J_T" else m64.fluxes
        dst = m32.net if k == "J_T" else m32.fluxes
        a, b = dst[k].cpu().numpy(), src[k].cpu().numpy()
        out[f"{k}_max_err_over_scale"] = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    H64, H32 = m64.heat_content(), m32.heat_content()
    out["heat_content_rel_diff"] = abs(H32 - H64) / abs(H64)
    out["heat_content_change_fp64"] = (H64 - H0) / abs(H0)
    out["ms_per_step_fp64"] = 1e3 * t64 / args.steps
    out["ms_per_step_mixed"] = 1e3 * t32 / args.steps
    out["speedup_mixed_over_fp64"] = t64 / t32
    print(json.dumps(out), flush=True)


def cpu_baseline(case, dt: float, target_s: float) -> dict:
    """The oracle (C restatement of the reference algorithm, OpenMP, FP64) on this host's cores:
    the median of timed coupled steps after warm-up steps, with all the threads the host offers
    (at most 16) and with one thread, on the same workload as the GPU run."""
    import statistics
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle as O
    Nx, Ny, Nz = case.shape
    cells = Nx * Ny * Nz

    def sample(threads, warmup, min_steps, budget_s):
        O.set_threads(threads)
        om = O.OracleModel(case).reconcile()
        for _ in range(warmup):
            om.time_step(dt)
        times, t_start = [], time.perf_counter()
        while len(times) < min_steps or (time.perf_counter() - t_start < budget_s and len(times) < 50):
            t0 = time.perf_counter()
            om.time_step(dt)
            times.append(time.perf_counter() - t0)
            if len(times) >= min_steps and time.perf_counter() - t_start > 3 * budget_s:
                break
        return statistics.median(times), len(times)

    threads = min(16, os.cpu_count() or 1)
    t_all, n_all = sample(threads, 3, 10, target_s)
    t_one, n_one = sample(1, 1, 10, 0.0)
    cpu_model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": cells / t_all, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"median of {n_all} coupled steps of the same {Nx}x{Ny}x{Nz} workload after 3 warm-up "
                      f"steps, {threads} OpenMP threads ({cpu_model}, {os.cpu_count()} logical CPUs visible)",
            "single_thread": {"value": cells / t_one, "unit": "cell-updates/s", "cores": 1,
                              "sample": f"median of {n_one} steps after 1 warm-up step"}}


def stream_report(args):
    """ms per step with the prescribed series streamed from host memory into rings of slots, against the
    same model holding the whole series in HBM, and the host-to-device rate of one snapshot."""
    import torch
    from numericalearth_hip.generators import make_case
    from numericalearth_hip.model import OceanOnlyModel
    from numericalearth_hip.time_series import Cyclical

    Nx, Ny, Nz = WORKLOADS[args.workload]
    case = make_case(Nx, Ny, Nz, seed=args.seed, n_times=args.stream_snapshots,
                     dt_times=args.dt * args.stream_every)
    out = {"report": "stream", "workload": args.workload, "precision": args.precision, "dt": args.dt,
           "steps_between_snapshots": args.stream_every, "snapshots": args.stream_snapshots,
           "window": args.stream_window, "steps": args.steps}
    win = dict(atmosphere_window=args.stream_window, radiation_window=args.stream_window)
    for label, kw in (("resident", {}), ("streamed", win), ("streamed_copy_stream", dict(win, copy_stream=True))):
        m = OceanOnlyModel(case, precision=args.precision, time_indexing=Cyclical(), **kw)
        for _ in range(args.warmup):
            m.time_step(args.dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            m.time_step(args.dt)
        torch.cuda.synchronize()
        out[f"ms_per_step_{label}"] = 1e3 * (time.perf_counter() - t0) / args.steps
        if m.streams and label == "streamed":
            st = m.streams["atmosphere"]
            out["uploads_atmosphere"] = st.uploads
            out["stalled_prepares"] = st.stalls
            out["snapshot_bytes"] = {k: len(s.names) * s.fields[s.names[0]][0].numel() * s.fields[s.names[0]].element_size()
                                     for k, s in m.streams.items()}
            out["device_bytes_resident_series"] = {k: sum(t.numel() * t.element_size() for t in s.fields.values())
                                                   for k, s in m.streams.items()}
        m.close()
        del m
    out["overhead"] = out["ms_per_step_streamed"] / out["ms_per_step_resident"] - 1.0
    out["overhead_copy_stream"] = out["ms_per_step_streamed_copy_stream"] / out["ms_per_step_resident"] - 1.0
    nbytes = out["snapshot_bytes"]["atmosphere"]
    host = torch.empty(nbytes // 8, dtype=torch.float64, pin_memory=True)
    dev = torch.empty_like(host, device="cuda")
    dev.copy_(host, non_blocking=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(20):
        dev.copy_(host, non_blocking=True)
    torch.cuda.synchronize()
    out["h2d_gbs_one_snapshot"] = 20 * nbytes / (time.perf_counter() - t0) / 1e9
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the multi-GPU layout is a design "
                                   "(latitude-band halo exchange), not an implementation", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.mixed_report:
        mixed_report(args)
        return
    if args.ensemble:
        ensemble_report(args)
        return
    if args.serve:
        serving_report(args)
        return
    if args.stream_report:
        stream_report(args)
        return
    if args.copy_gib > 0:
        print(json.dumps(dict(copy_bandwidth(args.copy_gib, args.steps), report="copy", gib=args.copy_gib)),
              flush=True)
        return
    import torch
    from numericalearth_hip.generators import make_case
    from numericalearth_hip.model import OceanOnlyModel, bytes_per_cell_update

    Nx, Ny, Nz = WORKLOADS[args.workload]
    case = make_case(Nx, Ny, Nz, seed=args.seed)
    model = OceanOnlyModel(case, precision=args.precision, timestepper=args.timestepper)
    for _ in range(args.warmup):
        model.time_step(args.dt)
    torch.cuda.synchronize()
    if args.traffic_probe:
        traffic_probe(model, Nx * Ny * Nz * (8 if args.precision == "fp64" else 4))
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
    roofline_kernel = "tracer_rk_stage (WENO5 tendency + top flux BC + RK stage, mean of 3 stages)"
    if args.timestepper == "QuasiAdamsBashforth2":
        # the step's tendency kernel (its first timed interval) reads u, v, w, T, S and writes G_T, G_S
        stage_bytes = N3 * s3 * 7.0
        stage_ms = kernel_ms["rk_stage_1"] / nrec
        roofline_kernel = "tracer tendency of the Quasi-Adams-Bashforth 2 step (WENO5 + top flux BC)"
    achieved = stage_bytes / (stage_ms * 1e-3) / 1e9
    B = bytes_per_cell_update(args.precision, Nz, timestepper=args.timestepper)
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
        "config": {"workload": f"{args.workload} coupled step, dt={args.dt:g}s, precision={args.precision}"
                               + ("" if args.timestepper == "SplitRungeKutta3" else f", timestepper={args.timestepper}"),
                   "grid": [Nx, Ny, Nz], "seed": args.seed, "land_fraction": round(case.meta["land_fraction"], 4),
                   "bytes_per_cell_update": B, "step_hbm_fraction": step_gbs / HBM_PEAK_GBS},
        "roofline": {"kernel": roofline_kernel,
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": (measured_traffic(args.workload, args.precision)
                                                                      if args.timestepper == "SplitRungeKutta3" else None),
                     "algorithmic_bytes_per_launch": stage_bytes},
        "kernels_ms_per_step": {k: v / nrec for k, v in kernel_ms.items()},
        "flux_solve": {"mean_iterations": counters["total_iterations"] / max(1, counters["solved_columns"]),
                       "max_iterations": counters["max_iterations"],
                       "nonconverged_columns": counters["nonconverged_columns"],
                       "nan_columns": counters["nan_columns"]},
    }
    valu = measured_valu(args.workload, args.precision)
    if valu and args.precision == "fp64" and "fp64_lane_flops_per_launch" in valu:
        # the flux solve is FP64-VALU bound: issued FP64 lane-flops (ADD, MUL = 1, FMA = 2 per lane,
        # counted by SQ_INSTS_VALU_*_F64 over every issued wavefront instruction) per live launch time
        solve_s = kernel_ms["surface_update"] / nrec * 1e-3
        tf = valu["fp64_lane_flops_per_launch"] / solve_s / 1e12
        out["flux_solve"]["roofline"] = {"kernel": "surface_update (interpolation + Monin-Obukhov fixed point + column assembly)",
                                         "bound": "fp64-valu", "achieved": tf, "peak": FP64_VALU_PEAK_TFLOPS,
                                         "unit": "TFLOP/s", "frac": tf / FP64_VALU_PEAK_TFLOPS,
                                         "valu_busy": valu.get("valu_busy_surface_update")}
    elif valu and "valu_busy_surface_update" in valu:
        # the mixed form's solve iterates in FP32 first: its bound is the VALU issue rate, not FP64 flops
        out["flux_solve"]["roofline"] = {"kernel": "surface_update (interpolation + Monin-Obukhov fixed point + column assembly)",
                                         "bound": "valu", "valu_busy": valu["valu_busy_surface_update"]}
    if valu and "valu_busy_rk_stage" in valu and args.timestepper == "SplitRungeKutta3":
        out["roofline"]["valu_busy"] = valu["valu_busy_rk_stage"]
    model.close()
    if args.extras == "on" and args.workload == "quarter_degree_1440x560x10":
        # the metric's other forms on the same run: the mixed (FP32 storage) step, the 64-member
        # ensemble against its members stepped one after another, and the serving loop's latency
        other = "mixed" if args.precision == "fp64" else "fp64"
        form = form_measure(args.workload, other, args.seed, 100, 10, args.dt)
        form["speedup_mixed_over_fp64"] = (ms_per_step / form["ms_per_step"] if other == "mixed"
                                           else form["ms_per_step"] / ms_per_step)
        out["forms"] = {other: form}
        if args.precision == "fp64" and args.timestepper == "SplitRungeKutta3":   # the alternative timestepper
            ab2 = form_measure(args.workload, "fp64", args.seed, 100, 10, args.dt, "QuasiAdamsBashforth2")
            ab2["speedup_over_split_rk3"] = ms_per_step / ab2["ms_per_step"]
            out["forms"]["fp64_quasi_adams_bashforth2"] = ab2
        ens = ensemble_measure(64, args.seed, "fp64", 20, 3, args.dt)
        out["ensemble_64"] = {k: ens[k] for k in ("members", "cells", "smallest", "largest", "sequential_ms_per_step",
                                                  "batched_graph_ms_per_step", "batched_graph_cell_updates_per_s",
                                                  "speedup_graph_over_sequential")}
        srv = serving_measure(args, [args.extras_rate])["rates"][str(args.extras_rate)]
        out["serving"] = {"rate_req_per_s": args.extras_rate, "window_s": args.serve_seconds,
                          "steps_per_request": args.serve_steps, "requests": srv["offered_requests"],
                          "p50_ms": srv["batched"]["p50_ms"], "p99_ms": srv["batched"]["p99_ms"],
                          "cell_updates_per_s": srv["batched"]["cell_updates_per_s"],
                          "plain_p99_ms": srv["plain"]["p99_ms"],
                          "throughput_over_plain": srv["throughput_ratio"]}
    copy = copy_bandwidth(2.0, 20)
    out["roofline"]["measured_copy_gbs"] = copy["gbs"]
    out["roofline"]["frac_of_measured_copy"] = achieved / copy["gbs"]
    if args.cpu_baseline == "on":
        out["cpu_baseline"] = cpu_baseline(case, args.dt, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
