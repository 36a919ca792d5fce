"""Benchmark of the AWP-ODC time-stepping path on one MI355X.

A step is one pass of the hot path over the whole grid in pmcl3d order:
dvelcx_H_opt -> dstrqc_H (with the free-surface ghosts) -> addsrc_H ->
fstr_H -> receiver gather, all through the C ABI of libawp_hip.so on
inputs already resident in HBM.

Workload: anelastic_512x512x256 (h = 50 m, dt = 2.5 ms, layered + 5 % von
Karman medium, Qs = 50 vs[km/s], Qp = 2 Qs, 8-mechanism coarse-grained
memory variables, FP32).

value = interior cell-updates per second of the whole step.  roofline:
the dominant kernel (dstrqc) against MI355X HBM bandwidth with its
algorithmic traffic of 140 B per cell-update (23 per-cell arrays read,
12 written, 4 B each); its duration is measured with HIP events around
every stress launch on the launch stream.  cpu_baseline: the FP32 CPU
oracle (a port of the reference loops, OpenMP) on a bounded sub-grid of the
same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "cell-updates/s, % of HBM roofline; adjoint-gradient cost vs plain forward"
HBM_PEAK_GBS = 8000.0
BYTES_VEL = 52.0      # 10 reads + 3 writes of 4 B per cell
BYTES_STR = 140.0     # 23 reads + 12 writes of 4 B per cell
BYTES_STEP = BYTES_VEL + BYTES_STR


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="run only this many steps with no event timing (for rocprofv3)")
    return ap.parse_args()


def cpu_baseline(workload_fn, target_s: float) -> dict:
    """FP32 oracle on a 128 x 128 x nz sub-grid of the same medium model,
    using the host cores, timed per step until about target_s seconds."""
    import numpy as np

    from oracle.oracle import OracleSolver, set_threads
    threads = int(os.environ.get("AWP_CPU_THREADS", min(16, os.cpu_count() or 1)))
    set_threads(threads)
    w = workload_fn(128, 128)
    med = w.medium()
    s = OracleSolver(med, w.sources, None, dtype=np.float32)
    s.start()
    s.step()  # first touch
    t0 = time.perf_counter()
    n = 0
    while True:
        s.step()
        n += 1
        el = time.perf_counter() - t0
        if el >= target_s or n >= 50:
            break
    cells = w.dims.ncells
    return {"value": cells * n / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"FP32 oracle, {w.dims.nxt}x{w.dims.nyt}x{w.dims.nzt} sub-grid of the same "
                      f"layered+random+Q medium, {n} steps in {el:.1f} s, {threads} OpenMP threads, "
                      f"cpu={_cpu_model()}"}


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of "
                          "it (the path would shard as 2-D x/y slabs with an 8-cell velocity "
                          "halo exchange per step)", "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    from awp_amd import _lib, configs
    from awp_amd.solver import Solver

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)

    def workload(nx, ny):
        return configs.anelastic(nx=nx, ny=ny, nz=args.nz, nt=max(args.steps + args.warmup, 1))

    w = workload(args.nx, args.ny)
    t0 = time.perf_counter()
    med = w.medium()
    setup_s = time.perf_counter() - t0
    sol = Solver(med, w.sources, w.receivers, nt_record=args.steps + args.warmup + 1)
    del med
    cells = w.dims.ncells
    sol.start()

    if args.profile_steps:
        for _ in range(args.profile_steps):
            sol.step()
        torch.cuda.synchronize()
        _lib.check()
        print(json.dumps({"profile_steps": args.profile_steps, "workload": w.name}))
        return

    for _ in range(args.warmup):
        sol.step()
    torch.cuda.synchronize()
    _lib.check()

    # Timed region: K full steps; HIP events bracket each velocity and
    # stress launch on the stream the launchers use (torch's current stream).
    stream = torch.cuda.current_stream()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for n in range(args.steps):
        e = ev[n]
        sol.cur_step += 1
        e[0].record(stream)
        sol.velocity()
        e[1].record(stream)
        sol.stress()
        e[2].record(stream)
        if sol.cur_step < sol.source_steps(0):
            sol.add_source(0, sol.cur_step + 1)
        sol.free_surface()
        sol.record(sol.cur_step - 1)
        e[3].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _lib.check()
    vel_ms = [e[0].elapsed_time(e[1]) for e in ev]
    str_ms = [e[1].elapsed_time(e[2]) for e in ev]
    step_ms = [e[0].elapsed_time(e[3]) for e in ev]
    vel_avg = float(np.mean(vel_ms))
    str_avg = float(np.mean(str_ms))
    ms_per_step = wall * 1e3 / args.steps
    value = cells * args.steps / wall
    str_gbs = BYTES_STR * cells / (str_avg * 1e-3) / 1e9
    vel_gbs = BYTES_VEL * cells / (vel_avg * 1e-3) / 1e9
    step_gbs = BYTES_STEP * cells / (ms_per_step * 1e-3) / 1e9
    tr = sol.traces()
    finite = bool(np.isfinite(tr).all())

    traffic = None
    tfile = ROOT / "profiles" / "dstrqc_anelastic_512x512x256_traffic.json"
    if tfile.exists() and (args.nx, args.ny, args.nz) == (512, 512, 256):
        try:
            traffic = json.loads(tfile.read_text()).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

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
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w.name, "nx": w.dims.nxt, "ny": w.dims.nyt, "nz": w.dims.nzt,
                   "h_m": w.DH, "dt_s": w.DT, "attenuation": "Qs=50vs[km/s], Qp=2Qs, 8 mechanisms",
                   "medium": "layered vs 1500-4500 m/s + 5% von Karman (a=500 m, H=0.5), seed 7",
                   "step": "dvelcx_H_opt, dstrqc_H (+FS ghosts), addsrc_H, fstr_H, receivers"},
        "roofline": {"bound": "hbm", "kernel": "dstrqc_H", "achieved": str_gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": str_gbs / HBM_PEAK_GBS,
                     "traffic": traffic, "algorithmic_bytes_per_cell": BYTES_STR,
                     "cells_per_launch": cells, "avg_launch_ms": str_avg},
        "kernels": {"dvelcx_H_opt": {"avg_ms": vel_avg, "GBps": vel_gbs,
                                     "frac": vel_gbs / HBM_PEAK_GBS},
                    "dstrqc_H": {"avg_ms": str_avg, "GBps": str_gbs,
                                 "frac": str_gbs / HBM_PEAK_GBS},
                    "step_events_avg_ms": float(np.mean(step_ms))},
        "step_roofline": {"bytes_per_cell": BYTES_STEP, "achieved_GBps": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS},
        "traces_finite": finite,
        "setup_s": setup_s,
    }
    del sol
    torch.cuda.empty_cache()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(workload, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
