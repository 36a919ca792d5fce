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
same workload.  adjoint: the metric's third quantity, the cost of one
material gradient of the 256 x 256 x 128, 2000-step adjoint workload over
one plain forward of it (checkpoint every 50 steps), with one
finite-difference direction as a sanity check.
"""
from __future__ import annotations

import argparse
import ctypes
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
# awp_step: u v w and the six stresses read once (36 B) and written once
# (36 B), memory variables read and written (48 B), d1 lam mu qp qs (20 B),
# vx1 vx2 ww wwo (16 B)
BYTES_FUSED = 156.0
# awp_dstrqc_tex: vx1, vx2, wwo from a table (and ww from init_texture's rule)
TEXTURE_SAVES = {"arrays": 0.0, "table": 12.0, "rule": 16.0}
# 16-bit wavefield storage: 9 x 2 B + d1 4 B read, 3 x 2 B written (velocity);
# 15 x 2 B + 8 medium x 4 B read, 12 x 2 B written (stress)
BYTES16_VEL = 28.0
BYTES16_STR = 86.0


def _stress_entry(args) -> str:
    """The stress launcher the anelastic workload's step calls."""
    s16 = args.storage != "f32"
    if args.texture == "arrays":
        return "awp_dstrqc_s16" if s16 else "dstrqc_H"
    return "awp_dstrqc_tex_s16" if s16 else "awp_dstrqc_tex"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="launch-shape knob for awp_set_tuning (repeatable); results do not "
                         "depend on these")
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=None,
                    help="cells along x (default per workload: anelastic 512, topo 384, "
                    "adjoint 256, dm coarse grid 256)")
    ap.add_argument("--ny", type=int, default=None)
    ap.add_argument("--nz", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-adjoint", dest="adjoint_check", action="store_false",
                    help="anelastic workload: skip the adjoint-cost measurement that the "
                         "default line carries (256x256x128, 2000 steps, one FD direction)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="run only this many steps with no event timing (for rocprofv3)")
    ap.add_argument("--workload",
                    choices=("anelastic", "dm", "adjoint", "topo", "lagrange", "kinematic",
                             "cpu", "stream"),
                    default="anelastic",
                    help="anelastic: the headline single-grid workload; dm: discontinuous mesh "
                         "+ plasticity (fine 3nx x 3ny x nz-fine over coarse nx x ny x nz); "
                         "adjoint: gradient cost and finite-difference check; topo: curvilinear "
                         "topography (384 x 384 x 192 unless --nx/--ny/--nz say otherwise); "
                         "lagrange: Lagrange-interpolated receivers and sources on the "
                         "512 x 512 x 256 layout; kinematic: finite-fault subfaults "
                         "(addkinsrc_H, Brune/Liu, 4th-order filter) on the same layout; cpu: "
                         "the CPU baseline table only (no GPU); stream: the HBM bandwidth of "
                         "each kernel's access mix with no stencil (awp_stream_probe)")
    ap.add_argument("--receivers", type=int, default=20000,
                    help="lagrange workload: receivers (three components each)")
    ap.add_argument("--subfaults", type=int, default=4000,
                    help="lagrange workload: point sources (six components each)")
    ap.add_argument("--degree", type=int, default=3, help="lagrange workload: degree")
    ap.add_argument("--volume-output", type=int, default=0, metavar="NTISKP",
                    help="anelastic workload: also write the free-surface velocities every "
                         "NTISKP steps (velbuffer, asynchronous copies, ten samples per file)")
    ap.add_argument("--subfault-order", choices=("cell", "file"), default="cell",
                    help="kinematic workload: subfaults sorted by cell at load time (the "
                         "default) or in a shuffled file order")
    ap.add_argument("--save-traces", default="",
                    help="write the receiver traces (numpy .npy) of the timed run here")
    ap.add_argument("--nz-fine", type=int, default=96)
    ap.add_argument("--storage", choices=("f32", "bf16", "fp16"), default="f32",
                    help="wavefield storage of the anelastic workload (16-bit: per-field "
                         "power-of-two scales, binary32 arithmetic)")
    ap.add_argument("--texture", choices=("arrays", "table", "rule"), default=None,
                    help="relaxation texture of the stress update: per-cell vx1/vx2/wwo "
                         "arrays (dstrqc_H), or awp_dstrqc_tex's 8-entry table with the ww "
                         "array (-12 B/cell) or with init_texture's parity rule (-16 B/cell); "
                         "the dm workload takes rule per grid where ww follows it, else the "
                         "table.  Default: rule, except arrays for FP16 storage, whose table "
                         "form measured slower (2.06 against 1.72 ms per stress update)")
    ap.add_argument("--fused", action="store_true",
                    help="anelastic workload: the whole step in one sweep (awp_step, state A -> "
                         "state B, texture table; 144 B/cell, 140 B with --texture rule)")
    ap.add_argument("--nt", type=int, default=2000, help="adjoint workload: time steps")
    ap.add_argument("--checkpoint-every", type=int, default=50)
    ap.add_argument("--fd-directions", type=int, default=5,
                    help="adjoint workload: random directions of the finite-difference check")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only copy a BYTES device buffer twice (device-to-device): the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def cpu_baseline(workload_fn, target_s: float) -> dict:
    """FP32 oracle on a 128 x 128 x nz sub-grid of the same medium model,
    using the host cores, timed per step until about target_s seconds."""
    import numpy as np

    from oracle.oracle import OracleSolver, set_threads
    threads = host_threads()
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
                      f"layered+random+Q medium, {n} steps in {el:.1f} s, {threads} OpenMP threads "
                      f"(every CPU this job may use; the host has {os.cpu_count()}), "
                      f"cpu={_cpu_model()}"}


def bench_cpu(args) -> None:
    """CPU baseline table: the oracle (the reference's kernel loops restated in
    C++ with OpenMP) on this host's cores, at 1 thread and at every CPU the
    job may use (host_threads), in FP32 and FP64.  The elastic 64^3
    workload runs its 200 steps; the 512 x 512 x 256 anelastic workload runs
    until --cpu-seconds or 10 steps, whichever comes first.  One untimed step
    first touches the arrays."""
    import numpy as np

    from awp_amd import configs
    from oracle.oracle import OracleSolver, set_threads
    allt = host_threads()
    rows = []
    cases = (("elastic_64x64x64", configs.elastic_halfspace(n=64, nt=201), 200, None),
             ("anelastic_512x512x256", configs.anelastic(nt=11), 10, args.cpu_seconds))
    for name, w, nmax, tmax in cases:
        med = w.medium()
        for dtype in (np.float32, np.float64):
            for th in sorted({1, allt}):
                set_threads(th)
                s = OracleSolver(med, w.sources, None, dtype=dtype)
                s.start()
                s.step()
                t0 = time.perf_counter()
                n = 0
                while n < nmax:
                    s.step()
                    n += 1
                    el = time.perf_counter() - t0
                    if tmax and el >= tmax:
                        break
                del s
                row = {"workload": name, "dtype": np.dtype(dtype).name, "threads": th,
                       "steps": n, "seconds": round(el, 3),
                       "cell_updates_per_s": w.dims.ncells * n / el}
                rows.append(row)
                print(json.dumps(row), file=sys.stderr, flush=True)
        del med
    print(json.dumps({"metric": "CPU baseline (oracle on host cores)", "unit": "cell-updates/s",
                      "cpu": _cpu_model(), "host_cpus": os.cpu_count(), "job_threads": allt,
                      "rows": rows}))


def host_threads() -> int:
    """Threads for the CPU baseline: AWP_CPU_THREADS if set, else every CPU
    this job may run on (OMP_NUM_THREADS where the host sets the job's CPU
    share, else the process's CPU affinity)."""
    if os.environ.get("AWP_CPU_THREADS"):
        return int(os.environ["AWP_CPU_THREADS"])
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        return min(int(omp), aff)
    return aff


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _size(args, nx: int, ny: int, nz: int) -> tuple:
    """The workload's grid: --nx/--ny/--nz where given, else its defaults."""
    return (args.nx or nx, args.ny or ny, args.nz or nz)


def main():
    args = parse()
    if args.texture is None:
        args.texture = "arrays" if args.storage == "fp16" else "rule"
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of "
                          "it (the path would shard as 2-D x/y slabs with an 8-cell velocity "
                          "halo exchange per step)", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "cpu":
        bench_cpu(args)
        return
    import numpy as np
    import torch

    from awp_amd import _lib, configs
    from awp_amd.solver import Solver

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    for kv in args.tune:
        key, _, val = kv.partition("=")
        if _lib.load().awp_set_tuning(key.encode(), int(val)) != 0:
            print(f"unknown tuning key or value: {kv}", file=sys.stderr)
            sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.workload == "dm":
        bench_dm(args)
        return
    if args.workload == "stream":
        bench_stream(args)
        return
    if args.workload == "adjoint":
        bench_adjoint(args)
        return
    if args.workload == "topo":
        bench_topo(args)
        return
    if args.workload == "lagrange":
        bench_lagrange(args)
        return
    if args.workload == "kinematic":
        bench_kinematic(args)
        return
    args.nx, args.ny, args.nz = _size(args, 512, 512, 256)

    def workload(nx, ny):
        return configs.anelastic(nx=nx, ny=ny, nz=args.nz, nt=max(args.steps + args.warmup, 1))

    w = workload(args.nx, args.ny)
    t0 = time.perf_counter()
    med = w.medium()
    setup_s = time.perf_counter() - t0
    if args.storage == "f32":
        sol = Solver(med, w.sources, w.receivers, nt_record=args.steps + args.warmup + 1,
                     texture=args.texture, fused=args.fused)
        b_vel, b_str = BYTES_VEL, BYTES_STR - TEXTURE_SAVES[args.texture]
        if args.fused:
            # one sweep: the state read once and written once, the texture
            # from the table (ww read, or from the parity rule)
            b_vel, b_str = 0.0, BYTES_FUSED - (16.0 if args.texture == "rule" else 12.0)
    else:
        from awp_amd.solver import Solver16
        sol = Solver16(med, w.sources, w.receivers, nt_record=args.steps + args.warmup + 1,
                       storage=args.storage, texture=args.texture)
        # the binary32 medium's vx1, vx2, wwo (and ww) as for FP32 storage
        b_vel, b_str = BYTES16_VEL, BYTES16_STR - TEXTURE_SAVES[args.texture]
    check_every = getattr(sol, "check_every", 0)
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

    vout = None
    if args.volume_output:
        # pmcl3d's volume output of the free surface every --volume-output
        # steps, ten samples per file, written to a scratch directory
        import tempfile

        from awp_amd.output import VolumeOutput
        d = sol.dims
        vout = VolumeOutput(d, (0, 0, 0), (d.nxt - 1, d.nyt - 1, 0), (1, 1, 1),
                            ntiskp=args.volume_output, write_step=10,
                            basenames=("SX", "SY", "SZ"), directory=tempfile.mkdtemp())

    for _ in range(args.warmup):
        sol.step()
        if vout is not None:
            vout.sample(sol.cur_step, sol.f["u1"], sol.f["v1"], sol.f["w1"])
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
        if args.fused:
            sol.fused_step()
            e[1].record(stream)
        else:
            sol.velocity()
            e[1].record(stream)
            sol.stress()
        e[2].record(stream)
        if sol.cur_step < sol.source_steps(0):
            sol.add_source(0, sol.cur_step + 1)
        sol.free_surface()
        sol.record(sol.cur_step - 1)
        if check_every and sol.cur_step % check_every == 0:
            sol.check_scales()
        if vout is not None:
            vout.sample(sol.cur_step, sol.f["u1"], sol.f["v1"], sol.f["w1"])
        e[3].record(stream)
    torch.cuda.synchronize()
    if vout is not None:
        vout.close()
    wall = time.perf_counter() - t0
    _lib.check()
    vel_ms = [e[0].elapsed_time(e[1]) for e in ev]
    str_ms = [e[1].elapsed_time(e[2]) for e in ev]
    step_ms = [e[0].elapsed_time(e[3]) for e in ev]
    vel_avg = float(np.mean(vel_ms))
    str_avg = float(np.mean(str_ms))
    if args.fused:
        # one launch between the first two events
        str_avg, vel_avg = vel_avg, 0.0
    ms_per_step = wall * 1e3 / args.steps
    value = cells * args.steps / wall
    str_gbs = b_str * cells / (str_avg * 1e-3) / 1e9
    vel_gbs = b_vel * cells / (vel_avg * 1e-3) / 1e9 if vel_avg > 0 else 0.0
    step_gbs = (b_vel + b_str) * cells / (ms_per_step * 1e-3) / 1e9
    tr = sol.traces()
    finite = bool(np.isfinite(tr).all())

    traffic = None
    tname = {("arrays", False): "dstrqc_anelastic_512x512x256_traffic.json",
             ("rule", False): "dstrqc_rule_anelastic_512x512x256_traffic.json",
             ("rule", True): "awp_step_rule_anelastic_512x512x256_traffic.json"}.get(
                 (args.texture, bool(args.fused)), "")
    tfile = ROOT / "profiles" / tname
    if (tname and tfile.exists() and (args.nx, args.ny, args.nz) == (512, 512, 256)
            and args.storage == "f32"):
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
        "dtype": "f32" if args.storage == "f32" else f"f32 arithmetic, {args.storage} storage",
        "data": "synthetic",
        "config": {"workload": w.name, "storage": args.storage, "nx": w.dims.nxt, "ny": w.dims.nyt, "nz": w.dims.nzt,
                   "h_m": w.DH, "dt_s": w.DT, "attenuation": "Qs=50vs[km/s], Qp=2Qs, 8 mechanisms",
                   "medium": "layered vs 1500-4500 m/s + 5% von Karman (a=500 m, H=0.5), seed 7",
                   "texture": args.texture,
                   "step": (("awp_step (velocity, FS ghosts and stress in one sweep)"
                             if args.fused else
                             "dvelcx_H_opt, " + (_stress_entry(args) + (
                                 "" if args.texture == "arrays" else f" ({args.texture})"))
                             + " (+FS ghosts)") + ", addsrc_H, fstr_H, receivers")
                           + (f", free-surface volume output every {args.volume_output} steps"
                              if args.volume_output else "")},
        "roofline": {"bound": "hbm",
                     "kernel": ("awp_step" if args.fused else _stress_entry(args)),
                     "achieved": str_gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": str_gbs / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "traffic_source": ("HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE and "
                                        "WRITE_SIZE of this kernel at this size, scaled by a copy "
                                        f"kernel's counts (profiles/{tname}); read from that "
                                        "file, not counted in this run")
                     if traffic is not None else None,
                     "algorithmic_bytes_per_cell": b_str,
                     "cells_per_launch": cells, "avg_launch_ms": str_avg},
        "kernels": ({"awp_step": {"avg_ms": str_avg, "GBps": str_gbs,
                                  "frac": str_gbs / HBM_PEAK_GBS},
                     "step_events_avg_ms": float(np.mean(step_ms))} if args.fused else
                    {"dvelcx_H_opt": {"avg_ms": vel_avg, "GBps": vel_gbs,
                                      "frac": vel_gbs / HBM_PEAK_GBS},
                     _stress_entry(args):
                         {"avg_ms": str_avg, "GBps": str_gbs, "frac": str_gbs / HBM_PEAK_GBS},
                     "step_events_avg_ms": float(np.mean(step_ms))}),
        "step_roofline": {"bytes_per_cell": b_vel + b_str, "achieved_GBps": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS},
        "traces_finite": finite,
        "setup_s": setup_s,
    }
    del sol
    torch.cuda.empty_cache()
    if args.adjoint_check and args.storage == "f32":
        # the metric's third quantity: gradient cost of the 256x256x128,
        # 2000-step adjoint workload over one plain forward of it
        out["adjoint"] = adjoint_cost(256, 256, 128, 2000, 50, 1)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(workload, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


# (name, reads, writes) per cell: the access mixes of the stencil kernels
STREAM_MIXES = (("copy", 1, 1), ("dvelcx (52 B)", 10, 3), ("dstrqc_H (140 B)", 23, 12),
                ("awp_dstrqc_tex, rule (124 B)", 19, 12), ("awp_step, rule (140 B)", 20, 15))


def bench_stream(args) -> None:
    """The bandwidth each access mix reaches with no stencil: awp_stream_probe
    over arrays of the 512x512x256 field size (padded, 90.6 M floats), every
    element read or written once, at 4, 8 and 16 bytes per lane, with each
    array's load waited for before the next (serial) or all of a lane's loads
    in flight together (batched), over a sweep of grid sizes; the best of
    each is reported."""
    import numpy as np
    import torch

    from awp_amd import _lib
    from awp_amd.grid import Dims
    lib = _lib.load()
    d = Dims(512, 512, 256)
    n = d.size
    nin = max(m[1] for m in STREAM_MIXES)
    nout = max(m[2] for m in STREAM_MIXES)
    ins = [torch.full((n,), 1.0 + a, dtype=torch.float32, device="cuda") for a in range(nin)]
    outs = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(nout)]
    pin = (ctypes.c_void_p * nin)(*[t.data_ptr() for t in ins])
    pout = (ctypes.c_void_p * nout)(*[t.data_ptr() for t in outs])
    st = torch.cuda.current_stream()
    rows = []
    for name, r, w in STREAM_MIXES:
        row = {"mix": name, "reads": r, "writes": w, "bytes_per_cell": 4 * (r + w)}
        for sched, batched in (("serial", 0), ("batched", 1)):
            best = {}
            for vec4 in (0, 2, 1):
                for blocks in (1024, 2048, 4096, 8192, 16384):
                    def go():
                        rc = lib.awp_stream_probe(pin, r, pout, w, n, vec4, blocks, batched,
                                                  st.cuda_stream)
                        if rc:
                            _lib.check()
                    for _ in range(2):
                        go()
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    reps = max(3, args.steps if args.steps < 50 else 10)
                    e0.record(st)
                    for _ in range(reps):
                        go()
                    e1.record(st)
                    torch.cuda.synchronize()
                    ms = e0.elapsed_time(e1) / reps
                    gbs = (r + w) * 4.0 * n / (ms * 1e-3) / 1e9
                    key = {0: "4B", 1: "16B", 2: "8B"}[vec4]
                    if gbs > best.get(key, {}).get("GBps", 0.0):
                        best[key] = {"GBps": gbs, "ms": ms, "blocks": blocks}
            want = float(sum(1.0 + a for a in range(r)))
            ok = all(bool(np.isclose(float(outs[a][i].item()), want + a))
                     for a in range(w) for i in (0, n - 1))
            row[sched] = dict(best, checked=ok)
        rows.append(row)
        print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
    _lib.check()
    print(json.dumps({"metric": "HBM streaming probe (awp_stream_probe)", "unit": "GB/s",
                      "floats_per_array": n, "peak_GBps": HBM_PEAK_GBS, "rows": rows}))


def copy_probe(nbytes: int) -> None:
    """Two device-to-device copies of nbytes: under rocprofv3 --pmc their
    FETCH_SIZE / WRITE_SIZE per dispatch calibrate the counters (nbytes read
    and nbytes written per copy)."""
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 2
    print(json.dumps({"copy_probe_bytes": n * 4, "GBps_read_plus_write": 2 * n * 4 / dt / 1e9}))


def adjoint_cost(nx: int, ny: int, nz: int, nt: int, checkpoint_every: int,
                 fd_directions: int) -> dict:
    """Adjoint gradient: smooth background vs a 10 % Gaussian vs anomaly,
    4 explosive sources, 8 x 8 surface receivers, h = 100 m.  Times one
    plain forward (FP32, the pmcl3d step) and one gradient (checkpointed
    forward with the binary64 misfit, segment recomputation, adjoint sweep,
    mapping to lam, mu, rho), and checks the gradient against central finite
    differences of the misfit along seeded random directions."""
    import numpy as np
    import torch

    from awp_amd import configs
    from awp_amd.adjoint import AdjointProblem
    from awp_amd.solver import Solver

    t0 = time.perf_counter()
    bg, true, src, rec = configs.adjoint_pair(nx=nx, ny=ny, nz=nz, nt=nt)
    setup_s = time.perf_counter() - t0
    print(f"adjoint setup {setup_s:.1f} s", file=sys.stderr, flush=True)
    obs_run = Solver(true, src, rec, nt_record=nt)
    obs_run.run(nt)
    obs = obs_run.traces()
    del obs_run
    torch.cuda.empty_cache()
    P = AdjointProblem(bg, src, rec, nt, obs, checkpoint_every=checkpoint_every)
    lam0 = torch.from_numpy(bg.lam.copy())
    mu0 = torch.from_numpy(bg.mu.copy())
    rho0 = torch.from_numpy(bg.d1.copy())
    P.set_medium(lam0, mu0, rho0)
    # plain forward, as the forward-only workload runs it
    plain = P.solver
    P._reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain.run(nt)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    print(f"plain forward {t_fwd:.3f} s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = P.forward()
    glam, gmu, grho = P.backward()
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    print(f"gradient {t_grad:.3f} s", file=sys.stderr, flush=True)
    glam, gmu, grho = glam.cpu(), gmu.cpu(), grho.cpu()
    rng = np.random.default_rng(2024)
    fd = []
    for n in range(fd_directions):
        dl, dm, dr = (torch.from_numpy(rng.standard_normal(a.shape)) * a
                      for a in (lam0, mu0, rho0))
        eps = 2e-3

        def Jof(s):
            P.set_medium(lam0 + s * dl, mu0 + s * dm, rho0 + s * dr)
            return P.forward()
        f = (Jof(eps) - Jof(-eps)) / (2 * eps)
        a = float((glam * dl).sum() + (gmu * dm).sum() + (grho * dr).sum())
        fd.append({"adjoint": a, "finite_difference": f, "rel_err": abs(f - a) / abs(a)})
        print(f"direction {n}: {fd[-1]}", file=sys.stderr, flush=True)
    del P
    torch.cuda.empty_cache()
    return {"workload": f"adjoint_{nx}x{ny}x{nz}", "nt": nt, "checkpoint_every": checkpoint_every,
            "value": t_grad / t_fwd, "unit": "gradient time / plain forward time",
            "plain_forward_s": t_fwd, "gradient_s": t_grad,
            "forward_cell_updates_per_s": nx * ny * nz * nt / t_fwd, "misfit": J,
            "fd_check": fd, "fd_max_rel_err": max(x["rel_err"] for x in fd) if fd else None,
            "setup_s": setup_s}


def bench_adjoint(args) -> None:
    """The adjoint workload alone (256 x 256 x 128 unless --nx/--ny/--nz
    say otherwise), as one JSON line."""
    nx, ny, nz = _size(args, 256, 256, 128)
    r = adjoint_cost(nx, ny, nz, args.nt, args.checkpoint_every, args.fd_directions)
    out = {
        "metric": METRIC, "value": r["value"], "unit": r["unit"],
        "n_gpus": 1, "steps": args.nt, "warmup": 0,
        "ms_per_step": r["plain_forward_s"] * 1e3 / args.nt,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 wavefields, f64 misfit and gradient", "data": "synthetic",
        "config": {"workload": r["workload"], "h_m": 100.0, "dt_s": 0.005, "nt": args.nt,
                   "checkpoint_every": args.checkpoint_every, "sources": "4 explosive",
                   "receivers": "8 x 8 surface", "model": "smooth vs(z) + 10 % Gaussian anomaly"},
        "roofline": None, "cpu_baseline": None,
    }
    out.update({k: v for k, v in r.items() if k not in ("value", "unit", "workload")})
    print(json.dumps(out))


def bench_topo(args) -> None:
    """Curvilinear topography: the pmcl3d step with the topo_* velocity and
    stress launchers on grid 0 (source, free-surface pass, receivers), HIP
    events around each launcher on the launch stream.  FP32 or 16-bit
    storage.  Algorithmic bytes per cell as for the Cartesian kernels (the
    metrics are 2-D and 1-D arrays)."""
    import numpy as np
    import torch

    from awp_amd import _lib, configs
    from awp_amd.solver import Solver, Solver16

    nx, ny, nz = _size(args, 384, 384, 192)
    nt = max(args.steps + args.warmup, 1)
    w = configs.topography_hills(nx=nx, ny=ny, nz=nz, nt=nt)
    topo = w.params["topography"]
    t0 = time.perf_counter()
    med = w.medium()
    setup_s = time.perf_counter() - t0
    if args.storage == "f32":
        # the texture applies to the two-pass stress update's Q pass
        # (awp_topo_stress_interior_tex_H)
        sol = Solver(med, w.sources, w.receivers, nt_record=nt + 1, topography=topo,
                     texture=args.texture)
        b_vel, b_str = BYTES_VEL, BYTES_STR - TEXTURE_SAVES[args.texture]
    else:
        sol = Solver16(med, w.sources, w.receivers, nt_record=nt + 1, storage=args.storage,
                       topography=topo)
        b_vel, b_str = BYTES16_VEL, BYTES16_STR
    check_every = getattr(sol, "check_every", 0)
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
        if check_every and sol.cur_step % check_every == 0:
            sol.check_scales()
        e[3].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _lib.check()
    vel_avg = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
    str_avg = float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))
    ms_per_step = wall * 1e3 / args.steps
    str_gbs = b_str * cells / (str_avg * 1e-3) / 1e9
    vel_gbs = b_vel * cells / (vel_avg * 1e-3) / 1e9
    tr = sol.traces()
    if args.save_traces:
        np.save(args.save_traces, tr)
    out = {
        "metric": METRIC, "value": cells * args.steps / wall, "unit": "cell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.storage == "f32" else f"f32 arithmetic, {args.storage} storage",
        "data": "synthetic",
        "config": {"workload": w.name, "storage": args.storage, "nx": nx, "ny": ny, "nz": nz,
                   "h_m": w.DH, "dt_s": w.DT,
                   "topography": "6 seeded Gaussian hills, height <= 1.5 km, widths 2-8 km",
                   "medium": "layered vs 1500-4500 m/s + 5% von Karman, Q",
                   "step": "topo velocity + stress (Q) on grid 0, addsrc_H, fstr_H, receivers",
                   "texture": getattr(sol.levels[0], "texmode", "arrays")},
        "roofline": {"bound": "hbm",
                     "kernel": ("awp_topo_stress_interior_tex_H"
                                if getattr(sol.levels[0], "texmode", "arrays") != "arrays"
                                else "awp_topo_stress_interior_rates_H"),
                     "achieved": str_gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": str_gbs / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes_per_cell": b_str,
                     "cells_per_launch": cells, "avg_launch_ms": str_avg},
        "kernels": {"topo_velocity_interior_H": {"avg_ms": vel_avg, "GBps": vel_gbs,
                                                 "frac": vel_gbs / HBM_PEAK_GBS},
                    "topo_stress_interior_H": {"avg_ms": str_avg, "GBps": str_gbs,
                                               "frac": str_gbs / HBM_PEAK_GBS}},
        "traces_finite": bool(np.isfinite(tr).all()), "trace_absmax": float(np.abs(tr).max()),
        "setup_s": setup_s, "cpu_baseline": None,
    }
    print(json.dumps(out))


def bench_dm(args) -> None:
    """Discontinuous mesh + plasticity: whole pmcl3d steps (both grids, the
    coupling kernels, NVE = 3 region split, yield factors, free surface,
    receivers), each bracketed by HIP events on the launch stream.  Sizes
    default to the 256 x 256 x 192 coarse grid (h = 75 m) under a
    768 x 768 x 96 fine grid (h = 25 m)."""
    import numpy as np
    import torch

    from awp_amd import _lib, configs
    from awp_amd.solver import Solver

    nx, ny, nz = _size(args, 256, 256, 192)
    nt = max(args.steps + args.warmup, 1)
    w = configs.dm_plastic(nx=nx, ny=ny, nz=nz, nz_fine=args.nz_fine, nt=nt)
    t0 = time.perf_counter()
    media = w.medium()
    pm = w.params["plastic_fn"](media)
    setup_s = time.perf_counter() - t0
    print(f"dm setup {setup_s:.1f} s", file=sys.stderr, flush=True)
    # the coarse grid's top levels carry the fine grid's texture
    # (pmcl3d.c:1049-1058), so the parity rule holds on the fine grid only:
    # "auto" takes the rule there and the ww table on the coarse grid
    tex = "auto" if args.texture == "rule" else args.texture
    sol = Solver(media, w.sources, w.receivers, nt_record=nt + 1, plastic=pm, texture=tex)
    cells = sum(m.dims.ncells for m in media)
    del media, pm
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
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for n in range(args.steps):
        ev[n][0].record(stream)
        sol.step()
        ev[n][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _lib.check()
    step_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    tr = sol.traces()
    d0, d1 = sol.grid_dims(0), sol.grid_dims(1)
    out = {
        "metric": METRIC, "value": cells * args.steps / wall, "unit": "cell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": w.name, "fine": [d0.nxt, d0.nyt, d0.nzt], "coarse": [d1.nxt, d1.nyt, d1.nzt],
                   "h_m": [w.params["DH"][0], w.params["DH"][1]], "dt_s": w.DT,
                   "plasticity": "Drucker-Prager c=5 MPa, mu=0.75, Tv=0.01 s",
                   "source": "200 double-couple subfaults in the coarse grid",
                   "texture": [L.texmode for L in sol.levels]},
        "step_events_avg_ms": step_ms, "cells_per_step": cells,
        "traces_finite": bool(np.isfinite(tr).all()), "setup_s": setup_s,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def _lagrange_setup(nx, ny, nz, nrec, nsrc, degree, seed=11):
    """Coordinates of the metric grid of an nx x ny x nz mesh (h = 50 m), the
    receiver positions (uniform in the volume) and a finite-fault cluster:
    nsrc points on a vertical plane, 40 % of the mesh long and 30 % deep,
    spaced about one cell apart, so neighbouring stencils overlap."""
    import numpy as np

    from awp_amd.grid import Dims, NGSL
    from awp_amd.interpolation import grid_coordinates, metric_grid
    h = 50.0
    dims = Dims(nx, ny, nz)
    grid = metric_grid(dims)
    x = grid_coordinates(grid.size_x, h, padding=NGSL)
    y = grid_coordinates(grid.size_y, h, padding=NGSL)
    z = grid_coordinates(grid.size_z, h)
    rng = np.random.default_rng(seed)
    rec = [rng.uniform(0.0, n * h, nrec) for n in (nx, ny, nz - 4)]
    side = int(np.ceil(np.sqrt(nsrc)))
    u, v = np.meshgrid(np.linspace(0.3, 0.7, side), np.linspace(0.4, 0.7, side), indexing="ij")
    u, v = u.ravel()[:nsrc], v.ravel()[:nsrc]
    src = [np.full(nsrc, 0.5 * nx * h) + rng.uniform(-0.5 * h, 0.5 * h, nsrc), u * ny * h,
           v * (nz - 4) * h]
    return dims, grid, (x, y, z), rec, src, h


def bench_lagrange(args) -> None:
    """Lagrange-interpolated receivers and sources (cuinterp_interp_H and
    cusource_add_cartesian_H) on the 512 x 512 x 256 field layout.  A step
    samples the three velocity fields at every receiver and adds the six
    moment-rate components of every point source into the six stress fields
    (degree 3 by default).  value and ms_per_step come from the wall clock
    of the timed steps; each kernel's duration from HIP events on the launch
    stream."""
    import numpy as np
    import torch

    from awp_amd import _lib
    from awp_amd.interpolation import Interpolation

    nx, ny, nz = _size(args, 512, 512, 256)
    nrec, nsrc, deg = args.receivers, args.subfaults, args.degree
    dims, grid, (x, y, z), rec, src, h = _lagrange_setup(nx, ny, nz, nrec, nsrc, deg)
    dt = 0.0025
    t0 = time.perf_counter()
    R = Interpolation(x, y, z, grid, *rec, degree=deg)
    S = Interpolation(x, y, z, grid, *src, degree=deg)
    setup_s = time.perf_counter() - t0
    print(f"lagrange setup {setup_s:.2f} s; source map {S.num_cells} points for "
          f"{nsrc * (deg + 1) ** 3} terms", file=sys.stderr, flush=True)
    gen = torch.Generator(device="cuda").manual_seed(1)
    vel = [torch.randn(dims.size, device="cuda", generator=gen) for _ in range(3)]
    stress = [torch.zeros(dims.size, device="cuda") for _ in range(6)]
    nsteps_rec = args.steps + args.warmup
    traces = torch.zeros((3, nsteps_rec, nrec), device="cuda")
    mrate = torch.randn((6, nsteps_rec, nsrc), device="cuda", generator=gen)

    def step(n):
        for c in range(3):
            R.interp(traces[c, n], vel[c])
        for c in range(6):
            S.add_cartesian(stress[c], mrate[c, n], h, dt)

    if args.profile_steps:
        for n in range(args.profile_steps):
            step(n)
        torch.cuda.synchronize()
        print(json.dumps({"profile_steps": args.profile_steps, "workload": "lagrange"}))
        return
    for n in range(args.warmup):
        step(n)
    torch.cuda.synchronize()
    _lib.check()
    K = args.steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for n in range(K):
        step(args.warmup + n)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _lib.check()
    # Kernel durations: HIP events on the launch stream around each kernel
    # group, with the GPU held busy first (torch.cuda._sleep) so the host has
    # queued every launch of the group before the first event fires; the
    # events then time the kernels, not the host's launch gaps.
    stream = torch.cuda.current_stream()
    rec_t, src_t = [], []
    for n in range(min(K, 100)):
        m = args.warmup + n
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        torch.cuda._sleep(2_000_000)
        e[0].record(stream)
        for c in range(3):
            R.interp(traces[c, m], vel[c])
        e[1].record(stream)
        for c in range(6):
            S.add_cartesian(stress[c], mrate[c, m], h, dt)
        e[2].record(stream)
        torch.cuda.synchronize()
        rec_t.append(e[0].elapsed_time(e[1]) / 3)
        src_t.append(e[1].elapsed_time(e[2]) / 6)
    _lib.check()
    rec_ms = float(np.mean(rec_t))
    src_ms = float(np.mean(src_t))
    nb = deg + 1
    # algorithmic bytes of one launch: receivers read nb^3 field values, 3 nb
    # coefficients and 4 indices per query and write one sample; sources read
    # each touched point's position and list bounds, one code per term, each
    # query's coefficients, indices and value, and read + write the point
    rec_bytes = nrec * (4 * nb ** 3 + 12 * nb + 16 + 4)
    terms = nsrc * nb ** 3
    src_bytes = S.num_cells * (8 + 4 + 8) + terms * 4 + nsrc * (12 * nb + 16 + 4)
    rec_gbs = rec_bytes / (rec_ms * 1e-3) / 1e9
    src_gbs = src_bytes / (src_ms * 1e-3) / 1e9
    out = {
        "metric": "Lagrange receiver samples + source components per second",
        "value": (3 * nrec + 6 * nsrc) * K / wall, "unit": "query-evaluations/s",
        "n_gpus": 1, "steps": K, "warmup": args.warmup, "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"lagrange_{nx}x{ny}x{nz}", "receivers": nrec,
                   "point_sources": nsrc, "degree": deg, "h_m": h,
                   "source_grid_points": S.num_cells},
        "kernels": {
            "cuinterp_interp_H": {"avg_ms": rec_ms, "bytes": rec_bytes, "GBps": rec_gbs},
            "cusource_add_cartesian_H": {"avg_ms": src_ms, "bytes": src_bytes, "GBps": src_gbs}},
        "roofline": {"bound": "hbm", "achieved": rec_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": rec_gbs / HBM_PEAK_GBS, "traffic": _lagrange_traffic(args, nx, ny, nz),
                     "kernel": "cuinterp_interp_H"},
        "cpu_baseline": None if args.no_cpu else lagrange_cpu_baseline(args, nx, ny, nz),
        "setup_s": setup_s,
    }
    R.close()
    S.close()
    print(json.dumps(out))


def _kinematic_setup(nx, ny, nz, n, seed=3, sort_cells=False):
    """n subfaults on a vertical plane x = nx/2 (distinct cells, one per
    (y, z) cell of a rectangle 90 % of ny wide and up to 80 % of nz deep),
    rupture spreading at 2.8 km/s from the plane's centre, Brune and Liu
    source time functions alternating, rise times 0.5-1.5 s, h = 50 m."""
    import numpy as np
    h = 50.0
    rng = np.random.default_rng(seed)
    wy, wz = int(0.9 * ny), max(1, min(int(0.8 * nz), -(-n // int(0.9 * ny))))
    if n > wy * wz:
        raise ValueError(f"{n} subfaults do not fit one per cell on a {wy} x {wz} plane")
    cells = rng.choice(wy * wz, n, replace=False)
    y = cells // wz + (ny - wy) // 2 + 1
    z = cells % wz + 1
    if sort_cells:
        # subfaults in cell order: neighbouring lanes then touch neighbouring
        # cells (z fastest), so their stress updates share cache lines
        order = np.lexsort((-z, y))
        y, z = y[order], z[order]
    psrc = np.stack([np.full(n, nx // 2), y, z], 1).astype(np.int32)
    r = np.hypot((y - ny / 2) * h, (z - wz / 2) * h)
    cols = [np.zeros(2 * n, np.float32) for _ in range(6)]
    cols[0][0::2] = 1 + (np.arange(n) % 2)
    cols[1][0::2] = rng.uniform(0.5, 3.0, n)
    cols[2][0::2] = r / 2800.0
    cols[3][0::2] = rng.uniform(0.5, 1.5, n)
    cols[4][0::2] = h * h
    cols[0][1::2] = 30.0
    cols[1][1::2] = 90.0
    cols[2][1::2] = rng.uniform(150, 210, n)
    return psrc, cols, h


def bench_kinematic(args) -> None:
    """Kinematic finite-fault sources (addkinsrc_H) on the 512 x 512 x 256
    layout: --subfaults subfaults (default 80,000 when left at the lagrange
    default), 4th-order Butterworth-type IIR filter on every source time
    function, one launch per step after the rupture front has passed.
    Subfaults are sorted by cell at load time unless --subfault-order file
    (results are the same up to the order of the moment and filter-state
    entries).  Kernel durations from HIP events with the GPU held busy
    first."""
    import ctypes as C

    import numpy as np
    import torch

    from awp_amd import _lib
    from awp_amd.grid import Dims

    nx, ny, nz = _size(args, 512, 512, 256)
    n = 80000 if args.subfaults == 4000 else args.subfaults
    lib = _lib.load()
    dims = Dims(nx, ny, nz)
    DT = 0.0025
    psrc, cols, h = _kinematic_setup(nx, ny, nz, n, sort_cells=args.subfault_order == "cell")
    F, I = C.c_float, C.c_int
    z9 = (F * 9)()
    lib.SetDeviceConstValue((F * 1)(h), F(DT), (I * 1)(nx), (I * 1)(ny), (I * 1)(nz), 1, 0.0,
                            0.0, z9, z9)
    # a 4th-order low-pass (Butterworth, cut-off 0.2 of Nyquist), coefficients fixed here
    b = np.array([0.004824343357716228, 0.019297373430864913, 0.02894606014629737,
                  0.019297373430864913, 0.004824343357716228])
    a = np.array([1.0, -2.369513007182038, 2.313988414415881, -1.054665405878568,
                  0.18737949236818502])
    lib.SetDeviceFilterParameters(4, b.ctypes.data, a.ctypes.data)
    _lib.check()
    dev = [torch.zeros(dims.size, device="cuda") for _ in range(6)]
    mu = torch.full((dims.size,), 3.0e10, device="cuda")
    dps = torch.from_numpy(psrc.reshape(-1).copy()).cuda()
    dcols = [torch.from_numpy(c).cuda() for c in cols]
    mom = torch.zeros(n, device="cuda")
    filt = torch.zeros(n * 5, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    first = int(float(cols[2][0::2].max()) / DT) + 2   # every subfault has started

    def launch(step):
        lib.addkinsrc_H(step, 3, dps.data_ptr(), n, st, mu.data_ptr(),
                        *(c.data_ptr() for c in dcols), *(f.data_ptr() for f in dev),
                        mom.data_ptr(), filt.data_ptr(), 0)

    for k in range(args.warmup):
        launch(first + k)
    torch.cuda.synchronize()
    _lib.check()
    K = args.steps
    t0 = time.perf_counter()
    for k in range(K):
        launch(first + args.warmup + k)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    times = []
    stream = torch.cuda.current_stream()
    for k in range(min(K, 50)):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda._sleep(1_000_000)
        e0.record(stream)
        for r in range(4):
            launch(first + args.warmup + K + 4 * k + r)
        e1.record(stream)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) / 4)
    _lib.check()
    ms = float(np.mean(times))
    # per subfault: 8 parameters and 3 cell indices read, mu read, six stresses
    # and the moment read and written, 5 filter states read and written
    per = 8 * 4 + 3 * 4 + 4 + 2 * 6 * 4 + 2 * 4 + 2 * 5 * 8
    gbs = per * n / (ms * 1e-3) / 1e9
    print(json.dumps({
        "metric": "kinematic subfault updates per second", "value": n * K / wall,
        "unit": "subfault-updates/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"kinematic_{nx}x{ny}x{nz}", "subfaults": n, "filter_order": 4,
                   "subfault_order": args.subfault_order},
        "kernels": {"addkinsrc_H": {"avg_ms": ms, "bytes": per * n, "GBps": gbs}},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None, "kernel": "addkinsrc_H"},
        "cpu_baseline": None,
    }))
    lib.SetDeviceFilterParameters(0, None, None)


def _lagrange_traffic(args, nx, ny, nz):
    """HBM bytes per receiver launch from the committed rocprofv3 counter
    summary, for the workload it was measured on (else None)."""
    f = ROOT / "profiles" / "lagrange_512x512x256_traffic.json"
    if (nx, ny, nz, args.receivers, args.subfaults, args.degree) != (512, 512, 256, 20000, 4000,
                                                                      3) or not f.exists():
        return None
    try:
        return json.loads(f.read_text()).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def lagrange_cpu_baseline(args, nx, ny, nz) -> dict:
    """The oracle's receiver and source loops (one thread, the reference's
    per-query arithmetic) on the same workload, steps repeated for about
    --cpu-seconds / 2."""
    import numpy as np

    from oracle import interp as oi
    deg = args.degree
    dims, grid, xyz, rec, src, h = _lagrange_setup(nx, ny, nz, args.receivers, args.subfaults,
                                                   deg)
    rc = oi.lagrange_h(*xyz, *rec, deg)[1:]
    sc = oi.lagrange_h(*xyz, *src, deg)[1:]
    rng = np.random.default_rng(2)
    vel = [rng.standard_normal(dims.size, dtype=np.float32) for _ in range(3)]
    stress = [np.zeros(dims.size, np.float32) for _ in range(6)]
    out = np.zeros(args.receivers, np.float32)
    vals = rng.standard_normal(args.subfaults).astype(np.float32)
    ridx = np.arange(args.receivers)
    lidx = np.arange(args.subfaults)
    target = max(2.0, args.cpu_seconds / 2)
    n = 0
    t0 = time.perf_counter()
    while True:
        for c in range(3):
            oi.interp(out, vel[c], rc, ridx, grid)
        for c in range(6):
            oi.source_add("cartesian", stress[c], vals, sc, lidx, grid, h, 0.0025)
        n += 1
        el = time.perf_counter() - t0
        if el >= target or n >= 200:
            break
    return {"value": (3 * args.receivers + 6 * args.subfaults) * n / el,
            "unit": "query-evaluations/s", "cores": 1, "kind": "port",
            "sample": f"oracle receiver/source loops, same workload, {n} steps in {el:.1f} s, "
                      f"1 thread, cpu={_cpu_model()}"}


if __name__ == "__main__":
    main()
