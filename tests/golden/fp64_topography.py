"""FP64 reference for the curvilinear-topography workload (384x384x192 cells,
h = 100 m, six seeded Gaussian hills, layered + random medium with Q, 1500
steps) and the receiver-trace errors of GPU runs against it.

The FP64 oracle takes about an hour on 6 host threads.  `run` wrote the
fixture topography_384x384x192_fp64_traces.npz beside this script (traces of
16 surface receivers x 3 components, 1501 samples), which the GPU tests read.
Usage:

    python tests/golden/fp64_topography.py run --threads 6 --out fp64_traces.npz
    python bench.py --workload topo --storage f32 --steps 1500 --warmup 0 --save-traces f32.npy
    python tests/golden/fp64_topography.py compare fp64_traces.npz f32.npy bf16.npy fp16.npy

`compare` prints, per GPU trace file, the relative L2 error over all
receivers, components and samples, and the largest and median per
receiver-component relative L2 error.
"""
import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[2]))


def run(args) -> None:
    from awp_amd import configs
    from oracle.oracle import OracleSolver, set_threads
    set_threads(args.threads)
    w = configs.topography_hills(nx=384, ny=384, nz=192, nt=args.nt)
    med = w.medium()
    o = OracleSolver(med, w.sources, w.receivers, nt_record=args.nt + 1, dtype=np.float64,
                     topography=w.params["topography"])
    del med
    o.start()
    t0 = time.time()
    for n in range(args.nt):
        o.step()
        if (n + 1) % 50 == 0:
            print(n + 1, round(time.time() - t0, 1), flush=True)
    np.savez_compressed(args.out, traces=o.traces())


def compare(args) -> None:
    ref = np.load(args.reference)["traces"]
    out = {}
    for path in args.traces:
        g = np.load(path).astype(np.float64)
        n = min(len(g), len(ref))
        d = g[:n] - ref[:n]
        per = np.linalg.norm(d, axis=0) / np.maximum(np.linalg.norm(ref[:n], axis=0), 1e-300)
        out[path] = {"samples": n,
                     "rel_l2": float(np.linalg.norm(d) / np.linalg.norm(ref[:n])),
                     "rel_l2_receiver_component_max": float(per.max()),
                     "rel_l2_receiver_component_median": float(np.median(per))}
    print(json.dumps(out, indent=1))


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--threads", type=int, default=6)
    r.add_argument("--nt", type=int, default=1500)
    r.add_argument("--out", default="fp64_traces.npz")
    c = sub.add_parser("compare")
    c.add_argument("reference")
    c.add_argument("traces", nargs="+")
    args = ap.parse_args()
    run(args) if args.cmd == "run" else compare(args)


if __name__ == "__main__":
    main()
