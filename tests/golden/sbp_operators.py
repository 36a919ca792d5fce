"""Extract the right-boundary blocks of the staggered SBP operators that the
reference's kernel generator composes into the curvilinear velocity kernel's
top-row tables.

The generator (tools/kernel_generation/scheme.py:150-163, helper.py:154-189)
builds the z-derivative of an interpolant, D(P(.), interp="first"), from the
coefficient files under tools/OpenFD/openfd/sbp/resources/build3_all/:
DmPp_bc42 / DpPm_bc42 (the composites the kernel tables hold), and the
factors Dm_bc42, Dp_bc42 (one-sided derivatives with the free-surface
closure), Pp42, Pm42 (interpolations between the node and cell-centre
grids).  This script copies the `op_right` block and interior stencil of
each into sbp_right_blocks.json beside it; the CPU tests pin the kernel
tables against these numbers.

    python tests/golden/sbp_operators.py [/root/reference]
"""
import json
import sys
from pathlib import Path

NAMES = ("Dm_bc42", "Dp_bc42", "Pp42", "Pm42", "DmPp_bc42", "DpPm_bc42")
RES = "tools/OpenFD/openfd/sbp/resources/build3_all"


def main(ref: str) -> None:
    out = {}
    for n in NAMES:
        d = json.loads((Path(ref) / RES / f"{n}.json").read_text())
        out[n] = {"op_right": d["op_right"], "idx_right": d["idx_right"],
                  "op_interior": d["op_interior"], "idx_interior": d["idx_interior"]}
    dst = Path(__file__).resolve().parent / "sbp_right_blocks.json"
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
