"""Member lists of the reference's structs that cross the topography and
Lagrange launcher boundary, for the layout tests.

Reads the struct definitions from the reference's headers
(include/awp/definitions.h: int3_t; include/grid/grid_3d.h: fcn_grid_t =
grid3_t; include/topography/metrics/metrics.h: f_grid_t, g_grid_t;
include/topography/topography.h: topo_t; include/interpolation/
interpolation.cuh: cu_interp_t) and writes ref_layouts.json beside this
script: for each struct, its members in order as [type, name, array length
or 0].  Types are kept as the reference spells them (_prec, prec, dim3,
cudaStream_t, size_t, int, const char *, pointer stars attached to the type).

    python tests/golden/ref_layouts.py [/root/reference]
"""
import json
import re
import sys
from pathlib import Path

SOURCES = {
    "int3_t": "include/awp/definitions.h",
    "fcn_grid_t": "include/grid/grid_3d.h",
    "f_grid_t": "include/topography/metrics/metrics.h",
    "g_grid_t": "include/topography/metrics/metrics.h",
    "topo_t": "include/topography/topography.h",
    "cu_interp_t": "include/interpolation/interpolation.cuh",
}


def struct_body(text: str, name: str) -> str:
    """Body of `typedef struct { ... } name;`."""
    for m in re.finditer(r"typedef\s+struct\s*\{", text):
        depth, i = 1, m.end()
        while depth:
            depth += {"{": 1, "}": -1}.get(text[i], 0)
            i += 1
        tail = re.match(r"\s*(\w+)\s*;", text[i:])
        if tail and tail.group(1) == name:
            return text[m.end():i - 1]
    raise KeyError(name)


def members(body: str) -> list:
    body = re.sub(r"//[^\n]*", "", body)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    out = []
    for decl in body.split(";"):
        decl = " ".join(decl.replace("__restrict__", "").split())
        if not decl:
            continue
        m = re.match(r"((?:const\s+)?\w+)\s*(.*)$", decl)
        base, rest = m.group(1), m.group(2)
        for part in rest.split(","):
            part = part.strip()
            stars = len(part) - len(part.lstrip("*"))
            part = part.lstrip("*").strip()
            am = re.match(r"(\w+)\s*(?:\[(\d+)\])?$", part)
            out.append([base + " " + "*" * stars if stars else base, am.group(1),
                        int(am.group(2)) if am.group(2) else 0])
    return out


def main(ref: str) -> None:
    out = {}
    for name, rel in SOURCES.items():
        text = (Path(ref) / rel).read_text()
        out[name] = members(struct_body(text, name))
    dst = Path(__file__).resolve().parent / "ref_layouts.json"
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
