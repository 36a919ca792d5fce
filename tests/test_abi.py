"""The C ABI: libawp_hip.so loads (no GPU needed) and exports every symbol
include/awp_hip.h declares."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "awp_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(?:void|int|const char\s*\*)\s+(\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_reference_launchers():
    names = declared_symbols()
    for n in ("SetDeviceConstValue", "dvelcx_H_opt", "dstrqc_H", "dstrqc_H_new", "fstr_H",
              "addsrc_H", "velbuffer_H"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from awp_amd import _lib
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    from awp_amd import _lib
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_error_channel_without_gpu():
    """Argument validation runs before any HIP call: an unset grid is an
    error, not a launch."""
    from awp_amd import _lib
    lib = _lib.load()
    F, I = ctypes.c_float, ctypes.c_int
    z9 = (F * 9)()
    lib.SetDeviceConstValue((F * 1)(100.0), F(0.005), (I * 1)(8), (I * 1)(8), (I * 1)(8), 1,
                            0.0, 0.0, z9, z9)
    lib.awp_clear_error()
    lib.dvelcx_H_opt(*([None] * 13), 8, 8, None, 10, 17, 9, 1)
    assert lib.awp_last_error() == 1
    assert b"d_i=9" in lib.awp_last_error_string()
    lib.awp_clear_error()
    assert lib.awp_last_error() == 0


def test_range_validation_without_gpu():
    """An i range that would read outside the padded arrays launches nothing."""
    from awp_amd import _lib
    lib = _lib.load()
    F, I = ctypes.c_float, ctypes.c_int
    z9 = (F * 9)()
    lib.SetDeviceConstValue((F * 1)(100.0), F(0.005), (I * 1)(8), (I * 1)(8), (I * 1)(8), 1,
                            0.0, 0.0, z9, z9)
    lib.awp_clear_error()
    lib.dvelcx_H_opt(*([None] * 13), 8, 8, None, 1, 17, 0, 1)   # s_i - 2 < 0
    assert lib.awp_last_error() == 1
    lib.awp_clear_error()
    lib.dvelcx_H_opt(*([None] * 13), 8, 8, None, 10, 26, 0, 1)  # e_i + 2 >= nxt + 20
    assert lib.awp_last_error() == 1
    lib.awp_clear_error()
    lib.dvelcx_H_opt(*([None] * 13), 9, 8, None, 10, 17, 0, 1)  # nyt disagrees
    assert lib.awp_last_error() == 1
    lib.awp_clear_error()


def _header_arg_counts():
    text = (ROOT / "include" / "awp_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(?:void|int|const char\s*\*)\s+(\w+)\s*\(([^)]*)\)", text):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_binding_arity_matches_header():
    from awp_amd import _lib
    for name, n in _header_arg_counts().items():
        assert len(_lib.SIGNATURES[name]) == n, (name, len(_lib.SIGNATURES[name]), n)


_REF_TYPES = {"int": "int", "_prec": "float", "prec": "float", "_prec *": "float *",
              "prec *": "float *", "int *": "int *", "dim3": "ref_dim3",
              "cudaStream_t": "void *", "const char *": "const char *", "size_t": "size_t",
              "int3_t": "ref_int3_t", "f_grid_t": "ref_f_grid_t", "g_grid_t": "ref_g_grid_t",
              "grid3_t": "ref_fcn_grid_t"}
# reference struct -> (this header's struct, ctypes mirror in awp_amd.refstructs)
_STRUCTS = {"int3_t": ("int3_t", "Int3"), "fcn_grid_t": ("fcn_grid_t", "Grid3"),
            "f_grid_t": ("f_grid_t", "FGrid"), "g_grid_t": ("g_grid_t", "GGrid"),
            "topo_t": ("topo_t", "Topo"), "cu_interp_t": ("cu_interp_t", "CuInterp")}


def _layouts_c(members):
    """A C++ program that defines the reference's structs from their member
    lists (tests/golden/ref_layouts.json, extracted from the reference's
    headers) beside this library's header and prints, per struct, the size
    and every member offset of both."""
    defs = ["typedef struct { unsigned x, y, z; } ref_dim3;"]
    for name in ("int3_t", "fcn_grid_t", "f_grid_t", "g_grid_t", "topo_t", "cu_interp_t"):
        body = "".join(f"  {_REF_TYPES[t]} {n}{f'[{a}]' if a else ''};\n"
                       for t, n, a in members[name])
        defs.append(f"typedef struct {{\n{body}}} ref_{name};")
    lines = []
    for ref, (ours, _) in _STRUCTS.items():
        lines.append(f'printf("%s sizeof %zu %zu\\n", "{ref}", sizeof(ref_{ref}), sizeof({ours}));')
        for _, n, _ in members[ref]:
            lines.append(f'printf("%s %s %zu %zu\\n", "{ref}", "{n}", offsetof(ref_{ref}, {n}),'
                         f' offsetof({ours}, {n}));')
    return ("#define __HIP_PLATFORM_AMD__\n#include \"awp_hip.h\"\n#include <cstdio>\n"
            "#include <cstddef>\n" + "\n".join(defs) + "\nint main(){\n" + "\n".join(lines)
            + "\n}\n")


def test_reference_struct_layouts():
    """topo_t, f_grid_t, g_grid_t, grid3_t, int3_t and cu_interp_t of the
    header have the reference's size and every member at the reference's
    offset (the reference's member lists compiled side by side with the
    header by the host compiler), and the ctypes mirrors agree with both."""
    import json
    import shutil
    import subprocess
    import tempfile
    import ctypes as C
    from awp_amd import refstructs
    gxx = shutil.which("g++")
    inc = Path("/opt/rocm/include")
    if gxx is None or not (inc / "hip" / "hip_runtime_api.h").exists():
        import pytest
        pytest.skip("no host compiler or HIP headers")
    members = json.loads((ROOT / "tests" / "golden" / "ref_layouts.json").read_text())
    with tempfile.TemporaryDirectory() as td:
        cpp, exe = Path(td) / "t.cpp", Path(td) / "t"
        cpp.write_text(_layouts_c(members))
        subprocess.run([gxx, f"-I{ROOT / 'include'}", f"-I{inc}", str(cpp), "-o", str(exe)],
                       check=True, capture_output=True)
        out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    seen = 0
    for line in out.splitlines():
        ref, name, a, b = line.split()
        assert a == b, (ref, name, a, b)
        mirror = getattr(refstructs, _STRUCTS[ref][1])
        want = C.sizeof(mirror) if name == "sizeof" else getattr(mirror, name).offset
        assert int(a) == want, (ref, name, a, want)
        seen += 1
    assert seen == sum(len(v) + 1 for v in members.values())
