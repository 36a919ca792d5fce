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


def test_topo_struct_layout_matches_c():
    """awp_topo_t: the ctypes mirror's size and member offsets equal the C
    compiler's (the header compiled on the host with g++)."""
    import shutil
    import subprocess
    import tempfile
    import ctypes as C
    from awp_amd.topography import AwpTopo
    gxx = shutil.which("g++")
    inc = Path("/opt/rocm/include")
    if gxx is None or not (inc / "hip" / "hip_runtime_api.h").exists():
        import pytest
        pytest.skip("no host compiler or HIP headers")
    src = ("#define __HIP_PLATFORM_AMD__\n#include \"awp_hip.h\"\n#include <cstdio>\n"
           "#include <cstddef>\nint main(){printf(\"%zu %zu %zu %zu %zu\\n\", sizeof(awp_topo_t),"
           " offsetof(awp_topo_t, rho), offsetof(awp_topo_t, metrics_f),"
           " offsetof(awp_topo_t, stream_i), offsetof(awp_topo_t, constants_set));}\n")
    with tempfile.TemporaryDirectory() as td:
        cpp, exe = Path(td) / "t.cpp", Path(td) / "t"
        cpp.write_text(src)
        subprocess.run([gxx, f"-I{ROOT / 'include'}", f"-I{inc}", str(cpp), "-o", str(exe)],
                       check=True, capture_output=True)
        got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True,
                                              text=True).stdout.split()]
    want = [C.sizeof(AwpTopo), AwpTopo.rho.offset, AwpTopo.metrics_f.offset,
            AwpTopo.stream_i.offset, AwpTopo.constants_set.offset]
    assert got == want


def test_velbuffer_needs_bathymetry_when_following_it():
    """velbuffer_H with FOLLOWBATHY = 1 on grid 0 reads the bathymetry array:
    a null one is an argument error, raised before any HIP call."""
    from awp_amd import _lib
    lib = _lib.load()
    F, I = ctypes.c_float, ctypes.c_int
    z9 = (F * 9)()
    lib.SetDeviceConstValue((F * 1)(100.0), F(0.005), (I * 1)(8), (I * 1)(8), (I * 1)(8), 1,
                            0.0, 0.0, z9, z9)
    lib.awp_clear_error()
    lib.velbuffer_H(*([None] * 8), 1, 0, 7, 1, 0, 7, 1, 0, 0, 1, 8, 8, 1, None, 1, None, 0)
    assert lib.awp_last_error() == 1
    assert b"bathymetry" in lib.awp_last_error_string()
    lib.awp_clear_error()
