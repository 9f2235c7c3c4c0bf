"""The C-ABI libraries load and export every entry point their headers declare (no compute calls)."""

import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: Path) -> list[str]:
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(vbd_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_hip_library_exports_header_symbols():
    names = _declared(ROOT / "include" / "vbd_hip.h")
    assert "vbd_step" in names and "vbd_create" in names and "vbd_collide_generate" in names
    lib = ctypes.CDLL(str(ROOT / "newton_amd" / "lib" / "libvbd_hip.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"libvbd_hip.so lacks {missing}"


def test_ctypes_table_matches_header():
    from newton_amd import _native

    names = set(_declared(ROOT / "include" / "vbd_hip.h"))
    assert names == set(_native.HIP_EXPORTS), (names ^ set(_native.HIP_EXPORTS))


def test_struct_layouts_match_header():
    """Field order of the ctypes mirrors equals the header's struct field order."""
    from newton_amd import _native

    text = (ROOT / "include" / "vbd_hip.h").read_text()
    for cname, pyname in (("VbdModelDesc", "VbdModelDesc"), ("VbdSolverParams", "VbdSolverParams"),
                          ("VbdContactsDesc", "VbdContactsDesc"), ("VbdContactsOut", "VbdContactsOut"),
                          ("VbdCollideDesc", "VbdCollideDesc")):
        body = re.search(r"typedef struct " + cname + r" \{(.*?)\} " + cname + ";", text, flags=re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        fields = re.findall(r"(\w+)\s*;", body)
        assert fields == [f[0] for f in getattr(_native, pyname)._fields_], cname


def test_host_library_exports():
    from newton_amd import _native

    lib = _native.host_lib()
    for name in _native.HOST_EXPORTS:
        assert hasattr(lib, name)


def test_hip_library_loads_without_gpu_calls():
    from newton_amd import _native

    lib = _native.hip_lib()
    assert lib.vbd_version() >= 1
    assert lib.vbd_last_error() is not None


def _declared_prefix(header: Path, prefix: str) -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(" + prefix + r"\w+)\s*\(", text, flags=re.M)))


def test_warmstart_library_exports_header_symbols():
    from newton_amd import _native

    names = _declared_prefix(ROOT / "include" / "vbd_warmstart.h", "vbd_ws_")
    assert "vbd_ws_forward" in names and "vbd_ws_backward" in names and "vbd_ws_adam" in names
    lib = ctypes.CDLL(str(ROOT / "newton_amd" / "lib" / "libvbd_ws.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"libvbd_ws.so lacks {missing}"
    assert set(names) == set(_native.WS_EXPORTS)


def test_warmstart_sizes_without_gpu():
    from newton_amd import _native

    import mlp_oracle

    out = (ctypes.c_int64 * 3)()
    assert _native.ws_lib().vbd_ws_sizes(out) == 0
    assert out[0] == mlp_oracle.N_PARAMS == 39683
    assert out[1] == 81920 and out[2] == 64


def test_host_adjacency_exports():
    from newton_amd import _native

    lib = _native.host_lib()
    for name in ("vbd_host_edge_tri_maps", "vbd_host_vertex_csr", "vbd_host_color_graph"):
        assert hasattr(lib, name)
