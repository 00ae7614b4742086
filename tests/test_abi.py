"""The C-ABI library loads and exports every symbol include/*.h declares.

No compute call is made here: this runs without a GPU.
"""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(lux_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    for required in ("lux_ctx_create", "lux_trace_rays", "lux_bvh_build", "lux_scene_upload_soa"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from bevy_penumbra_amd import lux
    lib = ctypes.CDLL(str(lux.LIB_PATH))
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_cover_header():
    from bevy_penumbra_amd import lux
    lib = lux.load_library()
    for n in declared_symbols():
        assert getattr(lib, n).argtypes is not None, n
