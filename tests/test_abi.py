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


def test_headers_compile_as_c_and_struct_sizes_match_the_bindings(tmp_path):
    """include/*.h compile as C99, and every struct the Python binding mirrors
    has the same size on both sides of the boundary."""
    import shutil
    import subprocess
    from bevy_penumbra_amd import lux
    gcc = shutil.which("gcc")
    if gcc is None:
        import pytest
        pytest.skip("no C compiler")
    src = tmp_path / "sizes.c"
    src.write_text(
        '#include <stdio.h>\n#include "lux.h"\n#include "lux_bevy.h"\n'
        "int main(void) {\n"
        '  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(lux_ray), sizeof(lux_hit),'
        " sizeof(lux_gi_params), sizeof(lux_surrogate_params), sizeof(lux_gi_surrogate_stats),"
        " sizeof(lux_trace_config), sizeof(lux_instance_desc), sizeof(lux_material),"
        " sizeof(lux_radiance_cascades), sizeof(lux_lod_view));\n"
        "  return 0;\n}\n")
    exe = tmp_path / "sizes"
    subprocess.run([gcc, "-std=c99", "-Wall", "-Werror", "-I", str(ROOT / "include"), str(src), "-o", str(exe)],
                   check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = [32, 16, ctypes.sizeof(lux.GiParams), ctypes.sizeof(lux.SurrogateParams),
            ctypes.sizeof(lux.SurrogateStats), 32, ctypes.sizeof(lux.InstanceDesc), 32,
            ctypes.sizeof(lux.RadianceCascades), ctypes.sizeof(lux.LodView)]
    assert got == want
