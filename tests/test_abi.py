"""The C ABI: libathx.so exports every entry point include/athx.h declares,
the ctypes mirror has the header's struct layouts, and calls fail cleanly
where there is no GPU. No compute is run here."""
from __future__ import annotations

import ctypes
import re
import subprocess
import tempfile
from pathlib import Path

import pytest

from athenea_amd import _lib

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "athx.h"


def declared_symbols() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"ATHX_API\s+[^;(]*?\b(athx_\w+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_symbols()
    assert "athx_render" in names and "athx_render_backward" in names and "athx_sort_pairs" in names
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures and header disagree"


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (athx_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_library_exports_nothing_else():
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    ours = {s for s in re.findall(r"\bT (\w+)", out) if not s.startswith("_")}
    assert ours == set(declared_symbols()), ours ^ set(declared_symbols())


STRUCTS = {
    "athx_view": _lib.View,
    "athx_gaussians": _lib.Gaussians,
    "athx_sky": _lib.Sky,
    "athx_relight": _lib.Relight,
    "athx_packed_splats": _lib.PackedSplats,
    "athx_settings": _lib.Settings,
    "athx_targets": _lib.Targets,
    "athx_frame_stats": _lib.FrameStats,
    "athx_counters": _lib.Counters,
    "athx_grads": _lib.Grads,
    "athx_frame_info": _lib.FrameInfo,
    "athx_ply_layout": _lib.PlyLayout,
    "athx_ingest_out": _lib.IngestOut,
    "athx_image_difference": _lib.ImageDifference,
    "athx_hdr_difference": _lib.HdrDifference,
}


def test_ctypes_layouts_match_the_header():
    """Compiles a probe against include/athx.h with gcc and compares every
    struct's size and every field's offset with the ctypes mirror."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "athx.h"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'printf("{cname}.{field} %zu\\n", offsetof({cname}, {field}));')
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src = Path(d) / "probe.c"
        exe = Path(d) / "probe"
        src.write_text("\n".join(lines))
        subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
        out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    seen = {}
    for line in out.splitlines():
        key, value = line.rsplit(" ", 1)
        seen[key] = int(value)
    for cname, py in STRUCTS.items():
        assert seen[f"{cname} size"] == ctypes.sizeof(py), cname
        for field, _ in py._fields_:
            assert seen[f"{cname}.{field}"] == getattr(py, field).offset, (cname, field)


def test_status_codes_are_athenea_error_codes():
    """include/athx.h's athx_status is modules/core/include/athenea/core/
    Error.h:14-26's ErrorCode, value for value, with 0 for success."""
    text = HEADER.read_text()
    values = dict((name, int(v)) for name, v in re.findall(r"ATHX_(\w+) = (\d+)", text) if not name.startswith("CULL"))
    expected = ["OK", "UNKNOWN", "NOT_FOUND", "INVALID_ARGUMENT", "UNSUPPORTED", "IO_FAILURE", "DEVICE_FAILURE",
                "SHADER_FAILURE", "PLUGIN_FAILURE", "OUT_OF_MEMORY", "CANCELLED", "INTERNAL_ERROR"]
    assert [values[e] for e in expected] == list(range(12))
    assert [_lib.STATUS_NAMES[i] for i in range(12)][3] == "InvalidArgument"


def test_calls_fail_cleanly_without_a_device():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _lib.load()
    h = ctypes.c_void_p()
    status = lib.athx_create(0, ctypes.byref(h))
    assert status == 6 and not h.value  # DeviceFailure, no context
    assert lib.athx_render(None, None, 0, None, None, None, None, None) == 3  # InvalidArgument
    assert lib.athx_last_error(None) == b"no context"
    lib.athx_destroy(None)


def test_sky_texel_count_is_the_layouts():
    """athx_sky_texels (host only) is environment.slang:66-92's dome size."""
    from athenea_amd import sky
    lib = _lib.load()
    for base in (16, 100, 256, 300, 512, 1024):
        assert lib.athx_sky_texels(base) == sky.texels_per_dome(base)
