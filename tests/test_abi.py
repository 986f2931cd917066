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
    "athx_instance": _lib.Instance,
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


# modules/core/include/athenea/core/Error.h:14-26: enum class ErrorCode, in
# declaration order, so ErrorCode::Unknown is 0 and InternalError 10.
ERROR_CODE_ORDER = ["Unknown", "NotFound", "InvalidArgument", "Unsupported", "IoFailure", "DeviceFailure",
                    "ShaderFailure", "PluginFailure", "OutOfMemory", "Cancelled", "InternalError"]


def test_status_codes_map_to_athenea_error_codes():
    """athx_status is 0 for success and ErrorCode + 1 otherwise, in the
    reference enum's order; athx_error_code() gives back the ErrorCode value."""
    text = HEADER.read_text()
    values = dict((name, int(v)) for name, v in re.findall(r"ATHX_(\w+) = (\d+)", text) if not name.startswith("CULL"))
    header_names = ["UNKNOWN", "NOT_FOUND", "INVALID_ARGUMENT", "UNSUPPORTED", "IO_FAILURE", "DEVICE_FAILURE",
                    "SHADER_FAILURE", "PLUGIN_FAILURE", "OUT_OF_MEMORY", "CANCELLED", "INTERNAL_ERROR"]
    assert values["OK"] == 0
    lib = _lib.load()
    for code, (hname, rname) in enumerate(zip(header_names, ERROR_CODE_ORDER)):
        status = values[hname]
        assert status == code + 1, hname
        assert lib.athx_error_code(status) == code, hname
        assert _lib.STATUS_NAMES[status] == rname
        assert _lib.AthxError(status, "").error_code == code
    assert lib.athx_error_code(0) == -1 and lib.athx_error_code(12) == -1 and lib.athx_error_code(-3) == -1


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


def test_contribution_refuses_wrong_per_pixel_inputs():
    """TileRasterizer.contribution checks each per-pixel input before the
    library reads it: mode 2 and 4 read [h, w, 4], mode 3 [h, w], floating
    point, on the cloud's device. No device is needed: every refusal comes
    before any call into the library."""
    torch = pytest.importorskip("torch")
    from athenea_amd import camera, synthetic
    from athenea_amd.raster import GaussianCloud, RenderTargets, TileRasterizer

    raster = object.__new__(TileRasterizer)  # no context: nothing may reach the library
    raster._lib = None
    raster._ctx = None
    view = camera.view_for(synthetic.front_camera(), 8, 6)
    cloud = GaussianCloud(torch.zeros(2, 3), torch.zeros(2, 3), torch.zeros(2, 4), torch.zeros(2),
                          torch.zeros(2, 1, 3))
    targets = [RenderTargets.allocate(8, 6, "cpu", True)]
    bad = {
        4: [torch.zeros(6, 8), torch.zeros(6, 8, 4, dtype=torch.int32), torch.zeros(6, 8, 3), None],
        2: [torch.zeros(6, 8)],
        3: [torch.zeros(6, 8, 4), torch.zeros(8, 6)],
    }
    for mode, inputs in bad.items():
        for t in inputs:
            with pytest.raises(ValueError):
                raster.contribution([view], cloud, None, targets, mode=mode, dl_dcolour=[t])
    with pytest.raises(ValueError):
        raster.contribution([view], cloud, None, targets, mode=4, dl_dcolour=None)
    with pytest.raises(ValueError):
        raster.contribution([view], cloud, None, targets, mode=5)
    # The right shape on another device than the cloud's.
    meta = GaussianCloud(*(t.to("meta") for t in (cloud.means, cloud.log_scales, cloud.quats,
                                                   cloud.opacity_logits, cloud.sh)))
    meta_targets = [RenderTargets(8, 6, torch.zeros(6, 8, 4, device="meta"), torch.zeros(6, 8, device="meta"))]
    with pytest.raises(ValueError, match="must be on"):
        raster.contribution([view], meta, None, meta_targets, mode=4, dl_dcolour=[torch.zeros(6, 8, 4)])
