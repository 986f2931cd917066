"""TEST INFRASTRUCTURE: numpy bindings of oracle.c, the CPU restatement of
athenea's splat tile rasterizer that the HIP path is checked against.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker (or, for the baseline, as the thing timed on
the host cores). The product (athenea_amd) never imports it.

``render`` runs the whole frame the way the HIP path does (project, scan,
emit with 64-bit keys in slot order, stable LSD sort, ranges, blend);
``render_ref_order`` runs the reference's own two-sort order
(TileRasterizer.cpp:1111-1176), which must give the same pairs and ranges.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


class OrcView(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("focal_x", ctypes.c_float),
        ("focal_y", ctypes.c_float),
        ("centre_x", ctypes.c_float),
        ("centre_y", ctypes.c_float),
        ("near_z", ctypes.c_float),
        ("far_z", ctypes.c_float),
        ("orthographic", ctypes.c_uint32),
        ("antialias", ctypes.c_uint32),
        ("linear", ctypes.c_uint32),
        ("sh_limit", ctypes.c_uint32),
        ("tile", ctypes.c_uint32),
        ("m", ctypes.c_float * 12),
        ("eye", ctypes.c_float * 3),
        ("bg", ctypes.c_float * 4),
        ("depth_mode", ctypes.c_uint32),
        ("depth_threshold", ctypes.c_float),
    ]


class OrcRelight(ctypes.Structure):
    _fields_ = [
        ("normals", ctypes.c_void_p),
        ("albedo", ctypes.c_void_p),
        ("metallic", ctypes.c_void_p),
        ("roughness", ctypes.c_void_p),
        ("o2w", ctypes.c_float * 12),
        ("texels", ctypes.c_void_p),
        ("sky_sh", ctypes.c_void_p),
        ("sky_sun", ctypes.c_void_p),
        ("base_side", ctypes.c_uint32),
        ("normals_packed", ctypes.c_void_p),
        ("pbr", ctypes.c_void_p),
    ]


def orc_relight(relight) -> tuple[OrcRelight, list]:
    """The oracle's relight inputs from a numpy-side description with
    attributes normals (n, 3) or None, albedo (n, 3) or None, metallic (n,),
    roughness (n,), object_to_world (12,), and sky with texels (uint32),
    sh (16, 4), sun (2, 4), base_side. Returns the struct and the arrays it
    points into (keep them alive while it is used)."""
    keep = []

    def ptr(a, dtype):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return a.ctypes.data

    r = OrcRelight()
    r.normals = ptr(relight.normals, np.float32)
    r.albedo = ptr(relight.albedo, np.float32)
    r.metallic = ptr(relight.metallic, np.float32)
    r.roughness = ptr(relight.roughness, np.float32)
    for k in range(12):
        r.o2w[k] = float(relight.object_to_world[k])
    r.texels = ptr(relight.sky.texels, np.uint32)
    r.sky_sh = ptr(relight.sky.sh, np.float32)
    r.sky_sun = ptr(relight.sky.sun, np.float32)
    r.base_side = int(relight.sky.base_side)
    r.normals_packed = ptr(getattr(relight, "normals_packed", None), np.uint32)
    r.pbr = ptr(getattr(relight, "pbr", None), np.uint32)
    return r, keep


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        sources = [HERE / "oracle.c", HERE.parent / "include" / "athx_detmath.h"]
        if any(src.stat().st_mtime > LIB_PATH.stat().st_mtime + 2.0 for src in sources):
            raise RuntimeError(f"{LIB_PATH} is older than its sources: rebuild it (make -C oracle)")
        lib = ctypes.CDLL(str(LIB_PATH))
        vp = ctypes.POINTER(OrcView)
        rp = ctypes.POINTER(OrcRelight)
        f3 = ctypes.c_float * 3
        lib.orc_project_lit.argtypes = [vp, _f32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32,
                                        _f32p, _u32p, _u32p, ctypes.c_void_p, rp]
        lib.orc_project_lit.restype = None
        lib.orc_project_packed.argtypes = [vp, _f32p, _u32p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                           _f32p, _u32p, _u32p, ctypes.c_void_p, rp]
        lib.orc_project_packed.restype = None
        lib.orc_unpack_normal.argtypes = [ctypes.c_uint32, f3]
        lib.orc_unpack_normal.restype = None
        lib.orc_decode_quaternion.argtypes = [ctypes.c_uint32, ctypes.c_float * 4]
        lib.orc_decode_quaternion.restype = None
        lib.orc_relit_point.argtypes = [rp, f3, f3, f3, ctypes.c_float, ctypes.c_float, f3]
        lib.orc_relit_point.restype = None
        lib.orc_env_radiance.argtypes = [rp, f3, ctypes.c_float, f3]
        lib.orc_env_radiance.restype = None
        lib.orc_env_irradiance.argtypes = [rp, f3, f3]
        lib.orc_env_irradiance.restype = None
        lib.orc_sky_sun.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_uint32, _f32p]
        lib.orc_sky_sun.restype = ctypes.c_int
        lib.orc_sky_residual.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p]
        lib.orc_sky_residual.restype = ctypes.c_int
        lib.orc_env_level_offset.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        lib.orc_env_level_offset.restype = ctypes.c_uint32
        lib.orc_half_to_float.argtypes = [ctypes.c_uint32]
        lib.orc_half_to_float.restype = ctypes.c_float
        lib.orc_ggx_albedo.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.POINTER(ctypes.c_float),
                                       ctypes.POINTER(ctypes.c_float)]
        lib.orc_ggx_albedo.restype = None
        lib.orc_ggx_dielectric.argtypes = [ctypes.c_float, ctypes.c_float]
        lib.orc_ggx_dielectric.restype = ctypes.c_float
        lib.orc_ggx_conductor.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
        lib.orc_ggx_conductor.restype = ctypes.c_float
        lib.orc_fresnel_dielectric.argtypes = [ctypes.c_float, ctypes.c_float]
        lib.orc_fresnel_dielectric.restype = ctypes.c_float
        lib.orc_fresnel_conductor.argtypes = [ctypes.c_float, ctypes.c_float]
        lib.orc_fresnel_conductor.restype = ctypes.c_float
        lib.orc_project.argtypes = [vp, _f32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32, _f32p,
                                    _u32p, _u32p, ctypes.c_void_p]
        lib.orc_project.restype = None
        lib.orc_exclusive_scan.argtypes = [_u32p, _u32p, ctypes.c_uint64]
        lib.orc_exclusive_scan.restype = ctypes.c_uint64
        lib.orc_emit_pairs.argtypes = [vp, _f32p, _u32p, _u32p, _u32p, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_uint32, _u64p, _u32p]
        lib.orc_emit_pairs.restype = None
        lib.orc_sort_pairs.argtypes = [_u64p, _u32p, ctypes.c_uint64, ctypes.c_uint32]
        lib.orc_sort_pairs.restype = None
        lib.orc_ranges.argtypes = [_u64p, ctypes.c_uint64, ctypes.c_uint32, _u32p]
        lib.orc_ranges.restype = None
        lib.orc_blend.argtypes = [vp, _f32p, _u32p, _u32p, ctypes.c_uint32, _f32p, _f32p, _f32p, _u32p]
        lib.orc_blend.restype = None
        lib.orc_blend_ex.argtypes = [vp, _f32p, _u32p, _u32p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p, _f32p, _f32p, _f32p, _u32p, ctypes.c_void_p,
                                     ctypes.c_void_p]
        lib.orc_blend_ex.restype = None
        lib.orc_blend_f64.argtypes = [vp, _f32p, _u32p, _u32p, _u32p, _u32p, ctypes.c_uint32, _f64p, _f64p, _f64p,
                                      _f64p, _f64p, ctypes.c_uint32, ctypes.c_uint32, _f64p, _f64p]
        lib.orc_blend_f64.restype = None
        lib.orc_reference.argtypes = [vp, _f32p, _u32p, ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p]
        lib.orc_reference.restype = ctypes.c_uint64
        lib.orc_render_ref_order.argtypes = [vp, _f32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32,
                                             _f32p, _u32p, _u32p, _u32p, ctypes.c_uint64, _u32p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, rp]
        lib.orc_render_ref_order.restype = ctypes.c_uint64
        lib.orc_threads.restype = ctypes.c_int
        lib.orc_set_threads.argtypes = [ctypes.c_int]
        lib.orc_expf.argtypes = [ctypes.c_float]
        lib.orc_expf.restype = ctypes.c_float
        lib.orc_logf.argtypes = [ctypes.c_float]
        lib.orc_logf.restype = ctypes.c_float
        lib.orc_ingest.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, ctypes.c_uint32,
                                   ctypes.c_uint32, _f32p, _u32p, ctypes.c_void_p, _u32p]
        lib.orc_ingest.restype = ctypes.c_uint32
        lib.orc_ingest_modes.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p,
                                         ctypes.c_uint32, ctypes.c_uint32, _u32p, _f32p, _u32p, ctypes.c_void_p,
                                         _u32p]
        lib.orc_ingest_modes.restype = ctypes.c_uint32
        lib.orc_ref_project.argtypes = [vp, _f32p, _u32p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                        _f32p]
        lib.orc_compare_codes.argtypes = [_f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32, _u32p]
        lib.orc_compare_hdr.argtypes = [_f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32, _u32p]
        lib.orc_compare_hdr.restype = ctypes.c_float
        lib.orc_contribution.argtypes = [vp, _f32p, _u32p, _u32p, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_void_p, _u32p]
        lib.orc_half_bits.argtypes = [ctypes.c_float]
        lib.orc_half_bits.restype = ctypes.c_uint32
        lib.orc_encode_quaternion.argtypes = [ctypes.c_float] * 4
        lib.orc_encode_quaternion.restype = ctypes.c_uint32
        _lib = lib
    return _lib


def orc_view(view, tile: int = 16, antialias: bool = True, linear: bool = True, sh_limit: int = 3,
             depth_mode: int = 0, depth_threshold: float = 0.5) -> OrcView:
    """The oracle's frame from an athx View (athenea_amd._lib.View) and settings."""
    v = OrcView()
    v.width, v.height = view.width, view.height
    v.focal_x, v.focal_y = view.focal_x, view.focal_y
    v.centre_x, v.centre_y = view.centre_x, view.centre_y
    v.near_z, v.far_z = view.near_z, view.far_z
    v.orthographic = view.orthographic
    v.antialias = 1 if antialias else 0
    v.linear = 1 if linear else 0
    v.sh_limit = sh_limit
    v.tile = tile
    for k in range(12):
        v.m[k] = view.object_to_view[k]
    for k in range(3):
        v.eye[k] = view.eye[k]
    for k in range(4):
        v.bg[k] = view.background[k]
    v.depth_mode = depth_mode
    v.depth_threshold = depth_threshold
    return v


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def tiles_of(v: OrcView) -> tuple[int, int]:
    return (v.width + v.tile - 1) // v.tile, (v.height + v.tile - 1) // v.tile


def bits_for(values: int) -> int:
    bits = 1
    while bits < 32 and (1 << bits) < values:
        bits += 1
    return bits


def project(v: OrcView, cloud, flags: bool = False, relight=None):
    """Records, tile counts and depth keys of every Gaussian; ``relight``
    (see orc_relight) swaps the harmonics' colour for the relit one."""
    lib = load()
    rl, keep = orc_relight(relight) if relight is not None else (None, None)
    if hasattr(cloud, "shape") and hasattr(cloud, "positions"):
        # athenea's packed layout (athenea_amd.packing.PackedCloud)
        n = cloud.positions.shape[0]
        rec = np.zeros((n, 12), np.float32)
        touched = np.zeros(n, np.uint32)
        keys = np.zeros(n, np.uint32)
        fl = np.zeros(n, np.uint32) if flags else None
        words = _c(cloud.sh, np.uint32) if cloud.sh is not None else None
        lib.orc_project_packed(ctypes.byref(v), _c(cloud.positions, np.float32), _c(cloud.shape, np.uint32),
                               words.ctypes.data if words is not None else None, cloud.rest_per_colour, n, rec,
                               touched, keys, fl.ctypes.data if flags else None,
                               ctypes.byref(rl) if rl is not None else None)
        del keep
        return (rec, touched, keys, fl) if flags else (rec, touched, keys)
    n = cloud.means.shape[0]
    coeffs = cloud.sh.shape[1]
    rec = np.zeros((n, 12), np.float32)
    touched = np.zeros(n, np.uint32)
    keys = np.zeros(n, np.uint32)
    fl = np.zeros(n, np.uint32) if flags else None
    lib.orc_project_lit(ctypes.byref(v), _c(cloud.means, np.float32), _c(cloud.log_scales, np.float32),
                        _c(cloud.quats, np.float32), _c(cloud.opacity_logits, np.float32), _c(cloud.sh, np.float32),
                        coeffs, n, rec, touched, keys, fl.ctypes.data if flags else None,
                        ctypes.byref(rl) if rl is not None else None)
    del keep
    return (rec, touched, keys, fl) if flags else (rec, touched, keys)


def exclusive_scan(values: np.ndarray) -> tuple[np.ndarray, int]:
    values = _c(values, np.uint32)
    out = np.zeros_like(values)
    total = load().orc_exclusive_scan(values, out, values.size)
    return out, int(total)


def sort_pairs(keys: np.ndarray, values: np.ndarray, bits: int) -> tuple[np.ndarray, np.ndarray]:
    k = _c(keys, np.uint64).copy()
    v = _c(values, np.uint32).copy()
    load().orc_sort_pairs(k, v, k.size, bits)
    return k, v


def render(views: list[OrcView], cloud, with_f64: bool = False, relight=None, f64_cloud=None) -> dict:
    """The frame as the HIP path computes it, all views in one key space.
    ``f64_cloud``: the FP32 arrays the FP64 replay reads (a packed cloud's
    decoded fields); the cloud itself by default."""
    lib = load()
    n = cloud.positions.shape[0] if hasattr(cloud, "positions") else cloud.means.shape[0]
    nv = len(views)
    rec = np.zeros((nv * n, 12), np.float32)
    touched = np.zeros(nv * n, np.uint32)
    keys = np.zeros(nv * n, np.uint32)
    flags = np.zeros(nv * n, np.uint32)
    tile_bases = []
    total_tiles = 0
    for k, v in enumerate(views):
        r, t, kk, fl = project(v, cloud, flags=True, relight=relight)
        rec[k * n:(k + 1) * n] = r
        touched[k * n:(k + 1) * n] = t
        keys[k * n:(k + 1) * n] = kk
        flags[k * n:(k + 1) * n] = fl
        tile_bases.append(total_tiles)
        tx, ty = tiles_of(v)
        total_tiles += tx * ty
    # Each slot's first pair, its tiles counted in the reference's depth order
    # (TileRasterizer.cpp:1111-1137): visible slots stably sorted by the 24-bit
    # depth key, counts scanned in that order. The full key sort below does not
    # depend on this order; the HIP path's per-slot offsets do.
    visible = np.nonzero(touched)[0]
    order = visible[np.argsort(keys[visible] & np.uint32(0xFFFFFF), kind="stable")]
    sorted_offsets, pairs = exclusive_scan(touched[order])
    offsets = np.zeros(nv * n, np.uint32)
    offsets[order] = sorted_offsets
    pk = np.zeros(pairs, np.uint64)
    pv = np.zeros(pairs, np.uint32)
    for k, v in enumerate(views):
        sl = slice(k * n, (k + 1) * n)
        lib.orc_emit_pairs(ctypes.byref(v), _c(rec[sl], np.float32), _c(touched[sl], np.uint32),
                           _c(keys[sl], np.uint32), _c(offsets[sl], np.uint32), n, tile_bases[k], k * n, pk, pv)
    key_bits = 24 + bits_for(total_tiles)
    lib.orc_sort_pairs(pk, pv, pairs, key_bits)
    ranges = np.zeros(total_tiles * 2, np.uint32)
    lib.orc_ranges(pk, pairs, total_tiles, ranges)
    out = {"records": rec, "touched": touched, "depth_keys": keys, "flags": flags, "offsets": offsets,
           "pairs": pairs, "pair_keys": pk, "pair_values": pv, "ranges": ranges, "tile_bases": tile_bases,
           "key_bits": key_bits, "colour": [], "depth": [], "final_t": [], "last": []}
    for k, v in enumerate(views):
        H, W = v.height, v.width
        col = np.zeros(H * W * 4, np.float32)
        dep = np.zeros(H * W, np.float32)
        ft = np.zeros(H * W, np.float32)
        last = np.zeros(H * W, np.uint32)
        lib.orc_blend(ctypes.byref(v), rec, ranges, pv, tile_bases[k], col, dep, ft, last)
        out["colour"].append(col.reshape(H, W, 4))
        out["depth"].append(dep.reshape(H, W))
        out["final_t"].append(ft.reshape(H, W))
        out["last"].append(last.reshape(H, W))
    if with_f64:
        out.update(render_f64(views, f64_cloud if f64_cloud is not None else cloud, out))
    return out


def render_instances(views: list[list[OrcView]], clouds: list, relights: list | None = None) -> dict:
    """A frame of several instances (TileRasterizer::render over a span of
    SplatInstances): views[j][k] is instance j's view k -- the same screen
    with that instance's object-to-view rows and eye. Each instance is
    projected on its own, its records placed after the instances before it
    (slot = k * total + base_j + i, bases the running count,
    TileRasterizer.cpp:896-898); the depth order, pairs, ranges and blend are
    one frame's over all of them, as ``render`` takes them."""
    lib = load()
    counts = [c.positions.shape[0] if hasattr(c, "positions") else c.means.shape[0] for c in clouds]
    bases = [int(b) for b in np.concatenate([[0], np.cumsum(counts)[:-1]])] if counts else []
    n = int(sum(counts))
    nv = len(views[0])
    relights = relights or [None] * len(clouds)
    rec = np.zeros((nv * n, 12), np.float32)
    touched = np.zeros(nv * n, np.uint32)
    keys = np.zeros(nv * n, np.uint32)
    tile_bases = []
    total_tiles = 0
    for k in range(nv):
        for j, cloud in enumerate(clouds):
            r, t, kk = project(views[j][k], cloud, relight=relights[j])
            sl = slice(k * n + bases[j], k * n + bases[j] + counts[j])
            rec[sl] = r
            touched[sl] = t
            keys[sl] = kk
        tile_bases.append(total_tiles)
        tx, ty = tiles_of(views[0][k])
        total_tiles += tx * ty
    visible = np.nonzero(touched)[0]
    order = visible[np.argsort(keys[visible] & np.uint32(0xFFFFFF), kind="stable")]
    sorted_offsets, pairs = exclusive_scan(touched[order])
    offsets = np.zeros(nv * n, np.uint32)
    offsets[order] = sorted_offsets
    pk = np.zeros(pairs, np.uint64)
    pv = np.zeros(pairs, np.uint32)
    for k in range(nv):
        sl = slice(k * n, (k + 1) * n)
        lib.orc_emit_pairs(ctypes.byref(views[0][k]), _c(rec[sl], np.float32), _c(touched[sl], np.uint32),
                           _c(keys[sl], np.uint32), _c(offsets[sl], np.uint32), n, tile_bases[k], k * n, pk, pv)
    key_bits = 24 + bits_for(total_tiles)
    lib.orc_sort_pairs(pk, pv, pairs, key_bits)
    ranges = np.zeros(total_tiles * 2, np.uint32)
    lib.orc_ranges(pk, pairs, total_tiles, ranges)
    out = {"records": rec, "touched": touched, "depth_keys": keys, "offsets": offsets, "pairs": pairs,
           "pair_keys": pk, "pair_values": pv, "ranges": ranges, "tile_bases": tile_bases, "key_bits": key_bits,
           "slot_bases": bases, "colour": [], "depth": [], "final_t": [], "last": []}
    for k in range(nv):
        v = views[0][k]
        H, W = v.height, v.width
        col = np.zeros(H * W * 4, np.float32)
        dep = np.zeros(H * W, np.float32)
        ft = np.zeros(H * W, np.float32)
        last = np.zeros(H * W, np.uint32)
        lib.orc_blend(ctypes.byref(v), rec, ranges, pv, tile_bases[k], col, dep, ft, last)
        out["colour"].append(col.reshape(H, W, 4))
        out["depth"].append(dep.reshape(H, W))
        out["final_t"].append(ft.reshape(H, W))
        out["last"].append(last.reshape(H, W))
    return out


TRACE_WORDS = 11


def blend_ex(views: list[OrcView], frame: dict, under: list | None = None, slot_ids: np.ndarray | None = None,
             matte: bool = False, traced: bool = False) -> dict:
    """orc_blend_ex over a frame ``render`` (or ``render_instances``) made:
    the walk's variants -- ``under`` a list of (colour [H, W, 4], depth [H,
    W], crypto ids [H, W] or None) a view, ``slot_ids`` each slot's
    Cryptomatte id, and the matte and trace planes when asked for."""
    lib = load()
    out = {"colour": [], "depth": [], "final_t": [], "last": [], "crypto": [], "trace": []}
    ids = _c(slot_ids, np.uint32) if slot_ids is not None else None
    for k, v in enumerate(views):
        H, W = v.height, v.width
        col = np.zeros(H * W * 4, np.float32)
        dep = np.zeros(H * W, np.float32)
        ft = np.zeros(H * W, np.float32)
        last = np.zeros(H * W, np.uint32)
        cr = np.zeros(3 * H * W * 4, np.float32) if matte else None
        tr = np.full(TRACE_WORDS * H * W, 0xDEADBEEF, np.uint32) if traced else None
        uc = ud = ux = None
        if under is not None:
            uc = _c(under[k][0], np.float32)
            ud = _c(under[k][1], np.float32)
            ux = _c(under[k][2], np.uint32) if under[k][2] is not None else None
        lib.orc_blend_ex(ctypes.byref(v), _c(frame["records"], np.float32), _c(frame["ranges"], np.uint32),
                         _c(frame["pair_values"], np.uint32), frame["tile_bases"][k],
                         uc.ctypes.data if uc is not None else None, ud.ctypes.data if ud is not None else None,
                         ids.ctypes.data if ids is not None else None, ux.ctypes.data if ux is not None else None,
                         col, dep, ft, last, cr.ctypes.data if cr is not None else None,
                         tr.ctypes.data if tr is not None else None)
        out["colour"].append(col.reshape(H, W, 4))
        out["depth"].append(dep.reshape(H, W))
        out["final_t"].append(ft.reshape(H, W))
        out["last"].append(last.reshape(H, W))
        out["crypto"].append(cr.reshape(3, H, W, 4) if cr is not None else None)
        out["trace"].append(tr.reshape(TRACE_WORDS, H, W) if tr is not None else None)
    return out


def render_f64(views: list[OrcView], cloud, frame: dict) -> dict:
    """FP64 images of a frame, every FP32 decision replayed (see orc_blend_f64)."""
    lib = load()
    n = cloud.means.shape[0]
    coeffs = cloud.sh.shape[1]
    cols, deps = [], []
    for k, v in enumerate(views):
        sl = slice(k * n, (k + 1) * n)
        H, W = v.height, v.width
        col = np.zeros(H * W * 4, np.float64)
        dep = np.zeros(H * W, np.float64)
        # Slots of view k are k * n + i: hand the oracle this view's slice
        # with pair values rebased to it.
        pv = frame["pair_values"].astype(np.int64) - k * n
        tb = frame["tile_bases"][k]
        tx, ty = tiles_of(v)
        rg = frame["ranges"][2 * tb:2 * (tb + tx * ty)].copy()
        pv_view = np.where((pv >= 0) & (pv < n), pv, 0).astype(np.uint32)
        lib.orc_blend_f64(ctypes.byref(v), _c(frame["records"][sl], np.float32), _c(rg, np.uint32), pv_view,
                          _c(frame["touched"][sl], np.uint32), _c(frame["flags"][sl], np.uint32), 0,
                          _c(cloud.means, np.float64), _c(cloud.log_scales, np.float64),
                          _c(cloud.quats, np.float64), _c(cloud.opacity_logits, np.float64),
                          _c(cloud.sh, np.float64), coeffs, n, col, dep)
        cols.append(col.reshape(H, W, 4))
        deps.append(dep.reshape(H, W))
    return {"colour64": cols, "depth64": deps}


def render_ref_order(v: OrcView, cloud, pair_cap: int | None = None, relight=None) -> dict:
    """The reference's two-sort order on one view: returns pair values in
    sorted order, ranges, P and the stage times (ms)."""
    lib = load()
    rl, keep = orc_relight(relight) if relight is not None else (None, None)
    n = cloud.means.shape[0]
    coeffs = cloud.sh.shape[1]
    rec = np.zeros((n, 12), np.float32)
    touched = np.zeros(n, np.uint32)
    keys = np.zeros(n, np.uint32)
    tx, ty = tiles_of(v)
    ranges = np.zeros(tx * ty * 2, np.uint32)
    if pair_cap is None:
        r, t, _ = project(v, cloud, relight=relight)
        pair_cap = int(t.astype(np.uint64).sum())
    pv = np.zeros(max(pair_cap, 1), np.uint32)
    H, W = v.height, v.width
    col = np.zeros(H * W * 4, np.float32)
    dep = np.zeros(H * W, np.float32)
    ft = np.zeros(H * W, np.float32)
    last = np.zeros(H * W, np.uint32)
    ms = np.zeros(7, np.float64)
    P = lib.orc_render_ref_order(ctypes.byref(v), _c(cloud.means, np.float32), _c(cloud.log_scales, np.float32),
                                 _c(cloud.quats, np.float32), _c(cloud.opacity_logits, np.float32),
                                 _c(cloud.sh, np.float32), coeffs, n, rec, touched, keys, pv, pair_cap, ranges,
                                 col.ctypes.data, dep.ctypes.data, ft.ctypes.data, last.ctypes.data, ms.ctypes.data,
                                 ctypes.byref(rl) if rl is not None else None)
    del keep
    return {"pairs": int(P), "pair_values": pv[:P], "ranges": ranges, "colour": col.reshape(H, W, 4),
            "depth": dep.reshape(H, W), "final_t": ft.reshape(H, W), "last": last.reshape(H, W), "ms": ms,
            "records": rec, "touched": touched}


def reference_brute_force(v: OrcView, records: np.ndarray, touched: np.ndarray, max_hits: int = 1024):
    """reference_blend.slang's per-pixel brute force over every kept slot."""
    H, W = v.height, v.width
    col = np.zeros(H * W * 4, np.float32)
    dep = np.zeros(H * W, np.float32)
    over = load().orc_reference(ctypes.byref(v), _c(records, np.float32), _c(touched, np.uint32),
                                touched.size, max_hits, col, dep)
    return col.reshape(H, W, 4), dep.reshape(H, W), int(over)


def threads() -> int:
    return int(load().orc_threads())


def set_threads(n: int) -> None:
    load().orc_set_threads(n)


def expf(x: float) -> float:
    return float(load().orc_expf(x))


def logf(x: float) -> float:
    return float(load().orc_logf(x))


def fd_grads(v: OrcView, cloud, frame: dict, upstream: np.ndarray, ids, h_rel: float = 1e-4) -> np.ndarray:
    """FP64 central differences of L = sum(upstream * C(theta)) -- the image
    paired with a fixed dL/dcolour, the product an adjoint computes -- with
    every FP32 decision of ``frame`` (one view, from ``render``) frozen.
    Returns [len(ids), 11 + 3 * coeffs]: means (3), log-scales (3),
    quaternion (4), opacity logit (1), SH (coeffs * 3, coefficient-major);
    and, of the same shape, the sum over pixels of each pixel's share's
    magnitude |upstream . dC/dtheta| (what a float32 sum of the shares rounds
    against)."""
    lib = load()
    if not hasattr(lib, "_fd_ready"):
        lib.orc_fd_grads.argtypes = [ctypes.POINTER(OrcView), _f32p, _u32p, _u32p, ctypes.c_uint32, _u32p, _u32p,
                                     _f64p, _f64p, _f64p, _f64p, _f64p, ctypes.c_uint32, ctypes.c_uint32, _f64p,
                                     _u32p, ctypes.c_uint32, ctypes.c_double, _f64p, _f64p]
        lib.orc_fd_grads.restype = None
        lib._fd_ready = True
    n = cloud.means.shape[0]
    coeffs = cloud.sh.shape[1]
    ids = _c(ids, np.uint32)
    out = np.zeros((ids.size, 11 + 3 * coeffs), np.float64)
    mag = np.zeros_like(out)
    lib.orc_fd_grads(ctypes.byref(v), _c(frame["records"][:n], np.float32), _c(frame["ranges"], np.uint32),
                     _c(frame["pair_values"], np.uint32), int(frame["pairs"]), _c(frame["touched"][:n], np.uint32),
                     _c(frame["flags"][:n], np.uint32), _c(cloud.means, np.float64),
                     _c(cloud.log_scales, np.float64), _c(cloud.quats, np.float64),
                     _c(cloud.opacity_logits, np.float64), _c(cloud.sh, np.float64), coeffs, n,
                     _c(upstream, np.float64), ids, ids.size, h_rel, out, mag)
    return out, mag


def _f3(x):
    return (ctypes.c_float * 3)(*[float(t) for t in x])


def relit_point(relight, n, wo, albedo, metallic: float, roughness: float) -> np.ndarray:
    """The relit colour of one surface point: n the world shading normal, wo
    the unit direction to the eye."""
    rl, keep = orc_relight(relight)
    out = _f3((0, 0, 0))
    load().orc_relit_point(ctypes.byref(rl), _f3(n), _f3(wo), _f3(albedo), metallic, roughness, out)
    return np.array(out[:], np.float32)


def env_radiance(relight, direction, roughness: float) -> np.ndarray:
    rl, keep = orc_relight(relight)
    out = _f3((0, 0, 0))
    load().orc_env_radiance(ctypes.byref(rl), _f3(direction), roughness, out)
    return np.array(out[:], np.float32)


def env_irradiance(relight, normal) -> np.ndarray:
    rl, keep = orc_relight(relight)
    out = _f3((0, 0, 0))
    load().orc_env_irradiance(ctypes.byref(rl), _f3(normal), out)
    return np.array(out[:], np.float32)


def sky_sun(equirect: np.ndarray) -> np.ndarray:
    """orc_sky_sun (env_sun.slang:154-261): the sun words [2, 4] of an
    equirect [H, W, 3] of linear radiance."""
    img = np.ascontiguousarray(equirect, np.float32)
    out = np.zeros(8, np.float32)
    if load().orc_sky_sun(_c(img, np.float32), img.shape[1], img.shape[0], out) != 0:
        raise MemoryError("orc_sky_sun")
    return out.reshape(2, 4)


def sky_residual(equirect: np.ndarray, sun: np.ndarray) -> np.ndarray:
    """orc_sky_residual (env_residual.slang:55-78): the equirect with the sun's
    texels set to 0."""
    img = np.ascontiguousarray(equirect, np.float32)
    out = np.zeros_like(img)
    words = np.ascontiguousarray(sun, np.float32).reshape(8)
    if load().orc_sky_residual(_c(img, np.float32), img.shape[1], img.shape[0], words, out) != 0:
        raise MemoryError("orc_sky_residual")
    return out


def env_level_offset(level: int, base_side: int) -> int:
    return int(load().orc_env_level_offset(level, base_side))


def half_to_float(h: int) -> float:
    return float(load().orc_half_to_float(h))


def ggx_albedo(nv: float, alpha: float) -> tuple[float, float]:
    a, b = ctypes.c_float(), ctypes.c_float()
    load().orc_ggx_albedo(nv, alpha, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def ggx_dielectric(roughness: float, nv: float) -> float:
    return float(load().orc_ggx_dielectric(roughness, nv))


def ggx_conductor(albedo: float, roughness: float, nv: float) -> float:
    return float(load().orc_ggx_conductor(albedo, roughness, nv))


def fresnel_dielectric(c: float, ior: float) -> float:
    return float(load().orc_fresnel_dielectric(c, ior))


def fresnel_conductor(albedo: float, nv: float) -> float:
    return float(load().orc_fresnel_conductor(albedo, nv))


def unpack_normal(word: int) -> np.ndarray:
    out = (ctypes.c_float * 3)()
    load().orc_unpack_normal(word, out)
    return np.array(out[:], np.float32)


def decode_quaternion(word: int) -> np.ndarray:
    out = (ctypes.c_float * 4)()
    load().orc_decode_quaternion(word, out)
    return np.array(out[:], np.float32)


def ingest(vertices: np.ndarray, count: int, stride: int, fields: list[tuple[int, int]], keep_faint: bool = False,
           modes: tuple = (0, 0, 0, 0)):
    """orc_ingest: `count` vertices of `stride` bytes (a uint8 array); fields =
    (byte offset, athx_ply_type) of x y z, opacity, scale_0..2, rot_0..3 (w x y
    z), f_dc_0..2, then the kept rest coefficients rgb a basis. Returns the
    kept vertices' positions [k, 4] f32, shape [k, 4] u32, sh [k, words] u32
    and origin [k] u32."""
    vertices = np.ascontiguousarray(vertices, np.uint8)
    assert vertices.size >= count * stride and len(fields) >= 14
    rest = len(fields) - 14
    words = (rest + 1) // 2
    offsets = np.array([f[0] for f in fields], np.uint32)
    types = np.array([f[1] for f in fields], np.uint32)
    positions = np.zeros((max(count, 1), 4), np.float32)
    shape = np.zeros((max(count, 1), 4), np.uint32)
    sh = np.zeros((max(count, 1), max(words, 1)), np.uint32)
    origin = np.zeros(max(count, 1), np.uint32)
    kept = load().orc_ingest_modes(vertices.ctypes.data, count, stride, _c(offsets, np.uint32),
                                   _c(types, np.uint32), rest, 1 if keep_faint else 0,
                                   _c(np.asarray(modes, np.uint32), np.uint32), _c(positions, np.float32),
                                   _c(shape, np.uint32), sh.ctypes.data, _c(origin, np.uint32))
    return positions[:kept], shape[:kept], sh[:kept, :words], origin[:kept]


def half_bits(v: float) -> int:
    return int(load().orc_half_bits(v))


def encode_quaternion(x: float, y: float, z: float, w: float) -> int:
    return int(load().orc_encode_quaternion(x, y, z, w))


def ref_project(v: OrcView, packed) -> np.ndarray:
    """orc_ref_project: the ground truth's own projection of a packed cloud
    (host arrays), records [n, 12] (z = -1: culled)."""
    n = packed.positions.shape[0]
    rec = np.zeros((max(n, 1), 12), np.float32)
    sh = np.ascontiguousarray(packed.sh, np.uint32) if packed.sh is not None else None
    load().orc_ref_project(ctypes.byref(v), _c(packed.positions, np.float32), _c(packed.shape, np.uint32),
                           sh.ctypes.data if sh is not None else None, n, packed.rest_per_colour, _c(rec, np.float32))
    return rec[:n]


def compare_codes(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    h, w = a.shape[:2]
    bins = np.zeros(256, np.uint32)
    load().orc_compare_codes(_c(a, np.float32), _c(b, np.float32), w, h, _c(bins, np.uint32))
    return bins


def compare_hdr(a: np.ndarray, b: np.ndarray) -> tuple[np.ndarray, float]:
    h, w = a.shape[:2]
    bins = np.zeros(256, np.uint32)
    total = load().orc_compare_hdr(_c(a, np.float32), _c(b, np.float32), w, h, _c(bins, np.uint32))
    return bins, float(np.float32(total))


def contribution(v: OrcView, frame: dict, slots: int, mode: int = 1, dl: np.ndarray | None = None) -> np.ndarray:
    """orc_contribution over one view's frame (render's dict): uint32
    [slots] (mode 1), [slots, 2] (modes 2 and 3) or [slots, 4] (mode 4)."""
    words = {1: 1, 2: 2, 3: 2, 4: 4}[mode]
    out = np.zeros(slots * words, np.uint32)
    d = np.ascontiguousarray(dl, np.float32) if dl is not None else None
    load().orc_contribution(ctypes.byref(v), _c(frame["records"], np.float32), _c(frame["ranges"], np.uint32),
                            _c(frame["pair_values"], np.uint32), 0, mode, d.ctypes.data if d is not None else None,
                            out)
    return out.reshape(slots, words) if words > 1 else out
