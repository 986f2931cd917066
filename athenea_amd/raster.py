"""The host-side mirror of athenea's splat-render API over libathx.so.

``TileRasterizer`` keeps the names and meaning of render::TileRasterizer
(modules/render/include/athenea/render/TileRasterizer.h:452-487): ``create``,
``render``, ``latest_counters``, ``release_scratch``; ``RenderSettings``,
``RenderTargets``, ``FrameStats`` and ``SplatCounters`` keep the fields a plain
frame reads (:248-450). Errors are raised as ``AthxError`` carrying the
reference's ErrorCode name. Device memory comes from PyTorch (ROCm); all the
work is the HIP library's, on the caller's current stream.

``exclusive_scan`` and ``sort_pairs`` are gpu::PrefixSum::apply and
gpu::RadixSort::sort (modules/gpu/include/athenea/gpu/algo/PrefixSum.h:43-44,
RadixSort.h:63-64).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib


@dataclass
class RenderSettings:
    tile: int = 16
    batch: int = 256
    antialias: bool = True
    max_sh_degree: int = 3
    depth: str = "mean"  # "mean" or "threshold"
    depth_threshold: float = 0.5
    time_stages: bool = False
    count_splats: bool = False
    counters_tag: int = 0  # RenderSettings::countersTag, returned with the frame's counters
    cryptomatte: bool = False  # RenderSettings::cryptomatte: allocated targets get the matte planes
    trace: bool = False  # RenderSettings::trace: allocated targets get the trace planes

    def to_c(self, linear: bool) -> _lib.Settings:
        s = _lib.Settings()
        s.tile = self.tile
        s.batch = self.batch
        s.antialias = 1 if self.antialias else 0
        s.linear = 1 if linear else 0
        s.max_sh_degree = self.max_sh_degree
        s.depth_mode = 0 if self.depth == "mean" else 1
        s.depth_threshold = self.depth_threshold
        s.time_stages = 1 if self.time_stages else 0
        s.count_splats = 1 if self.count_splats else 0
        s.counters_tag = int(self.counters_tag)
        return s


@dataclass
class Relight:
    """athx_relight: relitSplat's dome term with the plain lobes in place of
    the SH colour (SplatInstance::relight with one prepared dome). Tensors
    are float32 device tensors; ``sky`` is a DeviceSky (athenea_amd.sky).
    ``albedo`` None: the evaluated SH colour is the albedo; ``normals`` None:
    the disc's short axis is the shading normal."""

    metallic: torch.Tensor  # [n]
    roughness: torch.Tensor  # [n]
    sky: object
    normals: torch.Tensor | None = None  # [n, 3]
    albedo: torch.Tensor | None = None  # [n, 3]
    object_to_world: tuple = (1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0)
    normals_packed: torch.Tensor | None = None  # [n] int32 octahedral words, read instead of normals
    pbr: torch.Tensor | None = None  # [n] int32 pbr words, read instead of metallic and roughness


@dataclass
class GaussianCloud:
    """Device tensors in athx_gaussians' layout. ``linear``: colours are linear
    light (GpuSplats::linear); otherwise sRGB, decoded per splat. ``relight``:
    the relit colour instead of the harmonics' (forward only)."""

    means: torch.Tensor
    log_scales: torch.Tensor
    quats: torch.Tensor
    opacity_logits: torch.Tensor
    sh: torch.Tensor  # [n, k, 3], float32 or float16
    linear: bool = True
    relight: Relight | None = None
    crypto_ids: torch.Tensor | None = None  # [n] int32 Cryptomatte ids (GpuSplats::crypto), or None

    @property
    def count(self) -> int:
        return int(self.means.shape[0])

    @staticmethod
    def from_numpy(cloud, device="cuda", linear: bool = True, relight=None) -> "GaussianCloud":
        """``relight``: an athenea_amd.sky.RelightAttributes, or None."""

        def t(a):
            return torch.from_numpy(a).to(device).contiguous()

        return GaussianCloud(t(cloud.means), t(cloud.log_scales), t(cloud.quats), t(cloud.opacity_logits),
                             t(cloud.sh), linear, relight.to_device(device) if relight is not None else None)


@dataclass
class PackedGaussians:
    """athx_packed_splats: a cloud in athenea's packed device layout
    (scene::GpuSplats) -- positions float4 [n, 4] (xyz, linear opacity), shape
    words [n, 4] (smallest-three quaternion, f16 log-scales and base colour),
    the rest coefficients as halves [n, (rest_per_colour * 3 + 1) // 2] (int32
    tensors holding the uint32 words). Forward only."""

    positions: torch.Tensor
    shape: torch.Tensor
    sh: torch.Tensor | None
    rest_per_colour: int
    linear: bool = True
    relight: Relight | None = None
    crypto_ids: torch.Tensor | None = None  # [n] int32 Cryptomatte ids, or None

    @property
    def count(self) -> int:
        return int(self.positions.shape[0])


@dataclass
class SplatInstance:
    """SplatInstance (TileRasterizer.h:85-172): a cloud and where it stands,
    objectToWorld as a 4x4 (or 3x4) in double; None is the identity. A relit
    cloud's lighting is taken to the world by the same transform."""

    cloud: object  # GaussianCloud or PackedGaussians
    object_to_world: object = None


@dataclass
class RenderTargets:
    """One view's images, bottom row first (TileRasterizer.h:421-450)."""

    width: int
    height: int
    colour: torch.Tensor  # [h, w, 4] premultiplied linear
    depth: torch.Tensor  # [h, w]
    final_t: torch.Tensor | None = None  # [h, w]
    last: torch.Tensor | None = None  # [h, w] int32
    # The number of the frame that last wrote these targets on its rasterizer
    # (athx_frame_serial): the backward pass and the contributions walk that
    # frame and are refused once a later render has replaced it.
    frame_serial: int = 0
    # RenderTargets::crypto: three planes of (id bits, coverage, id bits,
    # coverage) a pixel, descending; RenderTargets::trace: TRACE_WORDS planes
    # of words a pixel (splat_trace.slang). None: not kept.
    crypto: torch.Tensor | None = None  # [3, h, w, 4] float32
    trace: torch.Tensor | None = None  # [TRACE_WORDS, h, w] int32

    @staticmethod
    def allocate(width: int, height: int, device, keep_backward: bool, crypto: bool = False,
                 trace: bool = False) -> "RenderTargets":
        colour = torch.empty((height, width, 4), dtype=torch.float32, device=device)
        depth = torch.empty((height, width), dtype=torch.float32, device=device)
        ft = torch.empty((height, width), dtype=torch.float32, device=device) if keep_backward else None
        last = torch.empty((height, width), dtype=torch.int32, device=device) if keep_backward else None
        cr = torch.empty((3, height, width, 4), dtype=torch.float32, device=device) if crypto else None
        tr = torch.empty((_lib.TRACE_WORDS, height, width), dtype=torch.int32, device=device) if trace else None
        return RenderTargets(width, height, colour, depth, ft, last, 0, cr, tr)


@dataclass
class UnderLayer:
    """What is under the splats (the reference's `under` RenderTargets): an
    opaque layer's linear premultiplied colour, its view depth (0 where it has
    nothing) and, for the matte, its Cryptomatte ids."""

    colour: torch.Tensor  # [h, w, 4] float32
    depth: torch.Tensor  # [h, w] float32
    crypto_ids: torch.Tensor | None = None  # [h, w] int32


@dataclass
class FrameStats:
    splats: int = 0
    visible: int = 0
    pairs: int = 0
    project_ms: float = 0.0
    depth_sort_ms: float = 0.0
    counts_ms: float = 0.0
    emit_ms: float = 0.0
    tile_sort_ms: float = 0.0
    ranges_ms: float = 0.0
    blend_ms: float = 0.0
    total_ms: float = 0.0
    frame_serial: int = 0
    # Where each instance's Gaussians start among a view's slots
    # (FrameStats::slotBases); [0] for a one-cloud frame.
    slot_bases: list = field(default_factory=lambda: [0])


CULL_NAMES = ["unprojected", "edit", "depth", "degenerate", "faint", "offscreen", "no_tile"]


@dataclass
class CloudCounters:
    """SplatCounters::clouds[j]: one instance's drawn footprints and pairs."""

    visible: int = 0
    pairs: int = 0


@dataclass
class SplatCounters:
    """SplatCounters (TileRasterizer.h:370-401)."""

    slots: int = 0
    visible: int = 0
    pairs: int = 0
    max_tiles: int = 0
    culled: dict = field(default_factory=dict)
    tag: int = 0
    clouds: list = field(default_factory=list)


def _ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensors handed to the rasterizer must be contiguous")
    return t.data_ptr()


def _device_of(cloud) -> torch.device:
    return cloud.positions.device if isinstance(cloud, PackedGaussians) else cloud.means.device


def _stream_handle(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _check_cloud(cloud: GaussianCloud) -> None:
    n = cloud.count
    shapes = {
        "means": (cloud.means, (n, 3)),
        "log_scales": (cloud.log_scales, (n, 3)),
        "quats": (cloud.quats, (n, 4)),
        "opacity_logits": (cloud.opacity_logits, (n,)),
    }
    for name, (t, shape) in shapes.items():
        if tuple(t.shape) != shape or t.dtype != torch.float32 or not t.is_cuda:
            raise ValueError(f"{name} must be a float32 device tensor of shape {shape}, got {tuple(t.shape)} {t.dtype}")
    if cloud.sh.dim() != 3 or cloud.sh.shape[0] != n or cloud.sh.shape[2] != 3 or cloud.sh.shape[1] not in (1, 4, 9, 16):
        raise ValueError(f"sh must be [n, 1|4|9|16, 3], got {tuple(cloud.sh.shape)}")
    if cloud.sh.dtype not in (torch.float32, torch.float16):
        raise ValueError("sh must be float32 or float16")
    r = cloud.relight
    if r is not None:
        per = {"metallic": (r.metallic, (n,)), "roughness": (r.roughness, (n,)), "normals": (r.normals, (n, 3)),
               "albedo": (r.albedo, (n, 3))}
        for name, (t, shape) in per.items():
            if t is None and (name in ("normals", "albedo") or r.pbr is not None):
                continue
            if t is None or tuple(t.shape) != shape or t.dtype != torch.float32 or not t.is_cuda:
                raise ValueError(f"relight {name} must be a float32 device tensor of shape {shape}")
        if len(r.object_to_world) != 12:
            raise ValueError("relight object_to_world is 12 floats (rows of a 3x4)")
        sky = r.sky
        if tuple(sky.sh.shape) != (16, 4) or tuple(sky.sun.shape) != (2, 4):
            raise ValueError("sky sh is [16, 4] and sun [2, 4]")


def _check_packed(cloud: PackedGaussians) -> None:
    n = cloud.count
    if tuple(cloud.positions.shape) != (n, 4) or cloud.positions.dtype != torch.float32:
        raise ValueError("packed positions must be float32 [n, 4]")
    if tuple(cloud.shape.shape) != (n, 4) or cloud.shape.dtype not in (torch.int32, torch.uint32):
        raise ValueError("packed shape must be 32-bit words [n, 4]")
    if cloud.rest_per_colour not in (0, 3, 8, 15):
        raise ValueError("rest_per_colour must be 0, 3, 8 or 15")
    words = (cloud.rest_per_colour * 3 + 1) // 2
    if cloud.rest_per_colour and (cloud.sh is None or tuple(cloud.sh.shape) != (n, words)):
        raise ValueError(f"packed sh must be 32-bit words [n, {words}]")


def _gaussians_c(cloud, object_to_world=None) -> _lib.Gaussians:
    """athx_gaussians of a cloud; ``object_to_world`` (rows of a 3x4) replaces
    a relit cloud's own."""
    g = _lib.Gaussians()
    g.count = cloud.count
    if isinstance(cloud, PackedGaussians):
        _check_packed(cloud)
        p = _lib.PackedSplats()
        p.positions = _ptr(cloud.positions)
        p.shape = _ptr(cloud.shape)
        p.sh = _ptr(cloud.sh)
        p.rest_per_colour = cloud.rest_per_colour
        g.sh_coeffs = cloud.rest_per_colour + 1
        g.packed = ctypes.pointer(p)
        g._packed = p
    else:
        _check_cloud(cloud)
        g.sh_coeffs = int(cloud.sh.shape[1])
        g.means = _ptr(cloud.means)
        g.log_scales = _ptr(cloud.log_scales)
        g.quats = _ptr(cloud.quats)
        g.opacity_logits = _ptr(cloud.opacity_logits)
        if cloud.sh.dtype == torch.float16:
            g.sh = None
            g.sh_half = _ptr(cloud.sh)
        else:
            g.sh = _ptr(cloud.sh)
            g.sh_half = None
    ids = getattr(cloud, "crypto_ids", None)
    if ids is not None:
        if tuple(ids.shape) != (cloud.count,) or ids.dtype != torch.int32 or not ids.is_cuda:
            raise ValueError("crypto_ids must be an int32 device tensor [count]")
        g.crypto_ids = _ptr(ids)
    r = cloud.relight
    if r is not None:
        c = _lib.Relight()
        c.normals = _ptr(r.normals)
        c.albedo = _ptr(r.albedo)
        c.metallic = _ptr(r.metallic)
        c.roughness = _ptr(r.roughness)
        c.normals_packed = _ptr(r.normals_packed)
        c.pbr = _ptr(r.pbr)
        rows = r.object_to_world if object_to_world is None else object_to_world
        for k in range(12):
            c.object_to_world[k] = float(rows[k])
        c.sky.texels = _ptr(r.sky.texels)
        c.sky.sh = _ptr(r.sky.sh)
        c.sky.sun = _ptr(r.sky.sun)
        c.sky.base_side = int(r.sky.base_side)
        g.relight = ctypes.pointer(c)
        g._relight = c  # the pointee lives as long as the struct that points to it
    return g


def _views_and_targets(views: list, targets: list[RenderTargets], under: list[UnderLayer] | None = None):
    if len(targets) != len(views):
        raise ValueError("one target per view")
    if under is not None and len(under) != len(views):
        raise ValueError("one under layer per view")
    nv = len(views)
    c_views = (_lib.View * nv)(*views)
    c_targets = (_lib.Targets * nv)()
    for k, t in enumerate(targets):
        h, w = views[k].height, views[k].width
        if t.colour.shape != (h, w, 4):
            raise ValueError(f"target {k} does not match its view's size")
        c_targets[k].colour = _ptr(t.colour)
        c_targets[k].depth = _ptr(t.depth)
        c_targets[k].final_t = _ptr(t.final_t)
        c_targets[k].last = _ptr(t.last)
        if t.crypto is not None:
            if tuple(t.crypto.shape) != (3, h, w, 4) or t.crypto.dtype != torch.float32:
                raise ValueError(f"target {k}: crypto must be float32 [3, {h}, {w}, 4]")
            c_targets[k].crypto = _ptr(t.crypto)
        if t.trace is not None:
            if tuple(t.trace.shape) != (_lib.TRACE_WORDS, h, w) or t.trace.dtype != torch.int32:
                raise ValueError(f"target {k}: trace must be int32 [{_lib.TRACE_WORDS}, {h}, {w}]")
            c_targets[k].trace = _ptr(t.trace)
        if under is not None:
            u = under[k]
            if tuple(u.colour.shape) != (h, w, 4) or u.colour.dtype != torch.float32 or \
                    tuple(u.depth.shape) != (h, w) or u.depth.dtype != torch.float32:
                raise ValueError(f"under layer {k}: colour float32 [{h}, {w}, 4] and depth float32 [{h}, {w}]")
            if u.crypto_ids is not None and (tuple(u.crypto_ids.shape) != (h, w) or u.crypto_ids.dtype != torch.int32):
                raise ValueError(f"under layer {k}: crypto_ids int32 [{h}, {w}]")
            c_targets[k].under_colour = _ptr(u.colour)
            c_targets[k].under_depth = _ptr(u.depth)
            c_targets[k].under_crypto = _ptr(u.crypto_ids)
    return c_views, c_targets


def _serial_of(targets: list[RenderTargets]) -> int:
    """The one frame every target was written by."""
    serials = {t.frame_serial for t in targets}
    if len(serials) != 1:
        raise ValueError("the targets were written by different frames")
    return serials.pop()


def _grads_c(grads: dict[str, torch.Tensor]) -> _lib.Grads:
    cg = _lib.Grads()
    for name in ("means", "log_scales", "quats", "opacity_logits", "sh"):
        t = grads.get(name)
        if t is not None and (t.dtype != torch.float32 or not t.is_contiguous()):
            raise ValueError(f"gradient {name} must be a contiguous float32 tensor")
        setattr(cg, name, _ptr(t))
    return cg


class FrameGraph:
    """A frame captured once into a hipGraph (athx_graph_create) and replayed:
    the batched, wait-free route. It owns a context of its own sized for these
    views and this pair capacity (0: sized from one eager frame, 1.5x its
    pairs); every replay reads whatever the Gaussian, target and gradient
    tensors hold then. With ``dl_dcolour`` and ``grads`` the adjoint is
    captured too (``launch_backward``; the caller zeroes ``grads``)."""

    def __init__(self, views: list, cloud: GaussianCloud, settings: RenderSettings | None = None,
                 targets: list[RenderTargets] | None = None, pair_capacity: int = 0,
                 dl_dcolour: list[torch.Tensor] | None = None, grads: dict[str, torch.Tensor] | None = None,
                 device: int = 0):
        self._lib = _lib.load()
        self.settings = settings or RenderSettings()
        dev = _device_of(cloud)
        backward = dl_dcolour is not None and grads is not None
        self.targets = targets or [RenderTargets.allocate(v.width, v.height, dev, backward) for v in views]
        self.cloud = cloud
        self.dl_dcolour = [d.contiguous() for d in dl_dcolour] if backward else None
        self.grads = grads
        c_views, c_targets = _views_and_targets(views, self.targets)
        g = _gaussians_c(cloud)
        s = self.settings.to_c(cloud.linear)
        dl = None
        cg = None
        if backward:
            dl = (ctypes.c_void_p * len(views))(*[d.data_ptr() for d in self.dl_dcolour])
            cg = ctypes.byref(_grads_c(grads))
        handle = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        status = self._lib.athx_graph_create(device, c_views, len(views), ctypes.byref(g), ctypes.byref(s),
                                             c_targets, int(pair_capacity), dl, cg, ctypes.byref(handle), err, 512)
        if status != 0:
            raise _lib.AthxError(status, err.value.decode())
        self._graph = handle

    def launch(self) -> None:
        _lib.check(self._lib.athx_graph_launch(self._graph, _stream_handle(_device_of(self.cloud))))

    def launch_backward(self) -> None:
        _lib.check(self._lib.athx_graph_launch_backward(self._graph, _stream_handle(_device_of(self.cloud))))

    def pairs(self) -> tuple[int, int]:
        """(pair total, capacity) of the last replay; waits for it."""
        pairs, cap = ctypes.c_uint64(), ctypes.c_uint64()
        _lib.check(self._lib.athx_graph_pairs(self._graph, ctypes.byref(pairs), ctypes.byref(cap)))
        return int(pairs.value), int(cap.value)

    def close(self) -> None:
        if getattr(self, "_graph", None):
            self._lib.athx_graph_destroy(self._graph)
            self._graph = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TileRasterizer:
    """One context of libathx (one device, one host thread at a time)."""

    def __init__(self, device: int = 0):
        self._lib = _lib.load()
        self.device = device
        handle = ctypes.c_void_p()
        _lib.check(self._lib.athx_create(device, ctypes.byref(handle)))
        self._ctx = handle

    @staticmethod
    def create(device: int = 0) -> "TileRasterizer":
        return TileRasterizer(device)

    def close(self) -> None:
        if getattr(self, "_ctx", None):
            self._lib.athx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, status: int) -> None:
        _lib.check(status, self._ctx)

    def render(
        self,
        views: list,
        cloud: GaussianCloud,
        settings: RenderSettings | None = None,
        targets: list[RenderTargets] | None = None,
        keep_backward: bool = False,
        under: list[UnderLayer] | None = None,
    ) -> tuple[list[RenderTargets], FrameStats]:
        """Draws ``cloud`` once per view (a list of athx View structs); all
        views go through one projection, one sort and one blend launch.
        ``under``: an opaque layer under the splats, one a view."""
        settings = settings or RenderSettings()
        dev = _device_of(cloud)
        if targets is None:
            targets = [RenderTargets.allocate(v.width, v.height, dev, keep_backward, settings.cryptomatte,
                                              settings.trace) for v in views]
        nv = len(views)
        c_views, c_targets = _views_and_targets(views, targets, under)
        g = _gaussians_c(cloud)
        s = settings.to_c(cloud.linear)
        stats = _lib.FrameStats()
        self._check(
            self._lib.athx_render(self._ctx, c_views, nv, ctypes.byref(g), ctypes.byref(s), c_targets,
                                  _stream_handle(dev), ctypes.byref(stats))
        )
        out = FrameStats(stats.splats, stats.visible, stats.pairs, stats.project_ms, stats.depth_sort_ms,
                         stats.counts_ms, stats.emit_ms, stats.tile_sort_ms, stats.ranges_ms, stats.blend_ms,
                         stats.total_ms, int(stats.frame_serial))
        for t in targets:
            t.frame_serial = int(stats.frame_serial)
        return targets, out

    def render_instances(
        self,
        views: list,
        instances: list[SplatInstance],
        settings: RenderSettings | None = None,
        targets: list[RenderTargets] | None = None,
        keep_backward: bool = False,
        under: list[UnderLayer] | None = None,
    ) -> tuple[list[RenderTargets], FrameStats]:
        """TileRasterizer::render over a span of SplatInstances
        (athx_render_instances): each cloud projected with its own
        object-to-world, all of them sorted and blended together. ``views``
        come from camera.make_view / view_for (their own object-to-world is
        replaced by each instance's). A view's slots are the instances'
        Gaussians back to back; ``stats.slot_bases`` gives where each starts."""
        from . import camera as _camera

        settings = settings or RenderSettings()
        if not instances:
            raise ValueError("no instances")
        dev = _device_of(instances[0].cloud)
        if any(_device_of(i.cloud) != dev for i in instances):
            raise ValueError("every instance's cloud must be on one device")
        linear = {bool(i.cloud.linear) for i in instances}
        if len(linear) != 1:
            raise ValueError("the instances' clouds must all be linear or all sRGB")
        if targets is None:
            targets = [RenderTargets.allocate(v.width, v.height, dev, keep_backward, settings.cryptomatte,
                                              settings.trace) for v in views]
        c_views, c_targets = _views_and_targets(views, targets, under)
        nv, ni = len(views), len(instances)
        keep = []
        c_inst = (_lib.Instance * ni)()
        for j, inst in enumerate(instances):
            rows, eyes = [], []
            for v in views:
                r, e = _camera.instance_transform(v, inst.object_to_world)
                rows.append(r)
                eyes.append(e)
            rows = np.ascontiguousarray(np.stack(rows), np.float32)
            eyes = np.ascontiguousarray(np.stack(eyes), np.float32)
            o2w = None
            if inst.object_to_world is not None:
                o2w = np.asarray(inst.object_to_world, np.float64)[:3, :4].reshape(-1)
            g = _gaussians_c(inst.cloud, o2w)
            keep += [rows, eyes, g]
            c_inst[j].gaussians = ctypes.pointer(g)
            c_inst[j].object_to_view = rows.ctypes.data
            c_inst[j].eye = eyes.ctypes.data
        s = settings.to_c(linear.pop())
        stats = _lib.FrameStats()
        bases = (ctypes.c_uint32 * ni)()
        self._check(self._lib.athx_render_instances(self._ctx, c_views, nv, c_inst, ni, ctypes.byref(s), c_targets,
                                                    _stream_handle(dev), ctypes.byref(stats), bases))
        del keep
        out = FrameStats(stats.splats, stats.visible, stats.pairs, stats.project_ms, stats.depth_sort_ms,
                         stats.counts_ms, stats.emit_ms, stats.tile_sort_ms, stats.ranges_ms, stats.blend_ms,
                         stats.total_ms, int(stats.frame_serial), [int(b) for b in bases])
        for t in targets:
            t.frame_serial = int(stats.frame_serial)
        return targets, out

    def render_async(
        self,
        views: list,
        cloud: GaussianCloud,
        settings: RenderSettings | None = None,
        targets: list[RenderTargets] | None = None,
        pair_capacity: int = 0,
        keep_backward: bool = False,
    ) -> list[RenderTargets]:
        """athx_render_async: the frame with no host wait, its pair buffers
        sized for ``pair_capacity`` (0: 1.5x the largest frame seen). Check
        ``frame_pairs()`` for a frame that outgrew its capacity."""
        settings = settings or RenderSettings()
        dev = _device_of(cloud)
        if targets is None:
            targets = [RenderTargets.allocate(v.width, v.height, dev, keep_backward) for v in views]
        c_views, c_targets = _views_and_targets(views, targets)
        g = _gaussians_c(cloud)
        s = settings.to_c(cloud.linear)
        self._check(self._lib.athx_render_async(self._ctx, c_views, len(views), ctypes.byref(g), ctypes.byref(s),
                                                c_targets, int(pair_capacity), _stream_handle(dev)))
        serial = self.frame_serial()
        for t in targets:
            t.frame_serial = serial
        return targets

    def frame_serial(self) -> int:
        """athx_frame_serial: the number of the last frame rendered here (0: none)."""
        return int(self._lib.athx_frame_serial(self._ctx))

    def frame_pairs(self) -> tuple[int, int]:
        """(pair total, pair capacity) of the last frame; waits for it."""
        pairs, cap = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self._lib.athx_frame_pairs(self._ctx, ctypes.byref(pairs), ctypes.byref(cap)))
        return int(pairs.value), int(cap.value)

    def contribution(self, views: list, cloud, settings: RenderSettings | None, targets: list[RenderTargets],
                     mode: int = 1, dl_dcolour: list[torch.Tensor] | None = None) -> torch.Tensor:
        """athx_contribution for the frame the last ``render`` drew (slot =
        view * count + gaussian; ``cloud`` may be the list of SplatInstances a
        ``render_instances`` frame drew, whose count is theirs): mode 1 an int32 [slots] of blend weights in
        1/4096; mode 2 an int32 [slots, 2], the (low, high) words of the
        squared removal error in 2^-24 against the drawn image given as
        ``dl_dcolour``; mode 3 an int32 [slots, 2], the weight times the mask
        given as ``dl_dcolour`` ([height, width]) and the weight, in 1/1024;
        mode 4 an int32 [slots, 4] of the colour gradient against
        ``dl_dcolour`` and the weight, in 1/256."""
        settings = settings or RenderSettings()
        instances = cloud if isinstance(cloud, (list, tuple)) else None
        if instances is not None:
            # A multi-instance frame: its slots are every instance's.
            dev = _device_of(instances[0].cloud)
            count = sum(i.cloud.count for i in instances)
            linear = bool(instances[0].cloud.linear)
        else:
            dev = _device_of(cloud)
            count = cloud.count
            linear = bool(cloud.linear)
        c_views, c_targets = _views_and_targets(views, targets)
        nv = len(views)
        slots = count * nv
        words = {1: 1, 2: 2, 3: 2, 4: 4}.get(int(mode))
        if words is None:
            raise ValueError("contribution modes are 1 to 4")
        dl = None
        keep = []
        if mode != 1:
            if dl_dcolour is None or len(dl_dcolour) != nv:
                raise ValueError(f"contribution mode {mode} needs one per-pixel input per view")
            dl = (ctypes.c_void_p * nv)()
            for k, t in enumerate(dl_dcolour):
                want = (views[k].height, views[k].width) + ((4,) if mode in (2, 4) else ())
                if not isinstance(t, torch.Tensor) or tuple(t.shape) != want:
                    got = list(t.shape) if isinstance(t, torch.Tensor) else type(t).__name__
                    raise ValueError(f"view {k}: mode {mode} reads a {list(want)} tensor, got {got}")
                if not t.is_floating_point():
                    raise ValueError(f"view {k}: the per-pixel input must be floating point")
                if t.device != dev:
                    raise ValueError(f"view {k}: the per-pixel input must be on {dev}, not {t.device}")
                t = t.to(torch.float32).contiguous()
                keep.append(t)
                dl[k] = t.data_ptr()
        out = torch.empty((slots, words) if words > 1 else (slots,), dtype=torch.int32, device=dev)
        g = ctypes.byref(_gaussians_c(cloud)) if instances is None else None
        s = settings.to_c(linear)
        self._check(self._lib.athx_contribution(self._ctx, c_views, nv, g, ctypes.byref(s), c_targets,
                                                _serial_of(targets), int(mode), dl, out.data_ptr(),
                                                _stream_handle(dev)))
        return out

    def render_backward(
        self,
        views: list,
        cloud: GaussianCloud,
        settings: RenderSettings | None,
        targets: list[RenderTargets],
        dl_dcolour: list[torch.Tensor],
        grads: dict[str, torch.Tensor],
        overwrite: bool = False,
    ) -> None:
        """Accumulates dL/d(parameter) into ``grads`` (keys: means, log_scales,
        quats, opacity_logits, sh) for the frame the last ``render`` drew; with
        ``overwrite`` writes them instead (every element, so ``grads`` may be
        uninitialised)."""
        settings = settings or RenderSettings()
        nv = len(views)
        c_views, c_targets = _views_and_targets(views, targets)
        keep = []
        dl = (ctypes.c_void_p * nv)()
        for k, t in enumerate(targets):
            d = dl_dcolour[k].to(torch.float32).contiguous()
            if d.shape != t.colour.shape:
                raise ValueError("dL/dcolour must match the colour target")
            keep.append(d)
            dl[k] = d.data_ptr()
        cg = _lib.Grads()
        for name in ("means", "log_scales", "quats", "opacity_logits", "sh"):
            t = grads.get(name)
            if t is not None and (t.dtype != torch.float32 or not t.is_contiguous()):
                raise ValueError(f"gradient {name} must be a contiguous float32 tensor")
            setattr(cg, name, _ptr(t))
        cg.overwrite = 1 if overwrite else 0
        g = _gaussians_c(cloud)
        s = settings.to_c(cloud.linear)
        self._check(
            self._lib.athx_render_backward(self._ctx, c_views, nv, ctypes.byref(g), ctypes.byref(s), c_targets,
                                           _serial_of(targets), dl, ctypes.byref(cg),
                                           _stream_handle(_device_of(cloud)))
        )
        # ``keep`` may hold temporaries the queued kernels read; PyTorch's
        # allocator is stream-ordered, so releasing them here is safe.
        del keep

    def latest_counters(self) -> SplatCounters | None:
        """TileRasterizer::latestCounters: the newest counted frame's counters
        that have reached the host, without waiting; None before any has."""
        c = _lib.Counters()
        status = self._lib.athx_latest_counters(self._ctx, ctypes.byref(c))
        if status == 2:  # NotFound: nothing counted has finished
            return None
        self._check(status)
        return SplatCounters(c.slots, c.visible, c.pairs, c.max_tiles,
                             {CULL_NAMES[r]: int(c.culled[r]) for r in range(_lib.CULL_REASONS)}, int(c.tag),
                             [CloudCounters(int(c.cloud_visible[j]), int(c.cloud_pairs[j])) for j in range(c.clouds)])

    def frame_info(self) -> dict:
        info = _lib.FrameInfo()
        self._check(self._lib.athx_frame_info_get(self._ctx, ctypes.byref(info)))
        return {"slots": info.slots, "tiles": info.tiles, "key_bits": info.key_bits, "pairs": info.pairs,
                "strips": info.strips, "entries": info.entries}

    def frame_buffers(self, device=None) -> dict[str, torch.Tensor]:
        """Copies of the last frame's per-slot and per-pair buffers (athx_copy_frame):
        records [slots, 12] f32, touched / depth_keys / offsets [slots] (uint32 bits
        in int32), pair_keys [pairs] (uint64 bits in int64), pair_values [pairs],
        ranges [tiles, 2]."""
        info = self.frame_info()
        device = device or torch.device("cuda", self.device)
        shapes = {
            "records": ((info["slots"], 12), torch.float32),
            "touched": ((info["slots"],), torch.int32),
            "depth_keys": ((info["slots"],), torch.int32),
            "offsets": ((info["slots"],), torch.int32),
            "pair_keys": ((info["pairs"],), torch.int64),
            "pair_values": ((info["pairs"],), torch.int32),
            "ranges": ((info["tiles"], 2), torch.int32),
        }
        out = {}
        for name, (shape, dtype) in shapes.items():
            t = torch.empty(shape, dtype=dtype, device=device)
            nbytes = t.numel() * t.element_size()
            self._check(self._lib.athx_copy_frame(self._ctx, _lib.FRAME_PARTS[name], t.data_ptr() if nbytes else None,
                                                  nbytes, _stream_handle(device)))
            out[name] = t
        return out

    def release_scratch(self) -> None:
        self._check(self._lib.athx_release_scratch(self._ctx))

    def scratch_bytes(self) -> int:
        return int(self._lib.athx_scratch_bytes(self._ctx))

    def exclusive_scan(self, values: torch.Tensor, out: torch.Tensor | None = None) -> tuple[torch.Tensor, torch.Tensor]:
        """PrefixSum::apply: out[i] = sum(values[:i]) over uint32 (int32 storage), and the total (int64)."""
        if values.dtype != torch.int32 or not values.is_cuda:
            raise ValueError("values must be an int32 device tensor holding uint32 bits")
        out = torch.empty_like(values) if out is None else out
        total = torch.zeros(1, dtype=torch.int64, device=values.device)
        self._check(self._lib.athx_exclusive_scan(self._ctx, _ptr(values), _ptr(out), _ptr(total), values.numel(),
                                                  _stream_handle(values.device)))
        return out, total

    def exclusive_scan_noalloc(self, values: torch.Tensor, out: torch.Tensor, total: torch.Tensor) -> None:
        """athx_exclusive_scan_noalloc: exclusive_scan into caller-given ``out``
        (int32, values' length) and ``total`` (one int64), with nothing
        allocated -- the library's static scratch; at most 2^24 values."""
        for name, t, dtype in (("values", values, torch.int32), ("out", out, torch.int32), ("total", total, torch.int64)):
            if t.dtype != dtype or not t.is_cuda or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous {dtype} device tensor")
        if out.numel() != values.numel() or total.numel() < 1:
            raise ValueError("out must match values, total hold one int64")
        self._check(self._lib.athx_exclusive_scan_noalloc(self._ctx, _ptr(values), _ptr(out), _ptr(total),
                                                          values.numel(), _stream_handle(values.device)))

    def sort_pairs(self, keys: torch.Tensor, values: torch.Tensor, key_bits: int) -> None:
        """RadixSort::sort in place: a stable sort by the low key_bits bits of
        int64 keys (uint64 bits), values (int32) carried along."""
        if keys.dtype != torch.int64 or values.dtype != torch.int32 or keys.numel() != values.numel():
            raise ValueError("keys must be int64 and values int32, of one length")
        self._check(self._lib.athx_sort_pairs(self._ctx, _ptr(keys), _ptr(values), keys.numel(), key_bits,
                                              _stream_handle(keys.device)))

