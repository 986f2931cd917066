"""A sky prepared for relighting, in the layout athx_sky reads
(shaders/athenea/light/environment.slang:50-110): one dome, eight
octahedral levels of the sky convolved with GGX lobes, nine (of sixteen)
radiance coefficients in real spherical harmonics, and an optional sun.

The reference prepares this on the GPU (technique::Environment,
modules/technique/include/athenea/technique/Environment.h, with
env_project / env_prefilter / env_sun .slang); so does
prepare_sky_device (athx_prepare_sky: FP64 sums over the equirect's
texels, one rounding to f16 a texel). prepare_sky is the same quadrature in
numpy on the host: what the device preparation is checked against, and the
way to make a sky without a device. The relit colour only reads the
buffers, so the GPU frame and the oracle are fed the same ones.

Conventions: +y is up; an equirect texel (i, j) of a W x H map looks along
theta = pi (j + 0.5) / H from +y and phi = 2 pi (i + 0.5) / W from +x
towards +z: d = (sin theta cos phi, cos theta, sin theta sin phi).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

LEVELS = 8
COARSE_SIDE = 256
FLOOR_SIDE = 16
COEFFICIENTS = 16


def level_side(level: int, base_side: int) -> int:
    """environment.slang:66-69"""
    base = max(base_side, COARSE_SIDE)
    return max(base >> min(level, LEVELS - 1), FLOOR_SIDE)


def level_offset(level: int, base_side: int) -> int:
    """Texels before a level (environment.slang:77-92)."""
    return sum(level_side(k, base_side) ** 2 for k in range(min(level, LEVELS)))


def texels_per_dome(base_side: int) -> int:
    return level_offset(LEVELS, base_side)


def level_roughness(level: int) -> float:
    """The roughness a level is convolved to: (k / 7)^2, the square-root
    walk of environment.slang:149-170."""
    return (level / (LEVELS - 1)) ** 2


def sh_basis(d: np.ndarray) -> np.ndarray:
    """The 16 real harmonics of environment.slang:202-224 at unit directions
    d [..., 3], float64."""
    x, y, z = d[..., 0], d[..., 1], d[..., 2]
    return np.stack([
        np.full_like(x, 0.28209479177387814),
        -0.4886025119029199 * y,
        0.4886025119029199 * z,
        -0.4886025119029199 * x,
        1.0925484305920792 * x * y,
        -1.0925484305920792 * y * z,
        0.31539156525252005 * (2 * z * z - x * x - y * y),
        -1.0925484305920792 * x * z,
        0.5462742152960396 * (x * x - y * y),
        -0.5900435899266435 * y * (3 * x * x - y * y),
        2.890611442640554 * x * y * z,
        -0.4570457994644658 * y * (4 * z * z - x * x - y * y),
        0.3731763325901154 * z * (2 * z * z - 3 * x * x - 3 * y * y),
        -0.4570457994644658 * x * (4 * z * z - x * x - y * y),
        1.445305721320277 * z * (x * x - y * y),
        -0.5900435899266435 * x * (x * x - 3 * y * y),
    ], axis=-1)


def oct_decode(u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """The square [0, 1]^2 back to unit directions (common/octahedral.slang:25-34)."""
    fx = u * 2.0 - 1.0
    fy = v * 2.0 - 1.0
    n = np.stack([fx, fy, 1.0 - np.abs(fx) - np.abs(fy)], axis=-1)
    t = np.clip(-n[..., 2], 0.0, 1.0)
    n[..., 0] += np.where(n[..., 0] >= 0.0, -t, t)
    n[..., 1] += np.where(n[..., 1] >= 0.0, -t, t)
    return n / np.linalg.norm(n, axis=-1, keepdims=True)


def equirect_directions(width: int, height: int) -> tuple[np.ndarray, np.ndarray]:
    """Texel-centre directions [H, W, 3] and solid angles [H, W] of an equirect."""
    theta = np.pi * (np.arange(height) + 0.5) / height
    phi = 2.0 * np.pi * (np.arange(width) + 0.5) / width
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    d = np.stack([st * np.cos(phi)[None, :], np.broadcast_to(ct, (height, width)), st * np.sin(phi)[None, :]],
                 axis=-1)
    omega = st * (np.pi / height) * (2.0 * np.pi / width)
    return d, np.broadcast_to(omega, (height, width)).copy()


def sample_equirect(image: np.ndarray, d: np.ndarray) -> np.ndarray:
    """Bilinear lookup of an equirect [H, W, 3] along unit directions [..., 3],
    wrapping in azimuth and clamping at the poles."""
    h, w = image.shape[:2]
    theta = np.arccos(np.clip(d[..., 1], -1.0, 1.0))
    phi = np.mod(np.arctan2(d[..., 2], d[..., 0]), 2.0 * np.pi)
    fy = np.clip(theta / np.pi * h - 0.5, 0.0, h - 1.0)
    fx = phi / (2.0 * np.pi) * w - 0.5
    y0 = np.floor(fy).astype(np.int64)
    y1 = np.minimum(y0 + 1, h - 1)
    ty = (fy - y0)[..., None]
    x0f = np.floor(fx)
    tx = (fx - x0f)[..., None]
    x0 = np.mod(x0f.astype(np.int64), w)
    x1 = np.mod(x0 + 1, w)
    top = image[y0, x0] * (1 - tx) + image[y0, x1] * tx
    bottom = image[y1, x0] * (1 - tx) + image[y1, x1] * tx
    return top * (1 - ty) + bottom * ty


def _ggx_prefilter(dirs: np.ndarray, src_d: np.ndarray, src_l: np.ndarray, src_w: np.ndarray,
                   roughness: float) -> np.ndarray:
    """The sky convolved with a GGX lobe of this roughness about each
    direction, taking n = v = r (the usual split-sum assumption):
    sum L(l) D(n.h) (n.l) dw / sum D(n.h) (n.l) dw over n.l > 0."""
    alpha = max(roughness * roughness, 1e-3)  # env_prefilter.slang:139
    a2 = alpha * alpha
    out = np.zeros((dirs.shape[0], 3))
    for s in range(0, dirs.shape[0], 2048):
        n = dirs[s:s + 2048]
        nl = n @ src_d.T
        h = n[:, None, :] + src_d[None, :, :]
        h /= np.maximum(np.linalg.norm(h, axis=-1, keepdims=True), 1e-12)
        nh = np.einsum("ik,ijk->ij", n, h)
        dd = a2 / (np.pi * (nh * nh * (a2 - 1.0) + 1.0) ** 2)
        wgt = np.where(nl > 0.0, dd * nl, 0.0) * src_w[None, :]
        out[s:s + 2048] = (wgt @ src_l) / np.maximum(wgt.sum(axis=1, keepdims=True), 1e-30)
    return out


def _pack_half(rgb: np.ndarray) -> np.ndarray:
    """rgb [m, 3] -> uint32 [m, 2]: red | green << 16, blue (IEEE binary16,
    round to nearest even)."""
    h = np.ascontiguousarray(rgb, dtype=np.float32).astype(np.float16).view(np.uint16).astype(np.uint32)
    words = np.empty((rgb.shape[0], 2), np.uint32)
    words[:, 0] = h[:, 0] | (h[:, 1] << 16)
    words[:, 1] = h[:, 2]
    return words


@dataclass
class PreparedSky:
    """One prepared dome in athx_sky's layout, host memory."""

    texels: np.ndarray  # uint32 [texels_per_dome(base_side) * 2]
    sh: np.ndarray  # float32 [16, 4]
    sun: np.ndarray  # float32 [2, 4]
    base_side: int

    def to_device(self, device) -> "DeviceSky":
        import torch

        return DeviceSky(torch.from_numpy(self.texels.view(np.int32)).to(device),
                         torch.from_numpy(self.sh).to(device), torch.from_numpy(self.sun).to(device),
                         self.base_side)


@dataclass
class DeviceSky:
    texels: "object"  # torch int32 words
    sh: "object"  # torch float32 [16, 4]
    sun: "object"  # torch float32 [2, 4]
    base_side: int


def prepare_sky(equirect: np.ndarray, base_side: int = COARSE_SIDE, sun: tuple | None = None) -> PreparedSky:
    """Prepares an HDR equirect [H, W, 3] (linear radiance): the octahedral
    chain (level 0 the sky itself, level k convolved to roughness (k/7)^2),
    its 16 radiance coefficients, and the sun: None, or
    ((dx, dy, dz), disc_size, (r, g, b) irradiance)."""
    image = np.asarray(equirect, np.float64)
    src_d, omega = equirect_directions(image.shape[1], image.shape[0])
    src_d = src_d.reshape(-1, 3)
    src_l = image.reshape(-1, 3)
    src_w = omega.reshape(-1)
    coeffs = (sh_basis(src_d) * src_w[:, None]).T @ src_l  # [16, 3]
    sh = np.zeros((COEFFICIENTS, 4), np.float32)
    sh[:, :3] = coeffs
    words = []
    for level in range(LEVELS):
        side = level_side(level, base_side)
        c = (np.arange(side) + 0.5) / side
        u, v = np.meshgrid(c, c)  # row y, column x
        dirs = oct_decode(u, v).reshape(-1, 3)
        if level == 0:
            rgb = sample_equirect(image, dirs)
        else:
            rgb = _ggx_prefilter(dirs, src_d, src_l, src_w, level_roughness(level))
        words.append(_pack_half(rgb))
    texels = np.concatenate(words).reshape(-1)
    assert texels.size == 2 * texels_per_dome(base_side)
    sun_words = np.zeros((2, 4), np.float32)
    if sun is not None:
        direction, size, irradiance = sun
        dvec = np.asarray(direction, np.float64)
        sun_words[0, :3] = dvec / np.linalg.norm(dvec)
        sun_words[0, 3] = size
        sun_words[1, :3] = irradiance
    return PreparedSky(texels.astype(np.uint32), sh, sun_words, int(base_side))


def constant_sky(radiance=(1.0, 1.0, 1.0), base_side: int = COARSE_SIDE) -> PreparedSky:
    """A sky of one radiance everywhere, prepared exactly: every texel is L
    and the only coefficient is the constant one, L * sqrt(4 pi)."""
    rgb = np.asarray(radiance, np.float64)
    count = texels_per_dome(base_side)
    texels = np.repeat(_pack_half(rgb[None, :]), count, axis=0).reshape(-1)
    sh = np.zeros((COEFFICIENTS, 4), np.float32)
    sh[0, :3] = rgb * np.sqrt(4.0 * np.pi)
    return PreparedSky(texels.astype(np.uint32), sh, np.zeros((2, 4), np.float32), int(base_side))


def lognormal_equirect(width: int = 64, height: int = 32, sigma: float = 1.5, seed: int = 0x5EED0001) -> np.ndarray:
    """An HDR sky of independent log-normal radiance per texel and channel,
    exp(N(0, sigma))."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.normal(0.0, sigma, (height, width, 3))).astype(np.float32)


IDENTITY_3X4 = (1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0)


@dataclass
class RelightAttributes:
    """Per-Gaussian relight inputs and the sky, host memory (the layout
    athx_relight points into; what the oracle reads too)."""

    metallic: np.ndarray  # float32 [n]
    roughness: np.ndarray  # float32 [n]
    sky: PreparedSky
    normals: np.ndarray | None = None  # float32 [n, 3]
    albedo: np.ndarray | None = None  # float32 [n, 3]
    object_to_world: tuple = field(default=IDENTITY_3X4)

    def to_device(self, device):
        import torch

        from .raster import Relight

        def t(a):
            return None if a is None else torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(device)

        return Relight(t(self.metallic), t(self.roughness), self.sky.to_device(device), t(self.normals),
                       t(self.albedo), tuple(float(x) for x in self.object_to_world))


def relight_attributes(n: int, sky: PreparedSky, seed: int = 0x5EED0001, with_albedo: bool = True,
                       with_normals: bool = True) -> RelightAttributes:
    """Relight attributes of the 1080p workload: albedo U[0,1]^3, roughness
    U[0.05, 1], metallic Bernoulli(0.2), unit normal normalize(N(0,1)^3)."""
    rng = np.random.default_rng(seed)
    albedo = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    roughness = rng.uniform(0.05, 1.0, n).astype(np.float32)
    metallic = (rng.uniform(0.0, 1.0, n) < 0.2).astype(np.float32)
    normals = rng.normal(0.0, 1.0, (n, 3))
    normals = (normals / np.linalg.norm(normals, axis=1, keepdims=True)).astype(np.float32)
    return RelightAttributes(metallic, roughness, sky, normals if with_normals else None,
                             albedo if with_albedo else None)


def _device_equirect(equirect, device):
    import torch

    img = torch.as_tensor(np.asarray(equirect, np.float32) if not torch.is_tensor(equirect) else equirect,
                          dtype=torch.float32).to(device).contiguous()
    if img.dim() != 3 or img.shape[2] != 3:
        raise ValueError("the equirect must be [height, width, 3]")
    return img


def sun_of(equirect, raster, device="cuda:0") -> np.ndarray:
    """The sun found in an equirect [H, W, 3] (athx_sky_sun; Environment::sun,
    env_sun.slang): float32 [2, 4] -- (direction, cone half angle, 0 for no
    sun), (irradiance rgb, the luminance it was cut at)."""
    from . import _lib

    img = _device_equirect(equirect, device)
    out = np.zeros(8, np.float32)
    raster._check(_lib.load().athx_sky_sun(raster._ctx, img.data_ptr(), int(img.shape[1]), int(img.shape[0]),
                                           out.ctypes.data))
    return out.reshape(2, 4)


def residual_of(equirect, sun: np.ndarray, raster, device="cuda:0"):
    """The equirect without its sun's texels (athx_sky_residual; env_residual
    .slang), a device tensor [H, W, 3]."""
    import torch

    from . import _lib

    img = _device_equirect(equirect, device)
    words = np.ascontiguousarray(sun, np.float32).reshape(8)
    out = torch.empty_like(img)
    raster._check(_lib.load().athx_sky_residual(raster._ctx, img.data_ptr(), int(img.shape[1]), int(img.shape[0]),
                                                words.ctypes.data, out.data_ptr()))
    return out


def prepare_sky_device(equirect, base_side: int = COARSE_SIDE, sun: tuple | str | None = None, raster=None,
                       device="cuda:0") -> DeviceSky:
    """prepare_sky on the device (athx_prepare_sky): an equirect [H, W, 3]
    (numpy or a device tensor) into a DeviceSky; `sun` as prepare_sky's, or
    "extract": the sun found in the sky (athx_sky_sun) and the levels and
    harmonics prepared from the sky without it (athx_sky_residual), as
    Environment::build does."""
    import ctypes

    import torch

    from . import _lib
    from .raster import TileRasterizer

    img = _device_equirect(equirect, device)
    lib = _lib.load()
    count = int(lib.athx_sky_texels(int(base_side)))
    texels = torch.empty((count * 2,), dtype=torch.int32, device=img.device)
    sh = torch.zeros((COEFFICIENTS, 4), dtype=torch.float32, device=img.device)
    own = raster is None
    raster = raster or TileRasterizer(img.device.index or 0)
    extracted = None
    try:
        if isinstance(sun, str):
            if sun != "extract":
                raise ValueError('sun must be None, (direction, size, irradiance) or "extract"')
            extracted = sun_of(img, raster, device=img.device)
            img = residual_of(img, extracted, raster, device=img.device)
        raster._check(lib.athx_prepare_sky(raster._ctx, img.data_ptr(), int(img.shape[1]), int(img.shape[0]),
                                           int(base_side), texels.data_ptr(), sh.data_ptr()))
    finally:
        if own:
            raster.close()
    sun_words = np.zeros((2, 4), np.float32)
    if extracted is not None:
        sun_words = extracted.copy()
    elif sun is not None:
        direction, size, irradiance = sun
        dvec = np.asarray(direction, np.float64)
        sun_words[0, :3] = dvec / np.linalg.norm(dvec)
        sun_words[0, 3] = size
        sun_words[1, :3] = irradiance
    return DeviceSky(texels, sh, torch.from_numpy(sun_words).to(img.device), int(base_side))
