"""Cameras and projections, in double precision, as the reference's host has them.

Restates modules/render/src/Camera.cpp:9-64 and modules/render/include/
athenea/render/Camera.h:23-89: a camera looks down its own -Z; the view the
kernels see is that flipped to +z forward (``scale(1, 1, -1)`` times the
inverse of camera-to-world); the focal length in pixels is
``width * focal / haperture``; the principal point is the image centre. The
floats a frame is drawn with are these doubles cast, as setFrame/setObject do
(modules/render/src/FrameParams.h:32-95).

The reference takes lookAt from aopenfx (``aofx/Transform.h``), whose source
is not in its checkout; this module's ``look_at`` is the usual right-handed
one (camera x = forward x up, camera y = x x forward, camera z = -forward),
which is what the reference's tests assume (tests/render/
test_splat_render.cpp:159-206: from +Z, -x is on the left).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib


@dataclass
class Lens:
    orthographic: bool = False
    focal: float = 50.0  # mm
    haperture: float = 24.576  # mm
    near_z: float = 0.1
    far_z: float = 10000.0
    window_translate: tuple[float, float] = (0.0, 0.0)
    window_scale: tuple[float, float] = (1.0, 1.0)
    window_roll: float = 0.0  # degrees, counter-clockwise (Camera.h:33)


@dataclass
class Camera:
    camera_to_world: np.ndarray = field(default_factory=lambda: np.eye(4))
    lens: Lens = field(default_factory=Lens)

    @staticmethod
    def looking_at(eye, target, up=(0.0, 1.0, 0.0)) -> "Camera":
        """Camera.cpp:9-24, with the near plane a thousandth of the distance."""
        cam = Camera(camera_to_world=look_at(eye, target, up))
        away = np.asarray(eye, np.float64) - np.asarray(target, np.float64)
        cam.lens.near_z = max(float(np.linalg.norm(away)) * 1e-3, 1e-5)
        return cam

    def with_horizontal_fov(self, degrees: float) -> "Camera":
        """Sets focal / haperture so the image is ``degrees`` wide."""
        self.lens.focal = self.lens.haperture / (2.0 * math.tan(math.radians(degrees) * 0.5))
        return self


def look_at(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, np.asarray(up, np.float64))
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[:3, 0] = s
    m[:3, 1] = u
    m[:3, 2] = -f
    m[:3, 3] = eye
    return m


@dataclass
class Projection:
    """render::Projection (Camera.h:50-89), the fields a plain frame reads."""

    world_to_view: np.ndarray
    focal_x: float
    focal_y: float
    centre_x: float
    centre_y: float
    near_z: float
    far_z: float
    orthographic: bool
    eye_world: np.ndarray


def rotation_z(degrees: float) -> np.ndarray:
    """A counter-clockwise turn about z by ``degrees``, 4x4."""
    t = math.radians(degrees)
    c, s = math.cos(t), math.sin(t)
    m = np.eye(4)
    m[0, 0], m[0, 1] = c, -s
    m[1, 0], m[1, 1] = s, c
    return m


def projection_for(camera: Camera, width: int, height: int) -> Projection:
    """Camera.cpp:32-64; the window roll turns the picture about the view's
    own z after the flip (Camera.cpp:35-38)."""
    flip = np.diag([1.0, 1.0, -1.0, 1.0])
    world_to_view = flip @ np.linalg.inv(camera.camera_to_world)
    if abs(camera.lens.window_roll) > 1e-12:
        world_to_view = rotation_z(camera.lens.window_roll) @ world_to_view
    w, h = float(width), float(height)
    focal = w * max(camera.lens.focal, 1e-9) / max(camera.lens.haperture, 1e-9)
    return Projection(
        world_to_view=world_to_view,
        focal_x=focal * camera.lens.window_scale[0],
        focal_y=focal * camera.lens.window_scale[1],
        centre_x=w * 0.5 + camera.lens.window_translate[0] * w * 0.5,
        centre_y=h * 0.5 + camera.lens.window_translate[1] * w * 0.5,
        near_z=camera.lens.near_z,
        far_z=camera.lens.far_z,
        orthographic=camera.lens.orthographic,
        eye_world=camera.camera_to_world[:3, 3].copy(),
    )


def make_view(
    projection: Projection,
    width: int,
    height: int,
    object_to_world: np.ndarray | None = None,
    background=(0.0, 0.0, 0.0, 0.0),
) -> _lib.View:
    """The athx_view of one instance under one projection (FrameParams.h:32-95)."""
    o2w = np.eye(4) if object_to_world is None else np.asarray(object_to_world, np.float64)
    o2v = projection.world_to_view @ o2w
    eye_obj = np.linalg.inv(o2w) @ np.append(projection.eye_world, 1.0)
    v = _lib.View()
    v.width = int(width)
    v.height = int(height)
    v.focal_x = projection.focal_x
    v.focal_y = projection.focal_y
    v.centre_x = projection.centre_x
    v.centre_y = projection.centre_y
    v.near_z = projection.near_z
    v.far_z = projection.far_z
    v.orthographic = 1 if projection.orthographic else 0
    rows = o2v[:3, :4].reshape(-1)
    for k in range(12):
        v.object_to_view[k] = float(rows[k])
    for k in range(3):
        v.eye[k] = float(eye_obj[k])
    for k in range(4):
        v.background[k] = float(background[k])
    # The doubles the view was made from: a multi-instance frame folds each
    # instance's object-to-world into them (instance_transform).
    v.projection = projection
    return v


def instance_transform(view, object_to_world=None) -> tuple[np.ndarray, np.ndarray]:
    """(object-to-view rows [12], eye [3]) as float32 for one instance under a
    view made by make_view / view_for: Projection::worldToView times the
    instance's objectToWorld and the eye taken into the instance's space, in
    double, then cast (TileRasterizer.cpp:598-600)."""
    projection = getattr(view, "projection", None)
    if projection is None:
        raise ValueError("a multi-instance frame needs views made by camera.make_view / view_for")
    o2w = np.eye(4) if object_to_world is None else np.asarray(object_to_world, np.float64)
    if o2w.shape == (3, 4):
        o2w = np.vstack([o2w, [0.0, 0.0, 0.0, 1.0]])
    o2v = projection.world_to_view @ o2w
    eye_obj = np.linalg.inv(o2w) @ np.append(projection.eye_world, 1.0)
    return o2v[:3, :4].reshape(-1).astype(np.float32), eye_obj[:3].astype(np.float32)


def view_for(camera: Camera, width: int, height: int, object_to_world=None, background=(0.0, 0.0, 0.0, 0.0)):
    return make_view(projection_for(camera, width, height), width, height, object_to_world, background)


def orbit_camera(radius: float, azimuth: float, fov_degrees: float, height: float = 0.0) -> Camera:
    """A camera on a circle about the origin, looking at it (the benchmark walks)."""
    eye = (radius * math.sin(azimuth), height, radius * math.cos(azimuth))
    return Camera.looking_at(eye, (0.0, 0.0, 0.0)).with_horizontal_fov(fov_degrees)
