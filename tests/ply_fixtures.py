"""PLY files for the ingest tests, written byte by byte so every value is
known: the reference's own loading fixtures (tests/scene/test_loading.cpp)
restated, and seeded random clouds of every property type."""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"

# athx_ply_type codes and the PLY type names / numpy dtypes they stand for.
TYPES = {"float": (0, "<f4"), "double": (1, "<f8"), "uchar": (2, "u1"), "char": (3, "i1"), "ushort": (4, "<u2"),
         "short": (5, "<i2"), "uint": (6, "<u4"), "int": (7, "<i4")}


def header(props: list[tuple[str, str]], count: int, before: str = "", fmt: str = "binary_little_endian") -> bytes:
    text = f"ply\nformat {fmt} 1.0\n{before}element vertex {count}\n"
    text += "".join(f"property {t} {n}\n" for n, t in props)
    return (text + "end_header\n").encode()


def block(props: list[tuple[str, str]], columns: dict[str, np.ndarray], count: int) -> bytes:
    dtype = np.dtype([(n, TYPES[t][1]) for n, t in props])
    rec = np.zeros(count, dtype)
    for n, _ in props:
        if n in columns:
            rec[n] = columns[n]
    return rec.tobytes()


def layout_fields(props: list[tuple[str, str]], rest_kept: int, per_colour: int) -> tuple[list[tuple[int, int]], int]:
    """(byte offset, type) of the fields the ingest reads, in its order, and the stride."""
    offsets, at = {}, 0
    for n, t in props:
        offsets[n] = (at, TYPES[t][0])
        at += np.dtype(TYPES[t][1]).itemsize
    names = ["x", "y", "z", "opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3",
             "f_dc_0", "f_dc_1", "f_dc_2"]
    fields = [offsets[n] for n in names]
    for basis in range(rest_kept):
        for colour in range(3):
            fields.append(offsets[f"f_rest_{colour * per_colour + basis}"])
    return fields, at


# The reference's "a 3DGS PLY decodes opacity, scale, rotation, colour and
# harmonics exactly" file (test_loading.cpp:38-62, 93-128): degree 1, four
# records, two of them dropped.
LOADING_PROPS = ([("x", "float"), ("y", "float"), ("z", "float"), ("nx", "float"), ("ny", "float"),
                  ("nz", "float"), ("f_dc_0", "float"), ("f_dc_1", "float"), ("f_dc_2", "float")]
                 + [(f"f_rest_{k}", "float") for k in range(9)]
                 + [("opacity", "float"), ("scale_0", "float"), ("scale_1", "float"), ("scale_2", "float"),
                    ("rot_0", "float"), ("rot_1", "float"), ("rot_2", "float"), ("rot_3", "float")])


def loading_file() -> bytes:
    nan = float("nan")
    first = dict(x=1.0, y=2.0, z=3.0, opacity=0.0, scale_0=0.0, scale_1=0.6931472, scale_2=-0.6931472, rot_0=2.0)
    for k in range(9):
        first[f"f_rest_{k}"] = np.float32(0.1) * np.float32(k + 1)
    faint = dict(first, opacity=-10.0)
    broken = dict(first, x=nan)
    last = dict(first, x=-4.0, y=5.0, z=-6.0, opacity=1.0986123, rot_0=0.0, rot_3=-3.0, f_dc_0=1.0 / 0.28209479)
    rows = [first, faint, broken, last]
    cols = {n: np.array([r.get(n, 0.0) for r in rows], np.float32) for n, _ in LOADING_PROPS}
    return header(LOADING_PROPS, 4) + block(LOADING_PROPS, cols, 4)


# The reference's mixed-types file (test_loading.cpp:207-235): doubles, a
# short, an unused uchar, so every later offset is unaligned.
MIXED_PROPS = [("x", "double"), ("y", "double"), ("z", "double"), ("opacity", "short"), ("unused", "uchar"),
               ("scale_0", "float"), ("scale_1", "float"), ("scale_2", "float"), ("rot_0", "float"),
               ("rot_1", "float"), ("rot_2", "float"), ("rot_3", "float"), ("f_dc_0", "float"), ("f_dc_1", "float"),
               ("f_dc_2", "float")]


def mixed_file() -> bytes:
    out = header(MIXED_PROPS, 2)
    for k in range(2):
        out += struct.pack("<ddd", 1.5 + k, -2.25, 1.0e-3 * (k + 1))
        out += struct.pack("<hB", 0 if k == 0 else 2, 7)
        out += struct.pack("<fff", 0.0, 0.6931472, -0.6931472)
        out += struct.pack("<ffff", 1.0, 0.0, 0.0, 0.0)
        out += struct.pack("<fff", 0.0, 1.0, -1.0)
    return out


def random_cloud(count: int, seed: int, degree: int = 3, types: dict[str, str] | None = None,
                 extra: bool = True) -> tuple[list[tuple[str, str]], bytes, dict[str, np.ndarray]]:
    """A 3DGS-like PLY of `count` vertices with harmonics of `degree`; `types`
    overrides a property's type. Includes faint, NaN, flat and out-of-range
    vertices, and quaternions of no length, among ordinary ones."""
    rng = np.random.default_rng(seed)
    per = (0, 3, 8, 15)[degree]
    names = ["x", "y", "z"] + (["nx", "ny", "nz"] if extra else []) + ["f_dc_0", "f_dc_1", "f_dc_2"]
    names += [f"f_rest_{k}" for k in range(3 * per)] + ["opacity", "scale_0", "scale_1", "scale_2",
                                                       "rot_0", "rot_1", "rot_2", "rot_3"]
    types = types or {}
    props = [(n, types.get(n, "float")) for n in names]
    cols: dict[str, np.ndarray] = {}
    cols["x"] = rng.normal(0, 3, count)
    cols["y"] = rng.normal(0, 3, count)
    cols["z"] = rng.normal(0, 3, count)
    for k in range(3):
        cols[f"f_dc_{k}"] = rng.normal(0, 1.5, count)
        cols[f"scale_{k}"] = rng.uniform(-7, 1, count)
    for k in range(3 * per):
        cols[f"f_rest_{k}"] = rng.normal(0, 0.3, count)
    cols["opacity"] = rng.normal(0, 4, count)
    for k in range(4):
        cols[f"rot_{k}"] = rng.normal(0, 1, count)
    pick = rng.integers(0, 12, count)
    cols["x"][pick == 0] = np.nan
    cols["y"][pick == 1] = np.inf
    cols["scale_1"][pick == 2] = -40.0          # a flat axis: kept
    cols["scale_0"][pick == 3] = -40.0
    cols["scale_2"][pick == 3] = -50.0          # two flat axes: dropped
    cols["scale_2"][pick == 4] = 70.0           # out of range: dropped
    cols["opacity"][pick == 5] = -8.0           # faint: dropped unless kept
    for k in range(4):
        cols[f"rot_{k}"][pick == 6] = 0.0       # no length: the identity
    cols["f_dc_0"][pick == 7] = 1.0e6           # past f16: clamped
    if 3 * per:
        cols["f_rest_0"][pick == 8] = np.nan    # a NaN coefficient clamps to -65000
    for n, t in props:
        code, dt = TYPES[t]
        if code >= 2:  # integer types: values of their range
            info = np.iinfo(np.dtype(dt))
            cols[n] = np.clip(np.nan_to_num(cols[n] * 10, nan=0, posinf=0, neginf=0), info.min, info.max)
    return props, header(props, count) + block(props, cols, count), cols


def dot_splat_record(pos=(0.5, -0.5, 2.0), scale=(0.25, 0.5, 1.0), rgba=(255, 0, 51, 255),
                     rot=(255, 128, 128, 128)) -> bytes:
    """One 32-byte .splat record (SplatReaders.cpp:215): position and linear
    scale as floats, rgba and rotation (w x y z) as bytes; the defaults are
    tests/scene/test_loading.cpp:168-199's record."""
    return struct.pack("<3f3f4B4B", *pos, *scale, *rgba, *rot)


def dot_splat_file(n: int, seed: int) -> bytes:
    """n seeded .splat records: sizes from tiny to large, every alpha byte
    (0 is too faint to keep), zero and negative scales (a flat axis, a hard
    surfel's p), and rotations of every byte."""
    rng = np.random.default_rng(seed)
    out = bytearray()
    for i in range(n):
        scale = list(np.exp(rng.normal(-3.0, 1.0, 3)).astype(np.float32))
        if i % 17 == 0:
            scale[i % 3] = 0.0
        if i % 23 == 0:
            scale[(i + 1) % 3] = -float(rng.uniform(1.0, 40.0))
        rgba = [int(v) for v in rng.integers(0, 256, 4)]
        rot = [int(v) for v in rng.integers(0, 256, 4)]
        if i % 29 == 0:
            rot = [128, 128, 128, 128]  # no length: the identity
        out += dot_splat_record(tuple(rng.uniform(-2, 2, 3)), tuple(scale), tuple(rgba), tuple(rot))
    return bytes(out)
