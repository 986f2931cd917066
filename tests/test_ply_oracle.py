"""The PLY ingest's host logic and its oracle, on the CPU.

athx_ply_parse (the library's header reader: no device is touched) against
layouts worked out by hand and the reference's error cases
(modules/io/src/PlyHeader.cpp:59-128, SplatReaders.cpp:48-195); orc_ingest
against the reference's own loading assertions (tests/scene/
test_loading.cpp:93-165, 207-246) and against what the mathematics fixes
(numpy's IEEE half rounding, the quaternion code's quantisation bound).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from athenea_amd import _lib, packing
from athenea_amd.ply import parse_header
from oracle import pyoracle
from tests import ply_fixtures as F


def code_of(exc) -> str:
    return exc.value.code


# ---- the header ----

def test_reference_fixtures_parse_to_their_layouts():
    tiny = parse_header((F.GOLDEN / "tiny.ply").read_bytes())
    assert tiny.count == 64 and tiny.rest_per_colour == 0 and tiny.properties == 14 and tiny.stride == 56
    sh3 = parse_header((F.GOLDEN / "sh3.ply").read_bytes())
    assert sh3.count == 64 and sh3.rest_per_colour == 15
    # x y z nx ny nz f_dc_0..2 f_rest_0..44 opacity scale_0..2 rot_0..3: all float
    assert sh3.properties == 62 and sh3.stride == 62 * 4
    assert sh3.rest_base == 9 and sh3.opacity == 54 and list(sh3.rot) == [58, 59, 60, 61]
    data = (F.GOLDEN / "sh3.ply").read_bytes()
    assert sh3.data_offset == data.index(b"end_header\n") + len(b"end_header\n")
    assert sh3.data_offset + sh3.count * sh3.stride == len(data)


def test_offsets_and_types_follow_the_properties():
    layout = parse_header(F.mixed_file())
    assert layout.stride == 3 * 8 + 2 + 1 + 10 * 4
    assert [layout.offset[k] for k in range(5)] == [0, 8, 16, 24, 26]
    assert [layout.type[k] for k in range(5)] == [1, 1, 1, 5, 2]
    assert layout.offset[5] == 27  # scale_0: unaligned
    assert layout.rest_per_colour == 0


def test_an_element_before_the_vertices_moves_their_offset():
    props = [(n, "float") for n in ["x", "y", "z", "opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1",
                                    "rot_2", "rot_3", "f_dc_0", "f_dc_1", "f_dc_2"]]
    before = "element camera 3\nproperty double a\nproperty ushort b\n"
    data = F.header(props, 5, before=before)
    layout = parse_header(data)
    assert layout.data_offset == len(data) + 3 * 10
    with pytest.raises(_lib.AthxError) as e:
        parse_header(F.header(props, 5, before="element face 2\nproperty list uchar int vertex_indices\n"))
    assert code_of(e) == "Unsupported"


@pytest.mark.parametrize("text,code", [
    (b"not a ply\n", "InvalidArgument"),
    (b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty float x\n", "InvalidArgument"),
    (b"ply\nelement vertex 1\nproperty float x\nend_header\n", "InvalidArgument"),
    (b"ply\nproperty float x\nend_header\n", "InvalidArgument"),
    (b"ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nend_header\n", "Unsupported"),
    (b"ply\nformat binary_little_endian 1.0\nelement vertex 0\nproperty float x\nend_header\n", "InvalidArgument"),
    (b"ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty list uchar float x\nend_header\n",
     "Unsupported"),
])
def test_header_errors_are_the_references(text, code):
    with pytest.raises(_lib.AthxError) as e:
        parse_header(text)
    assert code_of(e) == code


def test_missing_or_unknown_properties_are_refused():
    props = [(n, t) for n, t in F.MIXED_PROPS if n != "f_dc_2"]
    with pytest.raises(_lib.AthxError) as e:
        parse_header(F.header(props, 1))
    assert code_of(e) == "InvalidArgument" and "f_dc_2" in str(e.value)
    props = [(n, "int64" if n == "unused" else t) for n, t in F.MIXED_PROPS]
    with pytest.raises(_lib.AthxError) as e:
        parse_header(F.header(props, 1))
    assert code_of(e) == "Unsupported"


def test_harmonics_must_be_a_whole_contiguous_degree():
    base = [(n, "float") for n in ["x", "y", "z", "opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1",
                                   "rot_2", "rot_3", "f_dc_0", "f_dc_1", "f_dc_2"]]
    rest = [(f"f_rest_{k}", "float") for k in range(24)]
    assert parse_header(F.header(base + rest, 1)).rest_per_colour == 8
    # 12 coefficients: degree 1 needs 9, and 12 / 3 = 4 a colour is no degree
    with pytest.raises(_lib.AthxError) as e:
        parse_header(F.header(base + rest[:12], 1))
    assert code_of(e) == "Unsupported"
    # green's first coefficient out of place
    swapped = rest[:9]
    swapped[3], swapped[4] = swapped[4], swapped[3]
    with pytest.raises(_lib.AthxError) as e:
        parse_header(F.header(base + swapped, 1))
    assert code_of(e) == "Unsupported"


# ---- the oracle's decode and pack ----

def unpack(positions, shape, sh, j):
    """test_loading.cpp's `inspect`: position, opacity, scales, the
    quaternion (x y z w), the base colour and the rest coefficients."""
    half = lambda h: float(np.array([h], np.uint16).view(np.float16)[0])
    s = shape[j]
    ls = [half(s[1] & 0xFFFF), half(s[1] >> 16), half(s[2] & 0xFFFF)]
    q = pyoracle.decode_quaternion(int(s[0]))
    base = [half(s[2] >> 16), half(s[3] & 0xFFFF), half(s[3] >> 16)]
    rest = []
    for w in sh[j]:
        rest += [half(w & 0xFFFF), half(w >> 16)]
    return list(positions[j]) + [math.exp(v) for v in ls] + list(q) + base + rest


def ingest_bytes(data: bytes, max_degree: int = 3, keep_faint: bool = False):
    layout = parse_header(data)
    per = int(layout.rest_per_colour)
    keep = min(per, (0, 3, 8, 15)[max_degree])
    fields = [(layout.offset[k], layout.type[k]) for k in
              [layout.x, layout.y, layout.z, layout.opacity, *layout.scale, *layout.rot, *layout.dc]]
    for basis in range(keep):
        for colour in range(3):
            k = layout.rest_base + colour * per + basis
            fields.append((layout.offset[k], layout.type[k]))
    raw = np.frombuffer(data, np.uint8, offset=layout.data_offset)
    return pyoracle.ingest(raw, layout.count, layout.stride, fields, keep_faint)


def test_oracle_meets_the_references_decode_assertions():
    positions, shape, sh, origin = ingest_bytes(F.loading_file())
    assert len(positions) == 2 and list(origin) == [0, 3]
    assert sh.shape[1] == (3 * 3 + 1) // 2  # degree 1 kept
    a = unpack(positions, shape, sh, 0)
    assert a[0:3] == [1.0, 2.0, 3.0]
    assert a[3] == pytest.approx(0.5, abs=1e-6)
    assert a[4] == pytest.approx(1.0, rel=2e-3) and a[5] == pytest.approx(2.0, rel=2e-3)
    assert a[6] == pytest.approx(0.5, rel=2e-3)
    assert a[10] == pytest.approx(1.0, abs=1e-3)  # w
    assert a[11] == pytest.approx(0.5, abs=1e-3)
    assert a[14:18] == pytest.approx([0.1, 0.4, 0.7, 0.2], abs=1e-3)  # rgb a basis
    b = unpack(positions, shape, sh, 1)
    assert b[0] == -4.0
    assert b[3] == pytest.approx(0.75, abs=1e-5)
    assert abs(b[9]) == pytest.approx(1.0, abs=1e-3)  # |z| = 1
    assert b[11] == pytest.approx(1.5, abs=1e-3)
    # degree capped at 0: no harmonics
    positions0, _, sh0, _ = ingest_bytes(F.loading_file(), max_degree=0)
    assert len(positions0) == 2 and sh0.shape[1] == 0


def test_oracle_widens_mixed_types_as_the_reference():
    positions, shape, sh, _ = ingest_bytes(F.mixed_file())
    a, b = unpack(positions, shape, sh, 0), unpack(positions, shape, sh, 1)
    assert a[0] == 1.5 and a[1] == -2.25 and a[2] == float(np.float32(1.0e-3))
    assert b[0] == 2.5 and b[2] == float(np.float32(2.0e-3))
    assert a[3] == pytest.approx(0.5, abs=1e-6)
    assert b[3] == pytest.approx(0.8807971, abs=1e-5)
    assert a[5] == pytest.approx(2.0, rel=2e-3)


def test_keep_faint_keeps_every_finite_position():
    positions, _, _, origin = ingest_bytes(F.loading_file(), keep_faint=True)
    assert list(origin) == [0, 1, 3]  # the faint one kept, the NaN one not


def test_half_rounding_is_ieee_to_nearest_even():
    rng = np.random.default_rng(5)
    values = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 1e-5, 2000), rng.normal(0, 3e4, 2000),
                             [0.0, -0.0, 65000.0, -65000.0, 65504.0, 1e9, -1e9, 6.1035156e-05, 5.96e-08, 2.98e-08,
                              2.9802322e-08, 8.940697e-08]]).astype(np.float32)
    for v in values:
        clamped = np.float32(min(max(v, -65000.0), 65000.0))
        want = int(np.array([clamped], np.float32).astype(np.float16).view(np.uint16)[0])
        assert pyoracle.half_bits(float(v)) == want, v
    assert pyoracle.half_bits(float("nan")) == int(np.array([-65000.0], np.float16).view(np.uint16)[0])


def test_quaternion_codes_round_trip_within_their_step():
    rng = np.random.default_rng(9)
    q = rng.normal(0, 1, (3000, 4)).astype(np.float32)
    q[:20] = np.eye(4, dtype=np.float32)[rng.integers(0, 4, 20)] * rng.choice([-1, 1], (20, 1)).astype(np.float32)
    for row in q:
        word = pyoracle.encode_quaternion(*row)
        back = pyoracle.decode_quaternion(word)
        unit = row / np.linalg.norm(row.astype(np.float64))
        # the same rotation (q and -q), each small component within half a
        # 10-bit step of sqrt(1/2) * 2 / 1023, the largest rebuilt from them
        err = min(np.abs(back - unit).max(), np.abs(back + unit).max())
        assert err < 3e-3, (row, back)
    # the numpy encoder (host-side packing) agrees code for code but for
    # rounding at a step's midpoint
    words = packing.encode_quaternions(q)
    ours = np.array([pyoracle.encode_quaternion(*row) for row in q], np.uint32)
    same = words == ours
    assert same.mean() > 0.99
    for w, o in zip(words[~same], ours[~same]):
        assert (w & 3) == (o & 3)
        for shift in (22, 12, 2):
            assert abs(int((w >> shift) & 1023) - int((o >> shift) & 1023)) <= 1


# ---- .splat files ----

def splat_layout(size: int):
    layout = _lib.PlyLayout()
    message = __import__("ctypes").create_string_buffer(256)
    status = _lib.load().athx_splat_layout(size, __import__("ctypes").byref(layout), message, 256)
    return status, layout, message.value.decode()


def ingest_splat_bytes(data: bytes, keep_faint: bool = False):
    status, layout, _ = splat_layout(len(data))
    assert status == 0
    fields = [(layout.offset[k], layout.type[k]) for k in
              [layout.x, layout.y, layout.z, layout.opacity, *layout.scale, *layout.rot, *layout.dc]]
    modes = (layout.opacity_mode, layout.scale_mode, layout.colour_mode, layout.rotation_mode)
    return pyoracle.ingest(np.frombuffer(data, np.uint8), layout.count, layout.stride, fields, keep_faint, modes)


def test_splat_layout_is_the_references_record():
    status, layout, _ = splat_layout(96)
    assert status == 0 and layout.count == 3 and layout.stride == 32 and layout.data_offset == 0
    assert [layout.offset[k] for k in range(14)] == [0, 4, 8, 12, 16, 20, 24, 25, 26, 27, 28, 29, 30, 31]
    assert [layout.type[k] for k in range(14)] == [0] * 6 + [2] * 8
    assert (layout.x, layout.y, layout.z, list(layout.scale), list(layout.dc), layout.opacity, list(layout.rot)) == \
        (0, 1, 2, [3, 4, 5], [6, 7, 8], 9, [10, 11, 12, 13])
    assert (layout.opacity_mode, layout.scale_mode, layout.colour_mode, layout.rotation_mode) == (2, 1, 2, 1)
    assert layout.rest_per_colour == 0


@pytest.mark.parametrize("size", [0, 31, 33, 100])
def test_splat_sizes_that_are_not_whole_records_are_refused(size):
    status, _, message = splat_layout(size)
    assert status == 3 and "32-byte .splat records" in message


def test_oracle_meets_the_references_splat_decode_assertions():
    """test_loading.cpp:168-199: byte opacity 255 -> 1, scales 0.25 .. 1,
    w from byte 255 (0.992, normalised) -> 1, colour bytes 255 and 51 -> 1
    and 0.2."""
    positions, shape, sh, origin = ingest_splat_bytes(F.dot_splat_record())
    assert len(positions) == 1 and list(origin) == [0]
    v = unpack(positions, shape, sh, 0)
    assert v[0:3] == [0.5, -0.5, 2.0]
    assert v[3] == pytest.approx(1.0, abs=1e-6)
    assert v[4] == pytest.approx(0.25, rel=2e-3) and v[6] == pytest.approx(1.0, rel=2e-3)
    assert v[10] == pytest.approx(1.0, abs=1e-2)
    assert v[11] == pytest.approx(1.0, abs=1e-3)
    assert v[12] == pytest.approx(0.0, abs=1e-3)
    assert v[13] == pytest.approx(0.2, abs=1e-3)


def test_splat_decodes_flat_and_hard_axes_and_drops_faint_records():
    data = b"".join([F.dot_splat_record(rgba=(10, 20, 30, 0)),            # alpha 0: too faint
                     F.dot_splat_record(scale=(0.0, 0.5, 1.0)),            # a flat axis
                     F.dot_splat_record(scale=(0.2, -8.0, 1.0)),           # a hard surfel's p = 8
                     F.dot_splat_record(scale=(0.0, 0.0, 1.0)),            # two flat axes: invalid
                     F.dot_splat_record(rot=(128, 128, 128, 128))])        # no length: the identity
    positions, shape, sh, origin = ingest_splat_bytes(data)
    assert list(origin) == [1, 2, 4]
    half = lambda h: float(np.array([h], np.uint16).view(np.float16)[0])
    # a flat axis: -inf, clamped into the half range as packing.slang's halfSafe does (exp of it is 0)
    assert half(shape[0][1] & 0xFFFF) == -64992.0
    # -(1024 + 64 (8 - 2)) as a half
    assert half(shape[1][1] >> 16) == -(1024.0 + 64.0 * 6.0)
    assert unpack(positions, shape, sh, 2)[7:11] == pytest.approx([0.0, 0.0, 0.0, 1.0], abs=1e-3)
    _, _, _, origin_all = ingest_splat_bytes(data, keep_faint=True)
    assert list(origin_all) == [0, 1, 2, 3, 4]
