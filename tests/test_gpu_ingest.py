"""The PLY ingest on the GPU (athx_ingest_ply through the C ABI) against the
oracle's widen, validate, decode and pack of the same bytes: the kept count,
each kept splat's vertex, and every packed word bit-exact -- on the
reference's own fixture files (tests/data/splats/tiny.ply, sh3.ply), its
loading tests' hand-written files, and seeded clouds of every property type
with faint, NaN, flat, out-of-range and zero-length records among them. The
ingested clouds then render bit-exactly against the oracle through the packed
input path."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from athenea_amd import TileRasterizer, camera, packing, synthetic
from athenea_amd import ply as plymod
from oracle import pyoracle as po
from tests import ply_fixtures as F
from tests.test_gpu_packed import render_packed
from tests.test_gpu_parity import check_frame
from tests.test_ply_oracle import ingest_bytes

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def raster():
    r = TileRasterizer(0)
    yield r
    r.close()


def gpu_ingest(raster, data: bytes, max_degree: int = 3, keep_faint: bool = False):
    layout = plymod.parse_header(data)
    n, stride = int(layout.count), int(layout.stride)
    block = np.frombuffer(data, np.uint8, count=n * stride, offset=int(layout.data_offset))
    vertices = torch.from_numpy(block.copy()).to("cuda:0")
    out = plymod.ingest(vertices, layout, max_degree, keep_faint, raster)
    torch.cuda.synchronize()
    s = out.splats
    sh = s.sh.cpu().numpy().view(np.uint32) if s.sh is not None else np.zeros((s.count, 0), np.uint32)
    return (s.positions.cpu().numpy(), s.shape.cpu().numpy().view(np.uint32), sh,
            out.origin.cpu().numpy().view(np.uint32), out)


def check_same(raster, data: bytes, max_degree: int = 3, keep_faint: bool = False, label: str = ""):
    say("ingest", label, "degree", max_degree, "keep_faint", keep_faint, "bytes", len(data))
    gp, gs, gsh, go, out = gpu_ingest(raster, data, max_degree, keep_faint)
    op, os_, osh, oo = ingest_bytes(data, max_degree, keep_faint)
    assert len(gp) == len(op), (len(gp), len(op))
    assert np.array_equal(go, oo)
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32))
    assert np.array_equal(gs, os_)
    assert gsh.shape == osh.shape and np.array_equal(gsh, osh)
    return out


def test_reference_loading_file(raster):
    out = check_same(raster, F.loading_file(), label="loading")
    assert out.splats.count == 2 and list(out.origin.cpu().numpy()) == [0, 3]
    check_same(raster, F.loading_file(), max_degree=0, label="loading")
    check_same(raster, F.loading_file(), keep_faint=True, label="loading")


def test_reference_mixed_types_file(raster):
    check_same(raster, F.mixed_file(), label="mixed")


@pytest.mark.parametrize("name", ["tiny.ply", "sh3.ply"])
@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_reference_fixture_files(raster, name, degree):
    check_same(raster, (F.GOLDEN / name).read_bytes(), max_degree=degree, label=name)


@pytest.mark.parametrize("types", [
    {},
    {"x": "double", "y": "double", "z": "double", "opacity": "short", "scale_0": "char", "rot_0": "uchar"},
    {"f_dc_0": "ushort", "f_dc_1": "int", "f_dc_2": "uint", "rot_1": "double", "f_rest_4": "short"},
])
def test_random_clouds_of_every_type(raster, types):
    for seed, degree in ((1, 3), (2, 1), (3, 2), (4, 0)):
        _, data, _ = F.random_cloud(5000, seed, degree, types)
        for faint in (False, True):
            check_same(raster, data, max_degree=3, keep_faint=faint, label=f"random {seed}")
        check_same(raster, data, max_degree=1, label=f"random {seed} capped")


def test_large_cloud(raster):
    _, data, _ = F.random_cloud(1_000_003, 11, 3, {"opacity": "double"}, extra=False)
    check_same(raster, data, label="large")


def test_read_ply_from_a_file(raster, tmp_path):
    cloud = plymod.read_ply(F.GOLDEN / "sh3.ply", raster=raster)
    op, _, _, _ = ingest_bytes((F.GOLDEN / "sh3.ply").read_bytes())
    assert cloud.vertices == 64 and cloud.splats.count == len(op)
    truncated = tmp_path / "short.ply"
    truncated.write_bytes((F.GOLDEN / "sh3.ply").read_bytes()[:-10])
    with pytest.raises(plymod._lib.AthxError) as e:
        plymod.read_ply(truncated, raster=raster)
    assert e.value.code == "IoFailure"


@pytest.mark.parametrize("name", ["tiny.ply", "sh3.ply"])
def test_ingested_cloud_renders_bit_exact(raster, name):
    say("render ingested", name)
    _, _, _, _, out = gpu_ingest(raster, (F.GOLDEN / name).read_bytes())
    s = out.splats
    host = packing.PackedCloud(s.positions.cpu().numpy(), s.shape.cpu().numpy().view(np.uint32),
                               s.sh.cpu().numpy().view(np.uint32) if s.sh is not None else None, s.rest_per_colour)
    view = camera.view_for(synthetic.front_camera(), 256, 192)
    check_frame(*render_packed(raster, host, [view]))


@pytest.mark.parametrize("keep_faint", [False, True])
def test_dot_splat_files_are_bit_exact(raster, tmp_path, keep_faint):
    """.splat records (athx_splat_layout + athx_ingest_ply with the byte and
    linear encodings) against the oracle, through a file on disk: the
    reference's decode record first, then seeded records with faint, flat,
    hard-surfel and zero-length ones."""
    from tests.test_ply_oracle import ingest_splat_bytes

    say(".splat ingest, keep_faint", keep_faint)
    data = F.dot_splat_record() + F.dot_splat_file(50_001, 3)
    path = tmp_path / "cloud.splat"
    path.write_bytes(data)
    cloud = plymod.read_splats(path, raster=raster, keep_faint=keep_faint)
    torch.cuda.synchronize()
    op, os_, _, oo = ingest_splat_bytes(data, keep_faint)
    s = cloud.splats
    assert cloud.vertices == 50_002 and s.count == len(op) and s.rest_per_colour == 0
    assert np.array_equal(cloud.origin.cpu().numpy().view(np.uint32), oo)
    assert np.array_equal(s.positions.cpu().numpy().view(np.uint32), op.view(np.uint32))
    assert np.array_equal(s.shape.cpu().numpy().view(np.uint32), os_)
    assert keep_faint or len(op) < 50_002
    with pytest.raises(plymod._lib.AthxError) as e:
        bad = tmp_path / "bad.splat"
        bad.write_bytes(data[:-5])
        plymod.read_splats(bad, raster=raster)
    assert e.value.code == "InvalidArgument"
