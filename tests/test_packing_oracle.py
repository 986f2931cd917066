"""Pins of the packed input layout (athenea's scene::GpuSplats, common/
packing.slang): the oracle's quaternion decode against the closed form and an
integer round trip of the 10-bit codes, binary16 fields exact, the normal and
pbr words, and a packed cloud drawing what its FP32 source draws up to the
rounding the packing makes."""
from __future__ import annotations

import numpy as np
import pytest

from athenea_amd import camera, packing, synthetic
from oracle import pyoracle as po


@pytest.fixture(scope="module", autouse=True)
def _lib():
    po.load()


def test_quaternion_decode_matches_the_closed_form_and_reencodes_exactly():
    rng = np.random.default_rng(7)
    q = rng.normal(size=(4000, 4)).astype(np.float32)
    words = packing.encode_quaternions(q)
    dec = np.stack([po.decode_quaternion(int(w)) for w in words[:400]])
    # The oracle's decode is the numpy restatement's, bit for bit.
    assert np.array_equal(dec.view(np.uint32), packing.decode_quaternions(words[:400]).view(np.uint32))
    # Closed form: the three stored components are unorms of [-sqrt(1/2), sqrt(1/2)],
    # the dropped one the unit length's rest; each is within half a code of the
    # normalised, sign-fixed input.
    qn = q[:400] / np.linalg.norm(q[:400], axis=1, keepdims=True)
    largest = np.argmax(np.abs(qn), axis=1)
    qn = qn * np.where(qn[np.arange(400), largest] < 0, -1, 1)[:, None]
    step = np.sqrt(0.5) * 2 / 1023
    err = np.abs(dec - qn)
    stored = np.ones_like(err, bool)
    stored[np.arange(400), largest] = False
    assert err[stored].max() <= 0.5 * step + 1e-6
    # The dropped component, sqrt(1 - |stored|^2) with it at least 1/2, moves by
    # at most sum(|s| ds) / 0.5 <= 3 sqrt(1/2) (step / 2) / 0.5.
    assert err[~stored].max() <= 3 * np.sqrt(0.5) * step + 1e-6
    # Decoding and encoding again gives the same word (the codes are exact),
    # except where a stored component came back larger than the rebuilt one:
    # then the encoder drops another component, as it must.
    dq = packing.decode_quaternions(words)
    again = packing.encode_quaternions(dq)
    flipped = np.argmax(np.abs(dq), axis=1) != (words & 3)
    assert np.array_equal(again[~flipped], words[~flipped])
    assert np.array_equal(again != words, flipped)
    assert flipped.mean() < 0.01
    # A rotation is kept to a tenth of a degree (packing.slang's own claim).
    dots = np.abs((packing.decode_quaternions(words) * (q / np.linalg.norm(q, axis=1, keepdims=True))).sum(1))
    assert np.degrees(2 * np.arccos(np.clip(dots, -1, 1))).max() < 0.2


def test_normal_and_pbr_words():
    rng = np.random.default_rng(8)
    n = rng.normal(size=(500, 3)).astype(np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    words = packing.pack_normals(n)
    back = np.stack([po.unpack_normal(int(w)) for w in words])
    # 16-bit octahedral: a few thousandths of a degree (packing.slang:113).
    cross = np.linalg.norm(np.cross(back.astype(np.float64), n.astype(np.float64)), axis=1)
    assert np.degrees(np.arcsin(np.minimum(cross, 1.0))).max() < 0.01
    assert ((back.astype(np.float64) * n).sum(1) > 0).all()
    assert np.abs(np.linalg.norm(back, axis=1) - 1).max() < 1e-6
    m = np.array([0.0, 1.0, 0.2, 0.5], np.float32)
    r = np.array([0.05, 1.0, 0.5, 0.0], np.float32)
    pbr = packing.pack_pbr(m, r)
    assert np.array_equal(pbr & 0xFF, np.floor(m * 255 + 0.5).astype(np.uint32))
    assert np.array_equal((pbr >> 8) & 0xFF, np.floor(r * 255 + 0.5).astype(np.uint32))


def test_halves_round_to_nearest_even():
    # 1 + 2^-11 is halfway between two halves: ties go to the even mantissa.
    vals = np.array([1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, 65504.0, -0.0, 6e-8], np.float32)
    h = packing.halves(vals)
    assert list(h[:2]) == [0x3C00, 0x3C02]
    assert h[2] == 0x7BFF and h[3] == 0x8000
    assert all(po.half_to_float(int(x)) == np.float16(v) for x, v in zip(h, vals))


@pytest.mark.parametrize("coeffs", [1, 4, 16])
def test_packed_cloud_draws_its_source(coeffs):
    """The packed layout rounds log-scales, colours and coefficients to
    halves and the rotation to 10 bits; the records it projects to stay
    within those roundings of the FP32 source's."""
    cloud = synthetic.plain_512_cloud(3000)
    cloud.sh = np.ascontiguousarray(cloud.sh[:, :coeffs, :])
    packed = packing.pack_cloud(cloud)
    assert packed.rest_per_colour == coeffs - 1
    view = camera.view_for(synthetic.front_camera(), 128, 96)
    ov = po.orc_view(view)
    rec_a, t_a, k_a = po.project(ov, cloud)
    rec_b, t_b, k_b = po.project(ov, packed)
    both = (t_a > 0) & (t_b > 0)
    assert both.sum() > 0.9 * (t_a > 0).sum()
    # Same place, depth key and (nearly) the same conic and colour.
    np.testing.assert_allclose(rec_b[both, 0:3], rec_a[both, 0:3], rtol=2e-5, atol=2e-4)
    # The conic moves with the half-rounded log-scales (about 0.2% a scale):
    # within 2% of its largest entry.
    scale = np.abs(rec_a[both, 4:7]).max(axis=1, keepdims=True)
    assert (np.abs(rec_b[both, 4:7] - rec_a[both, 4:7]) <= 0.02 * scale).all()
    np.testing.assert_allclose(rec_b[both, 8:11], rec_a[both, 8:11], atol=3e-3)
    # The whole frame through the oracle: images within the packing's noise.
    fa = po.render([ov], cloud)
    fb = po.render([ov], packed)
    assert np.abs(fa["colour"][0] - fb["colour"][0]).mean() < 2e-3
