"""orc_contribution pinned where the answer is known: one Gaussian, so each
pixel's blend weight is its alpha (T = 1), and mode 1 is the sum of the
pixels' round(alpha * 4096) computed here pixel by pixel; mode 4 with
dL/dC = 0 adds nothing to the colour words and the same weights in 1/256;
mode 3 with the mask at one counts the weight twice and at zero once; mode 2
with the drawn image equal to the splat's own share (nothing behind it) is the
splat's whole share, |alpha c|^2 capped at one, in 2^-24 -- carried into the
high word past 2^32."""
from __future__ import annotations

import numpy as np

from athenea_amd import camera, synthetic
from oracle import pyoracle as po


def one_gaussian():
    cloud = synthetic.plain_512_cloud(1)
    cloud.means[:] = 0.0
    cloud.log_scales[:] = np.log(0.05)
    cloud.quats[:] = [0.0, 0.0, 0.0, 1.0]
    cloud.opacity_logits[:] = 2.0
    return cloud


def test_one_gaussian_contributions_are_its_alphas():
    cloud = one_gaussian()
    view = camera.view_for(synthetic.front_camera(), 64, 48)
    ov = po.orc_view(view)
    frame = po.render([ov], cloud)
    rec = frame["records"][0]
    f32 = np.float32
    want1, want4 = 0, 0
    for py in range(48):
        for px in range(64):
            dx, dy = f32(px + 0.5) - rec[0], f32(py + 0.5) - rec[1]
            power = (rec[4] * dx * dx + f32(2.0) * rec[5] * dx * dy) + rec[6] * dy * dy
            if power > rec[3]:
                continue
            alpha = min(f32(0.99), f32(rec[7] * f32(po.load().orc_expf(float(f32(-0.5) * power)))))
            if alpha < f32(1.0) / f32(255.0):
                continue
            want1 += int(f32(alpha * f32(4096.0)) + f32(0.5))
            want4 += int(f32(alpha * f32(256.0)) + f32(0.5))
    assert want1 > 1000
    assert int(po.contribution(ov, frame, 1, mode=1)[0]) == want1
    out4 = po.contribution(ov, frame, 1, mode=4, dl=np.zeros((48, 64, 4), np.float32))
    assert list(out4[0, :3]) == [0, 0, 0] and int(out4[0, 3]) == want4


def test_one_gaussian_mask_and_removal_error():
    cloud = one_gaussian()
    cloud.log_scales[:] = np.log(0.4)
    cloud.sh[:, 0, :] = 12.0
    view = camera.view_for(synthetic.front_camera(), 64, 48)
    ov = po.orc_view(view)
    frame = po.render([ov], cloud)
    rec = frame["records"][0]
    f32 = np.float32
    drawn = np.zeros((48, 64, 4), np.float32)
    want3, want2 = 0, 0
    for py in range(48):
        for px in range(64):
            dx, dy = f32(px + 0.5) - rec[0], f32(py + 0.5) - rec[1]
            power = (rec[4] * dx * dx + f32(2.0) * rec[5] * dx * dy) + rec[6] * dy * dy
            if power > rec[3]:
                continue
            alpha = min(f32(0.99), f32(rec[7] * f32(po.load().orc_expf(float(f32(-0.5) * power)))))
            if alpha < f32(1.0) / f32(255.0):
                continue
            share = [f32(alpha * rec[8 + c]) for c in range(3)]
            drawn[py, px, :3] = share
            want3 += int(f32(alpha * f32(1024.0)) + f32(0.5))
            e2 = min(f32(f32(share[0] * share[0] + share[1] * share[1]) + share[2] * share[2]), f32(1.0))
            want2 += int(f32(e2 * f32(16777216.0)) + f32(0.5))
    assert want3 > 1000 and want2 > 2**32
    ones = po.contribution(ov, frame, 1, mode=3, dl=np.ones((48, 64), np.float32))
    zeros = po.contribution(ov, frame, 1, mode=3, dl=np.full((48, 64), -0.5, np.float32))
    assert list(ones[0]) == [want3, want3] and list(zeros[0]) == [0, want3]
    removal = po.contribution(ov, frame, 1, mode=2, dl=drawn)
    assert int(removal[0, 0]) + (int(removal[0, 1]) << 32) == want2
