"""Writes synthetic_head.npz: the first 16 Gaussians of each synthetic
workload's generator, so a change to the generator (or to numpy's PCG64
stream) fails a test instead of silently changing every benchmark input."""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from athenea_amd import synthetic  # noqa: E402

out = {}
for name, make in (("plain_512", synthetic.plain_512_cloud), ("plain_1080p", synthetic.plain_1080p_cloud),
                   ("walk", synthetic.walk_cloud), ("mixed", synthetic.mixed_cloud)):
    c = make(16)
    for field in ("means", "log_scales", "quats", "opacity_logits", "sh"):
        out[f"{name}_{field}"] = getattr(c, field)
np.savez(Path(__file__).with_name("synthetic_head.npz"), **out)
print("wrote", len(out), "arrays")
