"""Benchmark of the MI355X splat tile rasterizer.

A step is one forward frame of the 1080p workload (1.4M Gaussians, SH degree
3, FP32, 1920x1080, 16x16 tiles, each Gaussian with relight attributes --
albedo, roughness, metallic, a stored normal -- under a prepared 64x32
log-normal HDR sky): projection with the relit colour, the depth sort, the binning of
(tile, Gaussian) pairs through 16-tile strips, tile ranges and the blend, with the Gaussians and the sky
already resident in HBM. Prints ONE JSON line:

  value         forward ms/frame (lower is better)
  fwd_sh_ms     the same frame coloured by the harmonics instead
  fwd_bwd_ms    forward + adjoint backward ms/frame of the SH-coloured frame
                (loss L2 against a target rendered from jittered means and SH)
  batched       the 60-view walk (2.36M Gaussians, 960x540 views) in one
                launch: views/s, and the per-view time of the same views
                rendered one launch each
  stages_ms     per-stage times from HIP events on the rasterizer's stream
  roofline      the dominant stage's algorithmic bytes / its time vs 8 TB/s
  cpu_baseline  the oracle (the FP32 C restatement of the reference's
                algorithm) on this host's cores, one frame of the same workload
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12
METRIC = "ms/frame fwd & fwd+bwd, 1.4M splats@1080p; batched views/s; % HBM peak"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_bytes(n: int, visible: int, pairs: int, entries: int, strips: int, tiles: int, pixels: int,
                relit: bool = False) -> dict:
    """Compulsory HBM bytes per stage (one view): what each stage must read
    and write whatever its implementation."""
    depth_passes = 3
    strip_passes = (max(strips - 1, 1).bit_length() + 7) // 8
    segments = entries // 1024 + strips
    return {
        # FP32 parameters read (236 B at SH degree 3, + 32 B of relight
        # attributes: albedo, normal, metallic, roughness); record, depth key
        # and strip-entry count written
        "project": n * (236 + (32 if relit else 0) + 48 + 4 + 4),
        # settling and compaction (counts and keys read, visible (key, slot)
        # written), then
        # a stable 24-bit sort of the visible u32 keys: a histogram read and a
        # read and write of (key, slot) a pass
        "depth_sort": n * 8 + visible * 8 + visible * 4 + depth_passes * visible * 16,
        # sorted slots read, entry counts gathered
        "counts": visible * 8,
        # sorted slot and the record's centre, conic and cut power read a rank
        # (the tile rectangle is recomputed from them); (strip | tile mask,
        # slot) written an entry; the strip sort: a histogram read, then a
        # read and a write of (key, slot) a pass
        "emit": visible * (4 + 32) + entries * 8 + entries * 4 + strip_passes * entries * 16,
        # an entry's key read and its mask kept (count); the table written,
        # scanned (read and written) and read; the mask and slot read again
        # and each pair's slot written at its place
        "tile_sort": entries * (4 + 2) + segments * 16 * 4 * 4 + entries * (2 + 4) + pairs * 4,
        "ranges": tiles * 8,
        # each pair's slot read, each drawn record once (the blend's re-reads
        # of a record hit the caches); RGBA, depth, final T, last index written
        "blend": pairs * 4 + visible * 48 + pixels * 28,
    }


# The kernels each timed stage launches, in order.
STAGE_KERNELS = {
    "project": "project_kernel",
    "depth_sort": "settle_count_kernel (unsettled footprints settled, each block's drawn slots counted), the "
                  "compaction, histogram_kernel, hist_reduce_kernel, onesweep_kernel x3",
    "counts": "the strip-entry scan in depth order",
    "emit": "strip_mask_emit_kernel (entries with their tiles decided), the strip sort (onesweep or "
            "reduce-scan-scatter passes), ranges_kernel over the strips, the segment scan, seg_desc_kernel",
    "tile_sort": "strip_mask_count_kernel, the table scan, strip_pairs_kernel",
    "ranges": "tile_ranges_kernel",
    "blend": "blend_kernel",
}


def measured_traffic(stage: str):
    """HBM bytes a frame's stage moves, from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE summed over the stage's kernels of
    one frame, each scaled by what the counter reads for a copy of known
    size): profiles/traffic_1080p_relit.json. None when absent."""
    path = ROOT / "profiles" / "traffic_1080p_relit.json"
    if not path.exists():
        return None
    stages = json.loads(path.read_text()).get("stages", {})
    entry = stages.get(stage)
    return int(entry["fetch_bytes"] + entry["write_bytes"]) if entry else None


def copy_probe() -> None:
    """One 256 MiB device-to-device copy: its FETCH_SIZE and WRITE_SIZE
    counters give the scale of the bytes those counters report."""
    import torch

    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()


def cpu_baseline(cloud_np, view, tile: int, relight) -> dict:
    from oracle import pyoracle as po

    po.load()
    threads = po.threads()
    ov = po.orc_view(view, tile=tile)
    times = []
    po.render_ref_order(ov, cloud_np, relight=relight)  # warm-up
    for _ in range(3):
        t0 = time.perf_counter()
        po.render_ref_order(ov, cloud_np, relight=relight)
        times.append((time.perf_counter() - t0) * 1e3)
    return {
        "value": round(statistics.median(times), 2),
        "unit": "ms/frame",
        "cores": threads,
        "kind": "port",
        "sample": "one full forward frame of the same relit 1.4M-Gaussian 1920x1080 workload (project with the "
                  "relit colour, 24-bit depth sort, emit, stable tile sort, ranges, blend), median of 3 after one "
                  f"warm-up, OpenMP over {threads} host threads",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tile", type=int, default=16)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--no-extras", action="store_true", help="forward only: no backward, batch or CPU legs")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 60-view walk")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 256 MiB device-to-device copy, the known-size kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc run of this command")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the rasterizer shards by views or screen tiles across GPUs, "
              "which this build does not run; use --gpus 1", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from athenea_amd import RenderSettings, TileRasterizer, camera, sky as skymod, synthetic
    from athenea_amd.raster import GaussianCloud

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    log("generating the 1080p workload and preparing its sky")
    cloud_np = synthetic.plain_1080p_cloud()
    sky = skymod.prepare_sky(skymod.lognormal_equirect())
    attrs = skymod.relight_attributes(cloud_np.means.shape[0], sky)
    W, H = 1920, 1080
    view = camera.view_for(synthetic.front_camera(), W, H)
    cloud_sh = GaussianCloud.from_numpy(cloud_np, dev)
    cloud = GaussianCloud.from_numpy(cloud_np, dev, relight=attrs)
    settings = RenderSettings(tile=args.tile, batch=args.batch)
    raster = TileRasterizer(0)

    def time_forward(c, steps, warmup):
        targets = None
        for _ in range(warmup):
            targets, stats = raster.render([view], c, settings, targets)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            targets, stats = raster.render([view], c, settings, targets)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps, stats, targets

    log(f"timing {args.steps} SH-coloured forward frames")
    fwd_sh_ms, _, _ = time_forward(cloud_sh, args.steps, args.warmup)
    log(f"timing {args.steps} relit forward frames")
    fwd_ms, stats, targets = time_forward(cloud, args.steps, args.warmup)
    pairs = int(stats.pairs)
    info = raster.frame_info()

    # Per-stage times from HIP events recorded on the rasterizer's stream.
    staged = RenderSettings(tile=args.tile, batch=args.batch, time_stages=True, count_splats=True)
    stage_samples = {k: [] for k in ("project", "depth_sort", "counts", "emit", "tile_sort", "ranges", "blend")}
    for _ in range(max(3, min(args.steps, 10))):
        targets, s = raster.render([view], cloud, staged, targets)
        for k in stage_samples:
            stage_samples[k].append(getattr(s, f"{k}_ms"))
    stages = {k: statistics.mean(v) for k, v in stage_samples.items()}
    counters = raster.latest_counters()
    tiles = info["tiles"]
    sbytes = stage_bytes(cloud.count, counters.visible, pairs, info["entries"], info["strips"], tiles, W * H,
                         relit=True)
    dominant = max(stages, key=stages.get)
    achieved = sbytes[dominant] / (stages[dominant] * 1e-3)
    roofline = {
        "kernel": f"{dominant} ({STAGE_KERNELS[dominant]})",
        "bound": "hbm",
        "achieved": round(achieved / 1e9, 1),
        "peak": HBM_PEAK / 1e9,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK, 4),
        "traffic": measured_traffic(dominant),
        "algorithmic_bytes": sbytes[dominant],
    }
    result = {
        "metric": METRIC,
        "value": round(fwd_ms, 3),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fwd_ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded, no checkpoint)",
        "config": {
            "workload": "1080p relit: 1.4M Gaussians, SH degree 3, FP32, 1920x1080, one view, relight "
                        "attributes (albedo, roughness, metallic, normal) under a prepared 64x32 HDR sky",
            "gaussians": cloud.count,
            "width": W,
            "height": H,
            "tile": args.tile,
            "batch": args.batch,
        },
        "fwd_sh_ms": round(fwd_sh_ms, 3),
        "pairs": pairs,
        "strip_entries": info["entries"],
        "visible": counters.visible,
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "pairs_binned_per_s": round(pairs / ((stages["emit"] + stages["tile_sort"]) * 1e-3)),
        "roofline": roofline,
    }

    if args.copy_probe:
        copy_probe()

    if not args.no_extras:
        from athenea_amd.autograd import rasterize

        log("forward + backward")
        rng = np.random.default_rng(2)
        tgt = synthetic.plain_1080p_cloud()
        tgt.means = (tgt.means + rng.normal(0, 0.01, tgt.means.shape)).astype(np.float32)
        tgt.sh = (tgt.sh + rng.normal(0, 0.05, tgt.sh.shape)).astype(np.float32)
        target_img = raster.render([view], GaussianCloud.from_numpy(tgt, dev), settings)[0][0].colour.clone()
        del tgt
        params = [torch.from_numpy(a).to(dev).requires_grad_(True) for a in
                  (cloud_np.means, cloud_np.log_scales, cloud_np.quats, cloud_np.opacity_logits, cloud_np.sh)]

        def step():
            for p in params:
                p.grad = None
            colours, _ = rasterize(*params, [view], raster, settings)
            loss = ((colours[0][..., :3] - target_img[..., :3]) ** 2).sum()
            loss.backward()

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        result["fwd_bwd_ms"] = round((time.perf_counter() - t0) * 1e3 / args.steps, 3)
        result["fwd_bwd_over_fwd"] = round(result["fwd_bwd_ms"] / fwd_sh_ms, 3)
        del params

    if not args.no_extras and not args.no_batch:
        log("batched 60-view walk")
        walk_np = synthetic.walk_cloud()
        walk = GaussianCloud.from_numpy(walk_np, dev)
        views = [camera.view_for(c, 960, 540) for c in synthetic.walk_cameras(60)]
        bt = None
        for _ in range(2):
            bt, bstats = raster.render(views, walk, settings, bt)
        torch.cuda.synchronize()
        reps = max(2, args.steps // 5)
        t0 = time.perf_counter()
        for _ in range(reps):
            bt, bstats = raster.render(views, walk, settings, bt)
        torch.cuda.synchronize()
        batch_ms = (time.perf_counter() - t0) * 1e3 / reps
        del bt
        # The same batch as one captured graph: no host wait, one launch.
        from athenea_amd.raster import FrameGraph

        graph = FrameGraph(views, walk, settings)
        graph.launch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            graph.launch()
        torch.cuda.synchronize()
        graph_ms = (time.perf_counter() - t0) * 1e3 / reps
        gpairs, gcap = graph.pairs()
        graph.close()
        del graph
        single = [None] * len(views)
        for k, v in enumerate(views):
            single[k], _ = raster.render([v], walk, settings, single[k])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            for k, v in enumerate(views):
                single[k], _ = raster.render([v], walk, settings, single[k])
        torch.cuda.synchronize()
        single_ms = (time.perf_counter() - t0) * 1e3 / reps
        result["batched"] = {
            "workload": "walk_60: 2.36M Gaussians, 60 views at 960x540 on a circular walk",
            "views_per_s": round(60 / (graph_ms * 1e-3), 1),
            "ms_per_view_graph": round(graph_ms / 60, 4),
            "ms_per_view_batched_eager": round(batch_ms / 60, 4),
            "ms_per_view_single": round(single_ms / 60, 4),
            "speedup_graph_vs_single": round(single_ms / graph_ms, 3),
            "pairs": int(bstats.pairs),
            "graph_pairs_within_capacity": bool(gpairs <= gcap),
        }
        raster.release_scratch()

    if not args.no_extras and not args.no_cpu:
        log("CPU baseline (the oracle on the host cores)")
        result["cpu_baseline"] = cpu_baseline(cloud_np, view, args.tile, attrs)

    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
