"""Benchmark of the MI355X splat tile rasterizer.

A step is one forward frame of the 1080p workload (1.4M Gaussians, SH degree
3, FP32, 1920x1080, 16x16 tiles, each Gaussian with relight attributes --
albedo, roughness, metallic, a stored normal -- under a prepared 64x32
log-normal HDR sky): projection with the relit colour, the depth sort, the
binning of (tile, Gaussian) pairs through 16-tile strips, tile ranges and the
blend, with the Gaussians and the sky already resident in HBM. Prints ONE
JSON line:

  value         forward ms/frame (lower is better)
  parity        the timed frame against the oracle's frame of the same
                inputs (the reference's two-sort route, rendered by the
                cpu_baseline leg): pairs, pair values, ranges, image, final
                transmittance, walk lengths -- "bit-exact" or what differs
  fwd_sh_ms     the same frame coloured by the harmonics instead
  fwd_bwd_ms    forward + adjoint backward ms/frame of the SH-coloured frame
                (loss L2 against a target rendered from jittered means and SH)
  batched       the 60-view walk (2.36M Gaussians, 960x540 views): views/s
                as one captured graph, with the tile side and blend batch the
                autotuner picks, against the same views rendered one launch
                each
  mixed         20M Gaussians with FP16 SH, 8 views from 640x360 to 3840x2160
                in one frame: ms/frame, views/s, pairs, strip entries
  sort          athx_sort_pairs on 10M pairs with 32-bit keys (the
                reference's own sort benchmark size): keys/s
  stages_ms     per-stage times from HIP events on the rasterizer's stream
  roofline      the dominant kernel (blend_kernel): its VALU issue rate
                against MI355X's one wave instruction a CU a clock, from the
                committed rocprofv3 counters over its live event time; its
                HBM traffic beside it
  roofline_hbm  the projection's compulsory bytes over its event time
                against 8 TB/s (the largest HBM-bound kernel)
  cpu_baseline  the oracle (the FP32 C restatement of the reference's
                algorithm, its two-sort route) on this host's cores, one frame
                of the same workload, per stage
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12
# MI355X's published FP32 vector peak, 157.3 TFLOP/s, in wave64 instructions
# a second: an FMA instruction does 64 lanes x 2 FLOPs. The blend's rate is
# every VALU instruction it issues (SQ_INSTS_VALU, whole device, a dispatch)
# over its event-timed duration, against that.
FP32_PEAK_FLOPS = 157.3e12
VALU_PEAK = FP32_PEAK_FLOPS / 128.0  # wave instructions a second
METRIC = "ms/frame fwd & fwd+bwd, 1.4M splats@1080p; batched views/s; % HBM peak"
STAGES = ("project", "depth_sort", "counts", "emit", "tile_sort", "ranges", "blend")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_bytes(n: int, visible: int, pairs: int, entries: int, tiles: int, pixels: int, relit: bool = False) -> dict:
    """Compulsory HBM bytes per stage of one frame -- what each stage must
    read and write whatever its implementation (the per-unit figures of the
    rasterizer's algorithmic byte model: parameters and records once, (key,
    value) once in and out of each sort's passes)."""
    return {
        # FP32 parameters read (236 B at SH degree 3, + 32 B of relight
        # attributes); the 48 B record, depth key and entry count written
        "project": n * (236 + (32 if relit else 0) + 48 + 4 + 4),
        # depth keys read; visible (key, slot) compacted; a stable 24-bit sort
        # of them, 3 passes of a read and a write of 8 B
        "depth_sort": n * 4 + visible * 8 + 3 * visible * 16,
        "counts": visible * 8,
        # the sorted slot and the record's 32 B of geometry read a drawn
        # footprint; (strip | tile mask, slot) written an entry
        "emit": visible * (4 + 32) + entries * 8,
        # the entries read once; each pair's slot written
        "tile_sort": entries * 8 + pairs * 4,
        "ranges": tiles * 8,
        # each pair's slot and record read, the targets written
        "blend": pairs * (4 + 48) + pixels * 28,
    }


def committed_counters() -> dict:
    """profiles/counters_1080p_relit.json: rocprofv3 --pmc values a launch of
    the 1080p relit frame's kernels (VALU instructions, HBM bytes), or {}."""
    path = ROOT / "profiles" / "counters_1080p_relit.json"
    return json.loads(path.read_text()) if path.exists() else {}


def copy_probe() -> None:
    """One 256 MiB device-to-device copy: its FETCH_SIZE and WRITE_SIZE
    counters give the scale of the bytes those counters report."""
    import torch

    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(cloud_np, view, tile: int, relight) -> tuple[dict, dict]:
    """The oracle's reference-order frame, 1 warm-up and the median of 5 (the
    reference's own protocol, CmdRender.cpp:379-437), per stage; returns the
    baseline and the last frame (for the parity check)."""
    from oracle import pyoracle as po

    po.load()
    threads = po.threads()
    ov = po.orc_view(view, tile=tile)
    frame = po.render_ref_order(ov, cloud_np, relight=relight)  # warm-up
    totals, per_stage = [], []
    for _ in range(5):
        t0 = time.perf_counter()
        frame = po.render_ref_order(ov, cloud_np, relight=relight)
        totals.append((time.perf_counter() - t0) * 1e3)
        per_stage.append(frame["ms"])
    names = ("project", "depth_sort", "counts", "emit", "tile_sort", "ranges", "blend")
    stages = {k: round(statistics.median(float(ms[i]) for ms in per_stage), 2) for i, k in enumerate(names)}
    out = {
        "value": round(statistics.median(totals), 2),
        "unit": "ms/frame",
        "cores": threads,
        "kind": "port",
        "cpu": cpu_model(),
        "host_cpus": os.cpu_count(),
        "stages_ms": stages,
        "sample": "one full forward frame of the same relit 1.4M-Gaussian 1920x1080 workload through the oracle's "
                  "reference-order route (project with the relit colour, 24-bit stable depth sort, emit, stable "
                  "tile sort, ranges, blend), median of 5 after one warm-up; OpenMP over the threads in `cores` "
                  "for projection, emission and blend, the two radix sorts single-threaded",
    }
    return out, frame


def compare_with_oracle(gpu: dict, frame: dict) -> dict:
    """The GPU frame (host copies) against the oracle's: every part equal
    bit for bit, or the parts that are not."""
    import numpy as np

    checks = {
        "pairs": gpu["pairs"] == frame["pairs"],
        "pair_values": np.array_equal(gpu["pair_values"], frame["pair_values"]),
        "ranges": np.array_equal(gpu["ranges"], frame["ranges"]),
        "image": np.array_equal(gpu["colour"], frame["colour"]),
        "depth": np.array_equal(gpu["depth"], frame["depth"]),
        "final_t": np.array_equal(gpu["final_t"], frame["final_t"]),
        "last": np.array_equal(gpu["last"], frame["last"]),
    }
    bad = [k for k, ok in checks.items() if not ok]
    out = {"status": "bit-exact" if not bad else "differs", "checked": list(checks)}
    if bad:
        out["differs"] = bad
        out["max_abs_image"] = float(np.abs(gpu["colour"].astype(np.float64) - frame["colour"]).max())
    return out


def timed_frames(raster, views, cloud, settings, steps, warmup):
    import torch

    targets = None
    for _ in range(warmup):
        targets, stats = raster.render(views, cloud, settings, targets)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        targets, stats = raster.render(views, cloud, settings, targets)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps, stats, targets


def sort_leg(raster, n: int = 10_000_000, reps: int = 10) -> dict:
    """RadixSort::sort's benchmark size (tests/gpu/test_algo.cpp:381-407): 10M
    pairs with 32-bit keys, stable, timed with events on the stream."""
    import torch

    g = torch.Generator(device="cuda:0").manual_seed(7)
    keys0 = torch.randint(0, 1 << 32, (n,), dtype=torch.int64, device="cuda:0", generator=g)
    vals0 = torch.arange(n, dtype=torch.int32, device="cuda:0")
    keys, vals = keys0.clone(), vals0.clone()
    raster.sort_pairs(keys, vals, 32)  # warm-up, sizes the scratch
    torch.cuda.synchronize()
    ok = bool(torch.all(keys[1:] >= keys[:-1]).item())
    times = []
    stream = torch.cuda.current_stream()
    for _ in range(reps):
        keys.copy_(keys0)
        vals.copy_(vals0)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        raster.sort_pairs(keys, vals, 32)
        b.record(stream)
        b.synchronize()
        times.append(a.elapsed_time(b))
    ms = statistics.median(times)
    return {"pairs": n, "key_bits": 32, "ms": round(ms, 4), "keys_per_s": round(n / (ms * 1e-3)),
            "sorted": ok, "note": "64-bit key words holding 32-bit keys: 4 passes of 8 bits"}


def mixed_leg(raster, steps: int) -> dict:
    import numpy as np
    import torch

    from athenea_amd import RenderSettings, camera, synthetic
    from athenea_amd.raster import GaussianCloud

    log("mixed: generating 20M Gaussians (FP16 SH)")
    cloud_np = synthetic.mixed_cloud()
    cloud = GaussianCloud.from_numpy(cloud_np, "cuda:0")
    del cloud_np
    cams = synthetic.walk_cameras(60)[:8]
    views = [camera.view_for(c, w, h) for c, (w, h) in zip(cams, synthetic.MIXED_SIZES)]
    settings = RenderSettings()
    reps = max(2, steps // 4)
    ms, stats, targets = timed_frames(raster, views, cloud, settings, reps, 1)
    info = raster.frame_info()
    staged = RenderSettings(time_stages=True)
    targets, s = raster.render(views, cloud, staged, targets)
    finite = all(bool(torch.isfinite(t.colour).all().item()) for t in targets)
    del targets, cloud
    torch.cuda.empty_cache()
    raster.release_scratch()
    pixels = int(sum(w * h for w, h in synthetic.MIXED_SIZES))
    return {
        "workload": "mixed_8: 20M Gaussians, FP16 SH, 8 views (3840x2160, 1920x1080, 1280x720, 640x360, two each) "
                    "on the walk, one frame",
        "ms_per_frame": round(ms, 3),
        "views_per_s": round(8 / (ms * 1e-3), 1),
        "pixels": pixels,
        "pairs": int(stats.pairs),
        "visible": int(stats.visible),
        "strip_entries": int(info["entries"]),
        "depth_sort_keys_per_s": round(s.visible / (s.depth_sort_ms * 1e-3)) if s.depth_sort_ms > 0 else None,
        "stages_ms": {k: round(getattr(s, f"{k}_ms"), 3) for k in STAGES},
        "images_finite": finite,
        "parity": "tests/test_gpu_configs.py checks this shape bit-exact against the oracle at 2M Gaussians",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tile", type=int, default=16)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--no-extras", action="store_true", help="forward only: no backward, batch, mixed, sort or CPU legs")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 60-view walk")
    ap.add_argument("--no-mixed", action="store_true", help="skip the 20M mixed-resolution frame")
    ap.add_argument("--no-sort", action="store_true", help="skip the 10M-pair sort")
    ap.add_argument("--no-autotune", action="store_true", help="the walk at --tile/--batch instead of tuned")
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

    log(f"timing {args.steps} SH-coloured forward frames")
    fwd_sh_ms, _, _ = timed_frames(raster, [view], cloud_sh, settings, args.steps, args.warmup)
    log(f"timing {args.steps} relit forward frames")
    fwd_ms, stats, targets = timed_frames(raster, [view], cloud, settings, args.steps, args.warmup)
    pairs = int(stats.pairs)
    info = raster.frame_info()

    # Per-stage times from HIP events recorded on the rasterizer's stream:
    # "blend" is blend_kernel's launch alone.
    staged = RenderSettings(tile=args.tile, batch=args.batch, time_stages=True, count_splats=True)
    stage_samples = {k: [] for k in STAGES}
    for _ in range(max(3, min(args.steps, 10))):
        targets, s = raster.render([view], cloud, staged, targets)
        for k in stage_samples:
            stage_samples[k].append(getattr(s, f"{k}_ms"))
    stages = {k: statistics.median(v) for k, v in stage_samples.items()}
    counters = raster.latest_counters()
    # The frame kept for the parity check (host copies of the last staged frame).
    bufs = raster.frame_buffers()
    gpu_frame = {
        "pairs": int(s.pairs),
        "pair_values": bufs["pair_values"].cpu().numpy().view(np.uint32),
        "ranges": bufs["ranges"].cpu().numpy().view(np.uint32).reshape(-1),
        "colour": targets[0].colour.cpu().numpy(),
        "depth": targets[0].depth.cpu().numpy(),
    }
    del bufs
    # final_t and last: one more frame with them kept.
    kept, _ = raster.render([view], cloud, settings, keep_backward=True)
    gpu_frame["final_t"] = kept[0].final_t.cpu().numpy()
    gpu_frame["last"] = kept[0].last.cpu().numpy().view(np.uint32)
    walked = int(kept[0].last.to(torch.int64).sum().item())  # (pixel, pair) evaluations of the blend
    del kept
    torch.cuda.synchronize()

    tiles = info["tiles"]
    sbytes = stage_bytes(cloud.count, counters.visible, pairs, info["entries"], tiles, W * H, relit=True)
    prof = committed_counters().get("kernels", {})
    blend_prof = prof.get("blend_kernel", {})
    blend_s = stages["blend"] * 1e-3
    valu = blend_prof.get("SQ_INSTS_VALU")
    roofline = {
        "kernel": "blend_kernel",
        "bound": "valu",
        "achieved": round(valu / blend_s / 1e9, 1) if valu else None,
        "peak": round(VALU_PEAK / 1e9, 1),
        "unit": "G VALU wave-instructions/s",
        "frac": round(valu / blend_s / VALU_PEAK, 4) if valu else None,
        "traffic": blend_prof.get("hbm_bytes"),
        "hbm_frac": round(blend_prof["hbm_bytes"] / blend_s / HBM_PEAK, 4) if blend_prof.get("hbm_bytes") else None,
        "ms": round(stages["blend"], 4),
        "evaluations": walked,
        "evaluations_per_s": round(walked / blend_s),
        "valu_per_evaluation": round(valu * 64 / walked, 2) if valu else None,
        "peak_basis": "157.3 TFLOP/s FP32 / 128 FLOP a wave64 FMA instruction",
        "counters": "profiles/counters_1080p_relit.json (rocprofv3 --pmc, a launch of this frame)",
    }
    proj_s = stages["project"] * 1e-3
    proj_prof = prof.get("project_kernel relit", {})
    roofline_hbm = {
        "kernel": "project_kernel",
        "bound": "hbm",
        "achieved": round(sbytes["project"] / proj_s / 1e9, 1),
        "peak": HBM_PEAK / 1e9,
        "unit": "GB/s",
        "frac": round(sbytes["project"] / proj_s / HBM_PEAK, 4),
        "traffic": proj_prof.get("hbm_bytes"),
        "algorithmic_bytes": sbytes["project"],
        "valu_frac": round(proj_prof["SQ_INSTS_VALU"] / proj_s / VALU_PEAK, 4) if proj_prof.get("SQ_INSTS_VALU")
        else None,
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
        "stage_hbm_frac": {k: round(sbytes[k] / (stages[k] * 1e-3) / HBM_PEAK, 4) for k in STAGES if stages[k] > 0},
        "depth_sort_keys_per_s": round(counters.visible / (stages["depth_sort"] * 1e-3)),
        "pairs_binned_per_s": round(pairs / ((stages["emit"] + stages["tile_sort"]) * 1e-3)),
        "roofline": roofline,
        "roofline_hbm": roofline_hbm,
    }

    if args.copy_probe:
        copy_probe()

    def fwd_bwd_leg():
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

    def batch_leg():
        from athenea_amd.raster import FrameGraph

        log("batched 60-view walk")
        walk_np = synthetic.walk_cloud()
        walk = GaussianCloud.from_numpy(walk_np, dev)
        views = [camera.view_for(c, 960, 540) for c in synthetic.walk_cameras(60)]
        reps = max(2, args.steps // 5)
        tuned = settings
        table = None
        if not args.no_autotune:
            from athenea_amd.autotune import autotune

            log("autotuning the walk's tile side and blend batch")
            tuned, table = autotune(raster, views, walk, settings, reps=2)
            raster.release_scratch()
        batch_ms, bstats, _ = timed_frames(raster, views, walk, tuned, reps, 1)
        # The same batch as one captured graph: no host wait, one launch.
        graph = FrameGraph(views, walk, tuned)
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
        raster.release_scratch()
        # The plain single-view rasterizer: each view its own render, at the
        # reference's own tile and batch.
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
        del single
        result["batched"] = {
            "workload": "walk_60: 2.36M Gaussians, 60 views at 960x540 on a circular walk",
            "views_per_s": round(60 / (graph_ms * 1e-3), 1),
            "ms_per_view_graph": round(graph_ms / 60, 4),
            "ms_per_view_batched_eager": round(batch_ms / 60, 4),
            "ms_per_view_single": round(single_ms / 60, 4),
            "speedup_graph_vs_single": round(single_ms / graph_ms, 3),
            "tile": tuned.tile,
            "batch": tuned.batch,
            "autotune": table,
            "pairs": int(bstats.pairs),
            "graph_pairs_within_capacity": bool(gpairs <= gcap),
        }
        del walk
        torch.cuda.empty_cache()
        raster.release_scratch()

    def sort_leg_run():
        log("sort: 10M pairs, 32-bit keys")
        result["sort"] = sort_leg(raster)
        raster.release_scratch()

    def mixed_leg_run():
        result["mixed"] = mixed_leg(raster, args.steps)

    def cpu_leg():
        log("CPU baseline (the oracle on the host cores) and the parity check")
        base, frame = cpu_baseline(cloud_np, view, args.tile, attrs)
        result["cpu_baseline"] = base
        result["parity"] = compare_with_oracle(gpu_frame, frame)

    legs = [("fwd_bwd_ms", fwd_bwd_leg, False), ("batched", batch_leg, args.no_batch),
            ("sort", sort_leg_run, args.no_sort), ("mixed", mixed_leg_run, args.no_mixed),
            ("cpu_baseline", cpu_leg, args.no_cpu)]
    for name, fn, skip in legs:
        if args.no_extras or skip:
            continue
        try:
            fn()
        except Exception as e:  # a leg that fails is reported in the line; the headline stands
            import traceback

            traceback.print_exc()
            result.setdefault("leg_errors", {})[name] = f"{type(e).__name__}: {e}"
            torch.cuda.synchronize()
            raster.release_scratch()

    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
