#!/usr/bin/env python3
"""Benchmark: device-resident registrations of the 256x256x128 affine pair.

Workload (one "step" = one full registration, inputs already resident in
HBM): fixed and moving 256x256x128 float32 volumes at 1.0 mm, a phantom of 60
seeded ellipsoids, the moving image under a seeded affine transform (rotations
+/-10 deg, scales U[0.9,1.1], shears U[-0.05,0.05], translations +/-15 mm) with
a multimodal contrast remap; symmetric mode, affine model, blocks 4^3, search
+/-4 voxels (nhoodwidth 9, nhoodgap 3), 3 pyramid levels, 5 iterations per
level. Prints one JSON line: NCC voxel-evals/s over the timed registrations,
with registrations/s and p50/p99 latency beside it, the roofline of the NCC
search kernel (HIP events on the library stream) and a CPU baseline (the
oracle's threaded block matcher on one full-resolution pass).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_FP32_TFLOPS = 157.3      # MI355X dense fp32 (vector/matrix) peak
PEAK_HBM_GBS = 8000.0
FLOP_PER_VOXEL_EVAL = 6       # two-pass NCC: mean add, centring, 2 FMAs


WORKLOADS = {
    # The measured workload (default).
    "pair256_affine_symmetric": dict(
        workload="pair256_affine_symmetric", dims=[256, 256, 128], spacing_mm=1.0,
        ellipsoids=60, model="affine", mode="symmetric", block_width=4, nhood_width=9,
        nhood_gap=3, block_gap=1, pyramid_levels=3, iterations_per_level=5, seed=2),
    # CT-like int16 / MR-like uint16 rigid pair (float32 on the device).
    "pair512_ctmr_rigid_symmetric": dict(
        workload="pair512_ctmr_rigid_symmetric", dims=[512, 512, 256], spacing_mm=1.0,
        ellipsoids=24, model="rigid", mode="symmetric", block_width=4, nhood_width=13,
        nhood_gap=3, block_gap=1, pyramid_levels=4, iterations_per_level=5, seed=3),
}


def workload(name="pair256_affine_symmetric"):
    return dict(WORKLOADS[name])


def make_inputs(cfg):
    from mirorr_amd import phantom
    if cfg["workload"].startswith("pair512_ctmr"):
        return phantom.make_ct_mr_pair(dims=tuple(cfg["dims"]), spacing=(cfg["spacing_mm"],) * 3,
                                       n_ellipsoids=cfg["ellipsoids"], seed=cfg["seed"])
    return phantom.make_pair(dims=tuple(cfg["dims"]), spacing=(cfg["spacing_mm"],) * 3,
                             n_ellipsoids=cfg["ellipsoids"], seed=cfg["seed"], model=cfg["model"])


def configure(reg, cfg):
    reg.SetRegistrationTypeSymmetrical()
    reg.SetTransformType(cfg["model"])
    reg.GetPyramidScheduleTuner().SetMinLength(32)
    reg.SetBlockWidth(cfg["block_width"])
    reg.SetBlockGap(cfg["block_gap"])
    reg.SetNhoodWidth(cfg["nhood_width"])
    reg.SetNhoodGap(cfg["nhood_gap"])
    reg.SetMaxIterations(cfg["iterations_per_level"])
    reg.GetPyramidScheduleTuner().SetMaxLevelNumber(cfg["pyramid_levels"])
    reg.GetPyramidScheduleTuner().SetLevelMax(0)


TRAFFIC_SUMMARY = os.path.join(HERE, "profiles", "ncc_tile_fast_pair256_traffic.md")


def measured_traffic():
    """HBM bytes per full-resolution ncc_tile_fast launch from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), or None."""
    import re
    try:
        with open(TRAFFIC_SUMMARY) as f:
            m = re.search(r"traffic_bytes_per_full_resolution_launch:\s*([0-9.eE+]+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def cpu_baseline(pair, cfg, threads, max_evals=20e9):
    """Oracle threaded block matcher on one full-resolution block-matching
    pass (moving blocks searched in the fixed image), timed on the host
    cores. A pass larger than max_evals is timed on a centred sub-volume
    (half of each axis) of the same images."""
    from oracle import oracle as orc
    if not os.path.exists(orc._LIB_PATH):
        orc.build()
    mov, fix = pair.moving.data, pair.fixed.data
    bw, bg, nw, ng = cfg["block_width"], cfg["block_gap"], cfg["nhood_width"], cfg["nhood_gap"]
    what = "full-resolution"
    evals = orc.candidate_count(mov.shape[::-1], bw, bg, nw, ng) * bw ** 3
    if evals > max_evals:
        sl = tuple(slice(n // 4, n // 4 + n // 2) for n in mov.shape)
        mov, fix = np.ascontiguousarray(mov[sl]), np.ascontiguousarray(fix[sl])
        evals = orc.candidate_count(mov.shape[::-1], bw, bg, nw, ng) * bw ** 3
        what = "centred half-size sub-volume of the full-resolution"
    g = orc.Geom(dims=mov.shape[::-1], spacing=pair.moving.spacing, origin=pair.moving.origin)
    t = time.perf_counter()
    orc.bm_get_offsets(mov, None, fix, None, g, bw=bw, bg=bg, nw=nw, ng=ng, nthreads=threads)
    dt = time.perf_counter() - t
    dims = "x".join(str(d) for d in mov.shape[::-1])
    return {"value": evals / dt, "unit": "voxel-evals/s", "cores": threads, "kind": "port",
            "sample": f"one {what} block-matching pass, {dims} voxels "
                      f"({evals / 1e9:.2f} G voxel-evals, {dt:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="pair256_affine_symmetric", choices=sorted(WORKLOADS))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported "
              "(independent registrations shard across GPUs with no communication).",
              file=sys.stderr)
        sys.exit(2)

    from mirorr_amd.blockmatcher import default_context
    from mirorr_amd.registration import DeviceVolume, MirorrPyramid

    cfg = workload(args.workload)
    pair = make_inputs(cfg)
    ctx = default_context()
    t_up = time.perf_counter()
    fixed = DeviceVolume(pair.fixed, ctx)
    moving = DeviceVolume(pair.moving, ctx)
    upload_ms = (time.perf_counter() - t_up) * 1e3  # host -> HBM, outside the timed region
    reg = MirorrPyramid(ctx)
    configure(reg, cfg)
    print(f"bench: warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        reg.RunMirorr(fixed_dev=fixed, moving_dev=moving)
    ctx.synchronize()
    # The timed steps carry HIP events around the fast search kernel's
    # full-resolution launches only (they are priced on the roofline); the
    # masked kernel's share is measured on one more, untimed registration.
    ctx.set_profiling(ctx.PROF_FAST | ctx.PROF_FINEST_ONLY)
    ctx.profile_reset()
    lat = []
    evals = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        t = time.perf_counter()
        params, center = reg.RunMirorr(fixed_dev=fixed, moving_dev=moving)
        lat.append(time.perf_counter() - t)
        evals += reg.stats["voxel_evals"]
        print(f"bench: step {i + 1}/{args.steps} {lat[-1] * 1e3:.1f} ms "
              f"iterations {reg.stats['iterations']}", file=sys.stderr, flush=True)
    ctx.synchronize()
    total = time.perf_counter() - t0
    n_fast, ms_fast, work_fast = ctx.profile("ncc_tile_fast")
    ctx.set_profiling(ctx.PROF_FAST | ctx.PROF_MASKED | ctx.PROF_FINEST_ONLY)
    ctx.profile_reset()
    reg.RunMirorr(fixed_dev=fixed, moving_dev=moving)
    _, _, work_fast1 = ctx.profile("ncc_tile_fast")
    _, ms_mask, work_mask = ctx.profile("ncc_tile_masked")
    ctx.set_profiling(False)

    from mirorr_amd import transforms
    M, off = transforms.matrix_offset(params, center, cfg["model"] == "affine")
    err_mm = transforms.max_corner_distance(M, off, pair.matrix, pair.translation, pair.fixed)

    avg_ms = ms_fast / max(n_fast, 1)
    achieved = FLOP_PER_VOXEL_EVAL * (work_fast / max(n_fast, 1)) / (avg_ms * 1e-3) / 1e12
    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": "NCC voxel-evals/s + registrations/s (p50/p99 latency) vs plain GPU, 1 GPU",
        "value": evals / total,
        "unit": "voxel-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": cfg,
        "registrations_per_s": args.steps / total,
        "latency_ms_p50": float(np.percentile(lat_ms, 50)),
        "latency_ms_p99": float(np.percentile(lat_ms, 99)),
        "voxel_evals_per_registration": evals / args.steps,
        "transform_error_mm": err_mm,
        "upload_ms": upload_ms,
        "roofline": {
            "bound": "mfma",
            "compute": "fp32 vector ALU (the search uses no MFMA)",
            "kernel": "ncc_tile_fast",
            "launch_scope": "full-resolution level",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            # the committed traffic summary covers the default workload only
            "traffic": measured_traffic() if args.workload == "pair256_affine_symmetric" else None,
            "launches": n_fast,
            "avg_launch_ms": avg_ms,
            "flop_per_voxel_eval": FLOP_PER_VOXEL_EVAL,
            "masked_kernel_ms_full_resolution": ms_mask,
            "masked_kernel_share_of_evals_full_resolution": work_mask / max(work_fast1 + work_mask, 1.0),
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        threads = min(16, len(os.sched_getaffinity(0)))
        out["cpu_baseline"] = cpu_baseline(pair, cfg, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
