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
search kernel (HIP events on the library stream, taken on further untimed
registrations), a CPU baseline (the oracle's threaded block matcher on one
full-resolution pass, on the host cores and on one core), the oracle's
own registration of the pair, against which the final transform is
compared, and a served batch: 16 distinct pairs of the workload through the
registration server (float32 and binary16 storage) against the plain path
registering them one at a time.
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
PEAK_LDS_GBS = 128 * 256 * 2.4  # 128 B per clock per CU, 256 CUs, 2.4 GHz
# LDS bytes the fast search kernel reads per voxel-eval (NW 9 tile, KY 3,
# rows of 12 floats): per lane and block z-plane, KY + 3 window rows of 48 B
# and the 64 B block plane serve KY * NW * 16 voxel-evals.
LDS_BYTES_PER_VOXEL_EVAL_NW9 = ((3 + 3) * 48 + 64) / (3 * 9 * 16)
FLOP_PER_VOXEL_EVAL = 6       # two-pass NCC: mean add, centring, 2 FMAs
ROOFLINE_REGISTRATIONS = 3    # untimed registrations whose fast-kernel launches are timed


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


# Served workloads: many registrations through the registration server
# (mirorr_submit / mirorr_poll), against the plain one-pair-at-a-time path.
SERVED = {
    # 64 pairs of uneven size; even-numbered pairs rigid with the 128x128x64
    # workload's parameters, odd-numbered affine with the 256x256x128 one's.
    "batch64_uneven": dict(
        workload="batch64_uneven", pairs=64, axis_lengths=[96, 320, 16],
        spacing_mm=[0.8, 2.0], slots=8, seed=7,
        rigid=dict(ellipsoids=20, nhood_width=7, nhood_gap=4, pyramid_levels=2,
                   portion_kept=0.5, variance_select=1),
        affine=dict(ellipsoids=60, nhood_width=9, nhood_gap=3, pyramid_levels=3,
                    iterations_per_level=5)),
    # 16 pairs of the 256x256x128 affine workload (distinct seeds), all
    # submitted at once.
    "batch16_pair256_affine": dict(
        workload="batch16_pair256_affine", pairs=16, dims=[256, 256, 128], spacing_mm=1.0,
        slots=8, seed=2,
        affine=dict(ellipsoids=60, nhood_width=9, nhood_gap=3, pyramid_levels=3,
                    iterations_per_level=5)),
    # Poisson arrivals of CT/MR-like rigid requests, 30% at 512x512x256 and
    # the rest at 256x256x128, drawn from a pool of distinct pairs.
    "serve_poisson_ctmr": dict(
        workload="serve_poisson_ctmr", rates_per_s=[0.5, 1.0, 2.0, 4.0],
        requests_per_rate=100, large_fraction=0.3, large_dims=[512, 512, 256],
        small_dims=[256, 256, 128], pool_large=2, pool_small=3, slots=4, seed=11,
        nhood_width=13, nhood_gap=3, pyramid_levels=4, ellipsoids=24),
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


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo, as lscpu prints it)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def oracle_transform_check(pair, cfg, params, center, iterations, points, threads):
    """The oracle's registration of the same pair with the same parameters
    (threaded CPU restatement of the reference), and the distance between its
    final transform and the GPU's over the field-of-view corners."""
    from mirorr_amd import transforms
    from oracle import oracle as orc
    if not os.path.exists(orc._LIB_PATH):
        orc.build()
    affine = cfg["model"] == "affine"
    rp = orc.reg_params(symmetric=1, affine=int(affine), nw=cfg["nhood_width"],
                        ng=cfg["nhood_gap"], pyr_num=cfg["pyramid_levels"],
                        iterations=cfg["iterations_per_level"], nthreads=threads)
    t = time.perf_counter()
    ref_p, ref_c, ref_st = orc.register(pair.fixed, pair.moving, rp)
    dt = time.perf_counter() - t
    M1, o1 = transforms.matrix_offset(params, center, affine)
    M2, o2 = transforms.matrix_offset(ref_p, ref_c, affine)
    return {
        "transform_vs_oracle_mm": transforms.max_corner_distance(M1, o1, M2, o2, pair.fixed),
        "rotation_vs_oracle_deg": transforms.rotation_angle_deg(M1, M2),
        "same_iterations_and_points": bool(list(iterations) == list(ref_st["iterations"])
                                           and list(points) == list(ref_st["points"])),
        "oracle_seconds": dt,
        "oracle_threads": threads,
    }


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


def _make_batch_pair(job):
    from mirorr_amd import phantom
    i, dims, spacing, model, n_ell, seed = job
    return phantom.make_pair(dims=dims, spacing=spacing, n_ellipsoids=n_ell, seed=seed,
                             model=model)


def _make_ctmr_pair(job):
    from mirorr_amd import phantom
    dims, seed, n_ell = job
    return phantom.make_ct_mr_pair(dims=dims, spacing=(1.0, 1.0, 1.0), n_ellipsoids=n_ell,
                                   seed=seed)


def _pool_map(fn, jobs):
    """Generates inputs in worker processes (before the GPU is touched)."""
    import multiprocessing as mp
    n = max(1, min(len(jobs), 16, len(os.sched_getaffinity(0))))
    with mp.get_context("fork").Pool(n) as pool:
        return pool.map(fn, jobs)


def _params(model, c):
    import ctypes as C
    from mirorr_amd import capi
    p = capi.RegParams()
    capi.lib().mirorr_reg_default_params(C.byref(p))
    p.symmetric = 1
    p.affine = 1 if model == "affine" else 0
    p.block_width, p.block_gap = 4, 1
    p.nhood_width, p.nhood_gap = c["nhood_width"], c["nhood_gap"]
    p.pyr_num, p.pyr_min_size, p.pyr_end = c["pyramid_levels"], 32, 0
    p.iterations = c.get("iterations_per_level", 0)
    if c.get("variance_select"):
        p.variance_select, p.portion_kept = 1, c["portion_kept"]
    return p


def _plain_run(ctx, pair, p):
    """The plain path: upload both volumes, register, on one context."""
    import ctypes as C
    from mirorr_amd import capi
    from mirorr_amd.registration import DeviceVolume
    fd, md = DeviceVolume(pair.fixed, ctx), DeviceVolume(pair.moving, ctx)
    out, oc, st = np.zeros(12), np.zeros(3), capi.RegStats()
    capi.check(capi.lib().mirorr_register(ctx.handle, C.byref(p), fd.handle, md.handle, None, None,
                                          capi.ptr(out), capi.ptr(oc), C.byref(st)), ctx.handle)
    fd.close()
    md.close()
    return out[: 12 if p.affine else 6], oc, st.voxel_evals


def _served_line(cfg, args, extra):
    out = {
        "metric": "NCC voxel-evals/s + registrations/s (p50/p99 latency) vs plain GPU, 1 GPU",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic", "config": cfg, "roofline": None, "cpu_baseline": None,
    }
    out.update(extra)
    return out


def batch_jobs(cfg):
    """The pairs (generator jobs) and registration parameters of a batch."""
    rng = np.random.default_rng(cfg["seed"])
    jobs, params = [], []
    for i in range(cfg["pairs"]):
        if "dims" in cfg:  # pairs of one size and model
            model, dims, sp = "affine", tuple(cfg["dims"]), (cfg["spacing_mm"],) * 3
        else:
            lo, hi, st = cfg["axis_lengths"]
            model = "rigid" if i % 2 == 0 else "affine"
            dims = tuple(int(v) for v in rng.choice(np.arange(lo, hi + 1, st), 3))
            sp = tuple(float(v) for v in rng.uniform(cfg["spacing_mm"][0], cfg["spacing_mm"][1],
                                                     3))
        jobs.append((i, dims, sp, model, cfg[model]["ellipsoids"], cfg["seed"] * 1000 + i))
        params.append(_params(model, cfg[model]))
    return jobs, params


def measure_lockstep(ctx, pairs, params, plain, steps):
    """mirorr_register_batch over the pairs (one batch per parameter set),
    against mirorr_register one pair at a time, both on volumes already in
    HBM (uploads excluded on both sides)."""
    from mirorr_amd.registration import DeviceVolume, MirorrPyramid, register_batch
    F = [DeviceVolume(pr.fixed, ctx) for pr in pairs]
    M = [DeviceVolume(pr.moving, ctx) for pr in pairs]
    regs = []
    for p in params:
        r = MirorrPyramid(ctx)
        r.SetParameters(p)
        regs.append(r)
    groups = {}
    for i, p in enumerate(params):
        key = (p.affine, p.nhood_width, p.nhood_gap, p.pyr_num, p.iterations, p.variance_select,
               p.portion_kept)
        groups.setdefault(key, []).append(i)

    def one_by_one():
        return [r.RunMirorr(fixed_dev=f, moving_dev=m) for r, f, m in zip(regs, F, M)]

    def lockstep():
        res = [None] * len(pairs)
        for g in groups.values():
            for i, x in zip(g, register_batch(regs[g[0]], [F[i] for i in g], [M[i] for i in g])):
                res[i] = x
        return res

    one_by_one()
    lockstep()
    ctx.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        one_by_one()
    t_one = time.perf_counter() - t
    t = time.perf_counter()
    for _ in range(steps):
        res = lockstep()
    t_lock = time.perf_counter() - t
    same = all(np.array_equal(a[0], q) and np.array_equal(a[1], c)
               for a, (q, c, _) in zip(res, plain))
    for v in F + M:
        v.close()
    n = len(pairs) * steps
    return {"registrations_per_s": n / t_lock, "one_by_one_registrations_per_s": n / t_one,
            "vs_one_by_one": t_one / t_lock, "batches": len(groups),
            "results_identical_to_plain": same,
            "note": "mirorr_register_batch per parameter set against mirorr_register one pair "
                    "at a time, volumes already in HBM"}


def measure_batch(pairs, params, cfg, steps, warmup, storages=("f32",)):
    """One step = every pair registered once: through the plain path (one
    context, one pair at a time, float32) and through the server (all
    submitted at once) in each storage. Uploads are included on both sides."""
    from mirorr_amd.blockmatcher import default_context
    from mirorr_amd.server import RegistrationServer
    ctx = default_context()
    print("bench: plain path", file=sys.stderr, flush=True)
    for _ in range(warmup):
        for pr, p in zip(pairs, params):
            _plain_run(ctx, pr, p)
    t = time.perf_counter()
    plain = []
    for _ in range(steps):
        plain = [_plain_run(ctx, pr, p) for pr, p in zip(pairs, params)]
    t_plain = time.perf_counter() - t
    evals = sum(e for _, _, e in plain)
    print(f"bench: plain {t_plain / steps * 1e3:.0f} ms per batch", file=sys.stderr, flush=True)
    n = len(pairs) * steps
    out = {"plain_registrations_per_s": n / t_plain, "voxel_evals_per_step": evals,
           "plain_ms_per_step": t_plain / steps * 1e3}
    out["lockstep"] = measure_lockstep(ctx, pairs, params, plain, steps)
    for storage in storages:
        with RegistrationServer(slots=cfg["slots"]) as srv:
            def batch():
                ts = [srv.submit(pr.fixed, pr.moving, p, borrow=True, storage=storage)
                      for pr, p in zip(pairs, params)]
                return [srv.poll(x) for x in ts]
            for _ in range(warmup):
                batch()
            lat, res = [], []
            t = time.perf_counter()
            for k in range(steps):
                res = batch()
                lat += [r["latency"] for r in res]
                print(f"bench: served {storage} step {k + 1}/{steps}", file=sys.stderr,
                      flush=True)
            t_served = time.perf_counter() - t
            info = srv.info()
        bad = [r["error"] for r in res if r["status"] != 0]
        if bad:
            raise RuntimeError(f"served registrations failed: {bad[:3]}")
        # binary16 storage rounds the images, so only float32 results are
        # expected to equal the plain path's bit for bit
        same = (all(np.array_equal(r["params"], q) and np.array_equal(r["center"], c)
                    for r, (q, c, _) in zip(res, plain)) if storage == "f32" else None)
        lat_ms = np.array(lat) * 1e3
        out[storage] = {
            "registrations_per_s": n / t_served,
            "vs_plain": t_plain / t_served,
            "ms_per_step": t_served / steps * 1e3,
            "value": evals * steps / t_served,
            "latency_ms_p50": float(np.percentile(lat_ms, 50)),
            "latency_ms_p99": float(np.percentile(lat_ms, 99)),
            "results_identical_to_plain": same,
            "max_running": info["max_running"],
            "hbm_peak_bytes": info["hbm_peak"],
        }
    return out


def run_batch(args, cfg):
    """One step = every pair of the batch registered once, through the server
    (all submitted at once) and, for comparison, through the plain path."""
    jobs, params = batch_jobs(cfg)
    print(f"bench: generating {len(jobs)} pairs", file=sys.stderr, flush=True)
    pairs = _pool_map(_make_batch_pair, jobs)
    m = measure_batch(pairs, params, cfg, args.steps, args.warmup, ("f32", "f16"))
    f32 = m["f32"]
    return _served_line(cfg, args, {
        "value": f32["value"],
        "unit": "voxel-evals/s",
        "ms_per_step": f32["ms_per_step"],
        "registrations_per_s": f32["registrations_per_s"],
        "plain_registrations_per_s": m["plain_registrations_per_s"],
        "vs_plain": f32["vs_plain"],
        "latency_ms_p50": f32["latency_ms_p50"],
        "latency_ms_p99": f32["latency_ms_p99"],
        "latency_note": "all pairs of a step are submitted at once",
        "voxel_evals_per_step": m["voxel_evals_per_step"],
        "results_identical_to_plain": f32["results_identical_to_plain"],
        "max_running": f32["max_running"],
        "hbm_peak_bytes": f32["hbm_peak_bytes"],
        "f16_storage": m["f16"],
        "lockstep": m["lockstep"],
    })


def run_serve(args, cfg):
    """Poisson arrivals at each rate, served with cfg["slots"] workers; at the
    highest rate the same stream also through one worker (one request at a
    time, the plain path's service order)."""
    from mirorr_amd.server import RegistrationServer
    rng = np.random.default_rng(cfg["seed"])
    jobs = ([(tuple(cfg["large_dims"]), cfg["seed"] + k, cfg["ellipsoids"])
             for k in range(cfg["pool_large"])] +
            [(tuple(cfg["small_dims"]), cfg["seed"] + 100 + k, cfg["ellipsoids"])
             for k in range(cfg["pool_small"])])
    print(f"bench: generating {len(jobs)} CT/MR pairs", file=sys.stderr, flush=True)
    pool = _pool_map(_make_ctmr_pair, jobs)
    large, small = pool[: cfg["pool_large"]], pool[cfg["pool_large"]:]
    p = _params("rigid", cfg)

    def stream(rate, n):
        gaps = rng.exponential(1.0 / rate, n)
        picks = [(large if rng.random() < cfg["large_fraction"] else small) for _ in range(n)]
        return np.cumsum(gaps) - gaps[0], [c[rng.integers(len(c))] for c in picks]

    def serve(srv, arrivals, reqs):
        t0 = time.perf_counter()
        ts = []
        for k, (a, pr) in enumerate(zip(arrivals, reqs)):
            d = t0 + a - time.perf_counter()
            if d > 0:
                time.sleep(d)
            ts.append(srv.submit(pr.fixed, pr.moving, p, borrow=True))
            if (k + 1) % 10 == 0:
                print(f"bench: submitted {k + 1}/{len(reqs)}", file=sys.stderr, flush=True)
        res = [srv.poll(x) for x in ts]
        bad = [r["error"] for r in res if r["status"] != 0]
        if bad:
            raise RuntimeError(f"served registrations failed: {bad[:3]}")
        lat = np.array([r["latency"] for r in res]) * 1e3
        span = max(r["t_end"] for r in res) - min(r["t_submit"] for r in res)
        return {"latency_ms_p50": float(np.percentile(lat, 50)),
                "latency_ms_p99": float(np.percentile(lat, 99)),
                "registrations_per_s": len(res) / span}, res

    per_rate = {}
    with RegistrationServer(slots=cfg["slots"]) as srv:
        for pr in large[:1] + small[:1]:  # warm the workers' workspaces
            srv.poll(srv.submit(pr.fixed, pr.moving, p, borrow=True))
        streams = {}
        for rate in cfg["rates_per_s"]:
            streams[rate] = stream(rate, cfg["requests_per_rate"])
            per_rate[str(rate)], res = serve(srv, *streams[rate])
            print(f"bench: rate {rate}/s {per_rate[str(rate)]}", file=sys.stderr, flush=True)
        evals = sum(r["stats"]["voxel_evals"] for r in res)
        busy = sum(r["t_end"] - r["t_start"] for r in res)
        info = srv.info()
    top = max(cfg["rates_per_s"])
    with RegistrationServer(slots=1) as one:
        one.poll(one.submit(large[0].fixed, large[0].moving, p, borrow=True))
        plain_top, _ = serve(one, *streams[top])
    print(f"bench: one at a time at {top}/s {plain_top}", file=sys.stderr, flush=True)
    at = per_rate[str(top)]
    return _served_line(cfg, args, {
        "value": evals / busy,
        "unit": "voxel-evals/s",
        "value_note": "voxel-evals per second of request service time at the highest rate",
        "ms_per_step": busy / len(res) * 1e3,
        "registrations_per_s": at["registrations_per_s"],
        "latency_ms_p50": at["latency_ms_p50"],
        "latency_ms_p99": at["latency_ms_p99"],
        "per_rate": per_rate,
        "one_at_a_time_at_highest_rate": plain_top,
        "vs_plain_p99": plain_top["latency_ms_p99"] / at["latency_ms_p99"],
        "max_running": info["max_running"],
        "hbm_peak_bytes": info["hbm_peak"],
        "data": "synthetic; requests drawn from a pool of distinct CT/MR-like pairs",
    })


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-served", action="store_true",
                    help="skip the served batch (16 pairs of the workload through the server)")
    ap.add_argument("--workload", default="pair256_affine_symmetric",
                    choices=sorted(WORKLOADS) + sorted(SERVED))
    ap.add_argument("--slots", type=int, default=0,
                    help="served workloads: worker slots (0: the workload's own)")
    ap.add_argument("--requests-per-rate", type=int, default=0,
                    help="serve_poisson_ctmr: requests per arrival rate (0: the workload's own)")
    ap.add_argument("--iteration-graphs", type=int, choices=(0, 1), default=None,
                    help="replay each level's outer iterations from captured HIP graphs "
                         "(default: the library's default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported "
              "(independent registrations shard across GPUs with no communication).",
              file=sys.stderr)
        sys.exit(2)
    if args.workload in SERVED:
        cfg = dict(SERVED[args.workload])
        if args.slots > 0:
            cfg["slots"] = args.slots
        if args.requests_per_rate > 0 and "requests_per_rate" in cfg:
            cfg["requests_per_rate"] = args.requests_per_rate
        run = run_batch if args.workload.startswith("batch") else run_serve
        print(json.dumps(run(args, cfg)), flush=True)
        return

    from mirorr_amd.blockmatcher import default_context
    from mirorr_amd.registration import DeviceVolume, MirorrPyramid

    cfg = workload(args.workload)
    served = None
    if not args.no_served and args.workload == "pair256_affine_symmetric":
        # generated in worker processes before the GPU is touched
        bcfg = dict(SERVED["batch16_pair256_affine"])
        bjobs, bparams = batch_jobs(bcfg)
        print(f"bench: generating {len(bjobs)} pairs for the served batch", file=sys.stderr,
              flush=True)
        served = (bcfg, _pool_map(_make_batch_pair, bjobs), bparams)
    pair = make_inputs(cfg)
    ctx = default_context()
    # host -> HBM, outside the timed region: the first upload of the process
    # (it also allocates the context's pinned staging) and a second one
    t_up = time.perf_counter()
    fixed = DeviceVolume(pair.fixed, ctx)
    moving = DeviceVolume(pair.moving, ctx)
    upload_first_ms = (time.perf_counter() - t_up) * 1e3
    fixed.close()
    moving.close()
    t_up = time.perf_counter()
    fixed = DeviceVolume(pair.fixed, ctx)
    moving = DeviceVolume(pair.moving, ctx)
    upload_ms = (time.perf_counter() - t_up) * 1e3
    reg = MirorrPyramid(ctx)
    configure(reg, cfg)
    if args.iteration_graphs is not None:
        ctx.iteration_graphs = bool(args.iteration_graphs)
    print(f"bench: warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        reg.RunMirorr(fixed_dev=fixed, moving_dev=moving)
    ctx.synchronize()
    # The timed steps run as a caller runs them: the two block-matching
    # passes of an iteration overlap on two streams. The roofline is taken on
    # further, untimed registrations with HIP events around the fast search
    # kernel's full-resolution launches; while kernels are timed the library
    # runs the passes one after the other, so each launch's time is its own.
    # The masked kernel's share comes from one more registration.
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
    # The same steps with iteration graphs switched the other way (one
    # untimed registration first, which captures them when they are on).
    graphs_on = ctx.iteration_graphs
    ctx.iteration_graphs = not graphs_on
    reg.RunMirorr(fixed_dev=fixed, moving_dev=moving)
    ctx.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        reg.RunMirorr(fixed_dev=fixed, moving_dev=moving)
    ctx.synchronize()
    other_ms = (time.perf_counter() - t1) / args.steps * 1e3
    ctx.iteration_graphs = graphs_on
    ctx.set_profiling(ctx.PROF_FAST | ctx.PROF_FINEST_ONLY)
    ctx.profile_reset()
    for _ in range(ROOFLINE_REGISTRATIONS):
        reg.RunMirorr(fixed_dev=fixed, moving_dev=moving)
    n_fast, ms_fast, work_fast = ctx.profile("ncc_tile_fast")
    ctx.set_profiling(ctx.PROF_FAST | ctx.PROF_MASKED | ctx.PROF_FINEST_ONLY)
    ctx.profile_reset()
    reg.RunMirorr(fixed_dev=fixed, moving_dev=moving)
    _, _, work_fast1 = ctx.profile("ncc_tile_fast")
    _, ms_mask, work_mask = ctx.profile("ncc_tile_masked")
    ctx.set_profiling(False)

    from mirorr_amd import capi, transforms
    M, off = transforms.matrix_offset(params, center, cfg["model"] == "affine")
    err_mm = transforms.max_corner_distance(M, off, pair.matrix, pair.translation, pair.fixed)
    iterations, points = list(reg.stats["iterations"]), list(reg.stats["points"])

    avg_ms = ms_fast / max(n_fast, 1)
    achieved = FLOP_PER_VOXEL_EVAL * (work_fast / max(n_fast, 1)) / (avg_ms * 1e-3) / 1e12
    traffic = measured_traffic() if args.workload == "pair256_affine_symmetric" else None
    lds_gbs = (LDS_BYTES_PER_VOXEL_EVAL_NW9 * (work_fast / max(n_fast, 1)) / (avg_ms * 1e-3) / 1e9
               if cfg["nhood_width"] == 9 else None)
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
        "value_note": "voxel-evals by the formula count of every block's candidate windows "
                      "x 64, including blocks whose base mask is off everywhere (not searched)",
        "transform_error_mm": err_mm,
        "upload_ms": upload_ms,
        "upload_first_ms": upload_first_ms,
        "upload_note": "both volumes (2 x 32 MiB float32) through the context's pinned "
                       "staging chunks; the first upload also allocates them",
        "iteration_graphs": {
            "on": graphs_on,
            "note": "each level's outer iterations replayed from a captured HIP graph",
            "ms_per_step_other_setting": other_ms},
        "roofline": {
            "bound": "valu",
            "compute": "fp32 vector ALU FMAs (the search uses no MFMA); packed FP32 is "
                       "not used, so the issue-rate ceiling of the kernel's FMAs is half the "
                       "published peak",
            "kernel": "ncc_tile_fast",
            "launch_scope": "full-resolution level",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            # the committed traffic summary covers the default workload only
            "traffic": traffic,
            "hbm": None if traffic is None else {
                "achieved": traffic / (avg_ms * 1e-3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": traffic / (avg_ms * 1e-3) / 1e9 / PEAK_HBM_GBS},
            "lds": None if lds_gbs is None else {
                "achieved": lds_gbs, "peak": PEAK_LDS_GBS, "unit": "GB/s",
                "frac": lds_gbs / PEAK_LDS_GBS,
                "bytes_per_voxel_eval": LDS_BYTES_PER_VOXEL_EVAL_NW9,
                "note": "inner-loop window and block reads of the tile kernel"},
            "launches": n_fast,
            "launch_note": f"{ROOFLINE_REGISTRATIONS} untimed registrations after the timed "
                           "steps, block-matching passes one after the other",
            "avg_launch_ms": avg_ms,
            "flop_per_voxel_eval": FLOP_PER_VOXEL_EVAL,
            "masked_kernel_ms_full_resolution": ms_mask,
            "masked_kernel_share_of_evals_full_resolution": work_mask / max(work_fast1 + work_mask, 1.0),
        },
        "cpu_baseline": None,
        # the NCC search's tile shape per level grid, picked by timing each
        # shape on the first search of that grid (the warm-up registration)
        "tile_autotune": capi.tile_autotune_report().splitlines(),
    }
    if served is not None:
        bcfg, bpairs, bparams = served
        m = measure_batch(bpairs, bparams, bcfg, 2, 1, ("f32", "f16"))
        out["served"] = {
            "workload": bcfg["workload"], "pairs": bcfg["pairs"], "slots": bcfg["slots"],
            "note": "16 distinct 256x256x128 affine pairs submitted at once, uploads included; "
                    "plain = one pair at a time on one context, float32",
            "plain_registrations_per_s": m["plain_registrations_per_s"],
            "f32": m["f32"], "f16": m["f16"], "lockstep": m["lockstep"],
        }
        del bpairs
    if not args.no_cpu_baseline:
        # The host cores a one-GPU job may use (16 on the MI355X boxes, whatever
        # nproc reports), then one core on a smaller sample.
        threads = min(16, len(os.sched_getaffinity(0)))
        out["cpu_baseline"] = cpu_baseline(pair, cfg, threads)
        out["cpu_baseline"]["cpu_model"] = cpu_model()
        single = cpu_baseline(pair, cfg, 1, max_evals=4e9)
        out["cpu_baseline"]["single_core"] = {k: single[k] for k in ("value", "cores", "sample")}
        print("bench: oracle registration for the transform check", file=sys.stderr, flush=True)
        out["oracle_check"] = oracle_transform_check(pair, cfg, params, center, iterations,
                                                     points, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
