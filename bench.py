"""Benchmark: plain hair path tracing on one MI355X at 1024x1024 over 3.2M
hair segments (the c1_curly scene), reported as paths/s.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one sample pass over the full frame: 1024 x 1024 camera paths, each
traced to termination with its next-event-estimation and MIS shadow rays
(max depth 40, Russian roulette). The scene, BVH, environment tables and path
state are resident in HBM before the timed region. One JSON line is printed.

Roofline: the dominant kernel is the closest-hit BVH traversal (extend).
Algorithmic bytes per launch = nodes visited x 64 B + segments tested x 32 B
+ 48 B per ray (32 B ray in, 16 B hit out), counted by the same traversal
compiled with counters; divided by the extend kernel's mean launch time from
HIP events on the context's stream.

CPU baseline: the C++ oracle (FP32 twin of the reference integrator, OpenMP)
on a bounded crop of the same scene, run on this machine's host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "paths/s @1024² 3.2M segs; HairMSNN β=1 vs plain PT time-to-equal-RMSE"
HBM_PEAK_GBS = 8000.0
SCENE = ROOT / "hairmsnn_amd" / "scenes" / "c1_curly.json"
NODE_BYTES, SEG_BYTES, RAY_BYTES = 64, 32, 48


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(sc, budget_s: float = 15.0) -> dict:
    from tests import oracle_lib as ol
    orc = ol.load_oracle()
    t0 = time.perf_counter()
    osc = ol.OracleScene(sc.info, sc.segments(), sc.env())
    build_s = time.perf_counter() - t0
    threads = int(orc.orc_max_threads())
    W, H = sc.width, sc.height
    # calibrate on a small crop, then size the sample to the time budget
    cw = 32
    x0, y0 = W // 2 - cw // 2, H // 2 - cw // 2
    t0 = time.perf_counter()
    osc.render(x0, y0, cw, cw, spp=1, accum_id0=0, prec=1, threads=threads)
    dt = max(time.perf_counter() - t0, 1e-3)
    rate = cw * cw / dt
    side = int(np.clip(np.sqrt(rate * budget_s), 32, 256))
    x0, y0 = W // 2 - side // 2, H // 2 - side // 2
    t0 = time.perf_counter()
    osc.render(x0, y0, side, side, spp=1, accum_id0=1, prec=1, threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": side * side / dt, "unit": "paths/s", "cores": threads, "kind": "port",
        "sample": f"centre {side}x{side} crop of the same 1024x1024 frame, 1 spp, FP32 oracle "
                  f"(C++ restatement of pathTrace, CPU BVH built in {build_s:.1f} s, not timed), "
                  f"{dt:.1f} s of CPU work",
    }


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--scene", default=str(SCENE))
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the path tracer shards by image tiles; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import hairmsnn_amd as hm
    n, name = hm.device_info()
    if n < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    log(f"device: {name}")
    t0 = time.perf_counter()
    sc = hm.Scene(args.scene)
    log(f"scene ready in {time.perf_counter() - t0:.1f} s: {int(sc.info['num_segments'])} segments, "
        f"BVH depth {int(sc.info['bvh_depth'])}")
    W, H = sc.width, sc.height
    paths_per_step = W * H

    # counting pass (not timed): algorithmic traversal bytes
    stats = sc.trace_stats(1)
    log(f"trace stats (1 pass): {stats}")
    # warmup
    sc.pt_reset()
    sc.pt_passes(args.warmup)
    # timed region: K passes, synchronised on both sides
    sc.synchronize()
    t0 = time.perf_counter()
    sc.pt_passes(args.steps, sync=False)
    sc.synchronize()
    wall = time.perf_counter() - t0
    value = paths_per_step * args.steps / wall
    log(f"{args.steps} passes in {wall:.3f} s -> {value / 1e6:.2f} M paths/s")

    # per-kernel timing with HIP events on the context's stream
    kt = sc.time_passes(max(2, min(8, args.steps)))
    passes_t = max(2, min(8, args.steps))
    launches = max(kt["extend_launches"], 1)
    extend_ms = kt["extend_ms"] / launches
    bytes_pass = (stats["closest_nodes"] * NODE_BYTES + stats["closest_segs"] * SEG_BYTES
                  + stats["closest_rays"] * RAY_BYTES)
    launches_pass = launches / passes_t
    bytes_launch = bytes_pass / launches_pass
    achieved = bytes_launch / (extend_ms * 1e-3) / 1e9
    any_bytes = (stats["any_nodes"] * NODE_BYTES + stats["any_segs"] * SEG_BYTES + stats["any_rays"] * 36)
    shadow_gbs = any_bytes / (kt["shadow_ms"] / passes_t * 1e-3) / 1e9
    log(f"extend: {extend_ms:.3f} ms/launch, {bytes_launch / 1e6:.1f} MB/launch -> {achieved:.0f} GB/s; "
        f"shadow {shadow_gbs:.0f} GB/s; shade {kt['shade_ms'] / passes_t:.2f} ms/pass")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "c1_curly: plain hair path tracing, 1024x1024 frame, 1 sample per pixel per step "
                        "(64 steps = the 64-spp render), 50000 strands x 64 linear segments (3.2M) on a unit "
                        "head sphere, procedural 2048x1024 log-normal env + 3 hot spots + 1 directional "
                        "light, depth 40, MIS, LCG",
            "width": W, "height": H, "paths_per_step": paths_per_step,
            "segments": int(sc.info["num_segments"]), "bvh_depth": int(sc.info["bvh_depth"]),
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel": "k_extend (closest-hit LBVH traversal)",
            "bytes_per_launch": bytes_launch, "ms_per_launch": extend_ms,
            "nodes_per_ray": stats["closest_nodes"] / max(stats["closest_rays"], 1),
            "segs_per_ray": stats["closest_segs"] / max(stats["closest_rays"], 1),
        },
        "kernels_ms_per_pass": {
            "extend": kt["extend_ms"] / passes_t, "shade": kt["shade_ms"] / passes_t,
            "shadow": kt["shadow_ms"] / passes_t, "total": kt["total_ms"] / passes_t,
        },
        "rays_per_pass": {"closest": stats["closest_rays"], "any": stats["any_rays"]},
    }
    if not args.no_cpu:
        try:
            line["cpu_baseline"] = cpu_baseline(sc)
        except Exception as e:  # the GPU figure stands on its own
            log(f"cpu baseline failed: {e}")
            line["cpu_baseline"] = None
    sc.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
