"""Benchmark: plain hair path tracing on one MI355X at 1024x1024 over 3.2M
hair segments (the c1_curly scene), reported as paths/s.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one sample pass over the full frame: 1024 x 1024 camera paths, each
traced to termination with its next-event-estimation and MIS shadow rays
(max depth 40, Russian roulette). The scene, BVH, environment tables and path
state are resident in HBM before the timed region. One JSON line is printed.

Roofline: every pass kernel is timed alone with HIP events on its stream
(extend: closest-hit traversal; shade: hit setup, light and BSDF samples,
compaction; shadow: any-hit traversal), and the line's `roofline` is the
one with the most time per pass. Algorithmic bytes: traversal = 4-wide nodes
visited x 128 B + segments tested x 32 B + ray in / hit out (48 B closest,
36 B any-hit), counted by the same traversal compiled with counters; shade =
per vertex 56 B of path state read, per continuing vertex 60 B more read
(hit, RNG, throughput, segment) and 92 B moved into the next wavefront, per
finished path 36 B, per shadow ray 52 B written. HBM traffic per launch comes
from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised in profiles/.

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
NODE_BYTES, SEG_BYTES, RAY_BYTES = 128, 32, 48  # 4-wide node, segment record, ray in + hit out
ANY_RAY_BYTES = 36  # shadow ray in (32 B) + visibility out
# shade: path state read per vertex, more per continuing vertex, moved per
# continuing vertex, per finished path, per shadow ray
SHADE_VERTEX_BYTES, SHADE_HIT_BYTES, SHADE_MOVE_BYTES, SHADE_DONE_BYTES, SHADE_SHADOW_BYTES = 56, 60, 92, 36, 52
# HBM bytes per launch of each pass kernel from the rocprofv3 FETCH_SIZE /
# WRITE_SIZE runs (copy-calibrated); counters cannot be read from inside this
# process
TRAFFIC_JSON = ROOT / "profiles" / "pt_c1_1024_hbm_traffic.json"
PASS_KERNELS = {"extend": "k_extend<false, true, false>", "shade": "k_shade_dense", "shadow": "k_shadow<false, false>"}
# FP32 VALU operations per launch (rocprofv3 SQ_INSTS_VALU_FLOPS_FP32 x 64
# lanes) of the pass kernels and of the BSDF streams, for the VALU FLOP rates
FLOPS_JSON = ROOT / "profiles" / "pt_c1_1024_flops.json"
KERNEL_FLOPS_JSON = ROOT / "profiles" / "bsdf_mlp_flops.json"
VALU_FP32_PEAK_TFLOPS = 157.3


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(sc, side: int = 128, spp: int = 64) -> dict:
    """The FP32 oracle (C++ restatement of pathTrace with a CPU BVH, OpenMP over
    pixel rows) on the centre side x side crop of the same frame at spp samples
    per pixel, on all host cores."""
    from tests import oracle_lib as ol
    orc = ol.load_oracle()
    t0 = time.perf_counter()
    osc = ol.OracleScene(sc.info, sc.segments(), sc.env())
    build_s = time.perf_counter() - t0
    threads = int(orc.orc_max_threads())
    W, H = sc.width, sc.height
    x0, y0 = W // 2 - side // 2, H // 2 - side // 2
    t0 = time.perf_counter()
    osc.render(x0, y0, side, side, spp=spp, accum_id0=1, prec=1, threads=threads)
    dt = time.perf_counter() - t0
    out = {
        "value": side * side * spp / dt, "unit": "paths/s", "cores": threads, "kind": "port",
        "sample": f"centre {side}x{side} crop of the same 1024x1024 frame, {spp} spp, FP32 oracle "
                  f"(C++ restatement of pathTrace, CPU BVH built in {build_s:.1f} s, not timed), "
                  f"{dt:.1f} s of CPU work",
    }
    out["c0_full"] = cpu_c0_full(threads)
    out["mlp_fwd_bwd"] = cpu_mlp(1 << 16)
    return out


def cpu_c0_full(threads: int, spp: int = 16) -> dict:
    """The FP32 oracle on the whole C0 frame (the hair-ball scene at 256x256,
    spp passes), on the same host cores."""
    import hairmsnn_amd as hm
    from tests import oracle_lib as ol
    cfg = json.loads((ROOT / "hairmsnn_amd" / "scenes" / "c0_hairball.json").read_text())
    with hm.Scene(json_text=cfg) as sc:
        osc = ol.OracleScene(sc.info, sc.segments(), sc.env() if sc.info["has_env"] else None, mesh=sc.mesh())
        W, H = sc.width, sc.height
    t0 = time.perf_counter()
    osc.render(0, 0, W, H, spp=spp, accum_id0=0, prec=1, threads=threads)
    dt = time.perf_counter() - t0
    return {"paths_per_s": W * H * spp / dt, "paths": W * H * spp, "s": dt,
            "sample": f"c0_hairball {W}x{H} x {spp} spp, FP32 oracle, {threads} threads"}


def cpu_mlp(n: int) -> dict:
    """The oracle network's forward + backward (loss and weight gradient) on n
    seeded records in FP32 and FP64 (124,032 FLOP per record)."""
    from tests import oracle_lib as ol
    rng = np.random.default_rng(8)
    rec = np.zeros((n, 12), np.float32)
    rec[:, 0:3] = rng.uniform(-0.5, 0.5, (n, 3))
    for a in (3, 6):
        v = rng.normal(size=(n, 3))
        rec[:, a:a + 3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    rec[:, 9:12] = rng.uniform(0.01, 0.5, (n, 3))
    tgt = rng.uniform(0.1, 2.0, (n, 3)).astype(np.float32)
    w = rng.uniform(-0.2, 0.2, ol.MLP_WEIGHTS).astype(np.float32)
    out = {"records": n}
    for name, prec in (("fp32", 1), ("fp64", 0)):
        t0 = time.perf_counter()
        ol.mlp_loss_grad(w, rec, tgt, loss=0, prec=prec, exp_floor=ol.MLP_EXP_GRAD_FLOOR)
        dt = time.perf_counter() - t0
        out[name] = {"s": dt, "records_per_s": n / dt, "GFLOP/s": 124032.0 * n / dt / 1e9}
    return out


def mlp_rates(hm, n: int = 1 << 20, reps: int = 10) -> dict:
    """Inference TFLOP/s of the radiance-cache network (41,344 FLOP per query)
    in FP16 and FP8 on n seeded queries, device-resident, wall clock over reps
    calls after one warm call (the FP8 calibration pass runs in the warm call)."""
    import ctypes as C
    lib = hm._lib.load()
    rng = np.random.default_rng(4)
    rec = np.zeros((n, 12), np.float32)
    rec[:, 0:3] = rng.uniform(-0.5, 0.5, (n, 3))
    for a in (3, 6):
        v = rng.normal(size=(n, 3))
        rec[:, a:a + 3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    rec[:, 9:12] = rng.uniform(0.01, 0.5, (n, 3))
    din, dout = hm.DeviceBuffer.from_numpy(rec), hm.DeviceBuffer(12 * n)
    out = {}
    for name, prec in (("fp16", 0), ("fp8", 2)):
        h = C.c_void_p()
        hm._lib.check(lib.hmsnn_mlp_create(prec, 1, C.byref(h)))
        hm._lib.check(lib.hmsnn_mlp_forward_dev(h, din.ptr, n, dout.ptr, None))
        hm._lib.check(lib.hmsnn_dev_synchronize())
        t0 = time.perf_counter()
        for _ in range(reps):
            hm._lib.check(lib.hmsnn_mlp_forward_dev(h, din.ptr, n, dout.ptr, None))
        hm._lib.check(lib.hmsnn_dev_synchronize())
        dt = (time.perf_counter() - t0) / reps
        lib.hmsnn_mlp_destroy(h)
        peak = 2500.0 if name == "fp16" else 5000.0
        tf = 41344.0 * n / dt / 1e12
        out[name] = {"queries": n, "ms": dt * 1e3, "TFLOP/s": tf, "peak_TFLOP/s": peak, "frac": tf / peak}
    return out


def pt_reference(sc, ref_spp: int, frames: int) -> dict:
    """A ref_spp path-traced reference on independent random streams, and the
    plain tracer's MSE(k) against it for k = 1..frames passes, fitted as
    a / k + b with b = a / ref_spp (the reference's own noise)."""
    sc.set_seed_offset(1 << 22)
    sc.pt_reset()
    t0 = time.perf_counter()
    done = 0
    while done < ref_spp:
        k = min(256, ref_spp - done)
        sc.pt_passes(k)
        done += k
        log(f"reference: {done}/{ref_spp} passes ({time.perf_counter() - t0:.1f} s)")
    ref = sc.image(0)[..., :3].astype(np.float64)
    sc.set_seed_offset(0)
    sc.pt_reset()
    ks, pt_mse = [], []
    for k in range(1, frames + 1):
        sc.pt_passes(1)
        if k & (k - 1) == 0 or k == frames:
            ks.append(k)
            pt_mse.append(float(np.mean((sc.image(0)[..., :3].astype(np.float64) - ref) ** 2)))
    # the reference is the same unbiased estimator at ref_spp passes on
    # independent streams, so MSE_PT(k) = a (1 / k + 1 / ref_spp): one
    # parameter, fitted in relative terms so that no single k dominates
    x, m_a = 1.0 / np.array(ks, float) + 1.0 / ref_spp, np.array(pt_mse)
    a = float(np.sum(x / m_a) / np.sum((x / m_a) ** 2))
    return {"ref": ref, "spp": ref_spp, "a": a, "b": a / ref_spp, "ks": ks, "mse": pt_mse}


def msnn_vs_pt(sc, pt_ms_per_pass: float, R: dict, frames: int, beta: int = 1) -> dict:
    """HairMSNN at BETA vs plain path tracing: the time the plain tracer needs
    to reach the MSE HairMSNN has after its warm-up and `frames` frames
    (MSE_PT(k) = a / k + b solved for k, times the timed per-pass cost;
    b, the reference's own noise, is taken off both sides)
    over HairMSNN's wall time (warm-up training + frames). The
    scene-sample generation inside hmsnn_msnn_begin is one-time scene setup,
    like the BVH build the PT time excludes too, and is reported, not counted."""
    ref, a, b = R["ref"], R["a"], R["b"]

    def mse(img):
        return float(np.mean((img[..., :3].astype(np.float64) - ref) ** 2))

    sc.synchronize()
    sc.msnn_begin(beta)
    st0 = sc.msnn_frames(0)
    t_setup, t_warm = st0["setup_ms"] * 1e-3, st0["warmup_ms"] * 1e-3
    fr, ms_mse, ms_t = [], [], []
    t_frames = 0.0
    done = 0
    parts = {"gbuffer_ms": 0.0, "train_ms": 0.0, "infer_ms": 0.0, "render_ms": 0.0, "total_ms": 0.0}
    for f in sorted({1, 2, 4, 8, 16, 32, 64, frames}):
        if f > frames:
            continue
        t0 = time.perf_counter()
        st = sc.msnn_frames(f - done)
        t_frames += time.perf_counter() - t0
        for k in parts:
            parts[k] += st[k]
        done = f
        fr.append(f)
        ms_t.append(t_warm + t_frames)
        ms_mse.append(mse(sc.msnn_image(0)))
    m_final = max(ms_mse[-1] - b, 1e-30)
    # what the trained network adds: the same frames' truncated paths alone
    m_short = max(mse(sc.msnn_image(1)) - b, 1e-30)
    k_eq = a / m_final
    t_pt = k_eq * pt_ms_per_pass * 1e-3
    t_ms = ms_t[-1]
    log(f"msnn vs pt (BETA {beta}): MSNN {frames} frames MSE {ms_mse[-1]:.4g} in {t_ms:.2f} s "
        f"(warm-up {t_warm:.2f} s, setup {t_setup:.2f} s not counted) = PT at {k_eq:.1f} spp, {t_pt:.2f} s "
        f"-> speed-up {t_pt / t_ms:.2f}x; RMSE {np.sqrt(m_final):.4f} with the cache, {np.sqrt(m_short):.4f} "
        f"from the truncated paths alone; per frame: " +
        ", ".join(f"{k} {v / frames:.2f}" for k, v in parts.items()))
    return {
        "beta": beta, "frames": frames, "reference_spp": R["spp"],
        "speedup": t_pt / t_ms, "t_msnn_s": t_ms, "t_warmup_s": t_warm, "t_setup_s_not_counted": t_setup,
        "t_pt_equal_rmse_s": t_pt, "msnn_frame_ms": {k: v / frames for k, v in parts.items()},
        "pt_spp_equal_rmse": k_eq, "rmse_msnn": float(np.sqrt(m_final)),
        "rmse_short_paths_only": float(np.sqrt(m_short)),
        "msnn_curve": {"frames": fr, "t_s": ms_t, "mse": ms_mse},
    }


def nrc_vs_pt(sc, pt_ms_per_pass: float, R: dict, frames: int) -> dict:
    """Neural radiance caching (the reference's comparator renderer) vs plain
    path tracing, measured like msnn_vs_pt: NRC trains online from the first
    frame, so its time is the frames' wall time."""
    ref, a, b = R["ref"], R["a"], R["b"]
    sc.synchronize()
    sc.nrc_begin()
    fr, mses, ts = [], [], []
    t_frames, done = 0.0, 0
    parts = {"gbuffer_ms": 0.0, "infer_ms": 0.0, "render_ms": 0.0, "train_ms": 0.0}
    queries = 0
    for f in sorted({1, 2, 4, 8, 16, 32, 64, frames}):
        if f > frames:
            continue
        t0 = time.perf_counter()
        st = sc.nrc_frames(f - done)
        t_frames += time.perf_counter() - t0
        for k in parts:
            parts[k] += st[k]
        queries = st["cache_queries"]
        done = f
        fr.append(f)
        ts.append(t_frames)
        mses.append(float(np.mean((sc.nrc_image()[..., :3].astype(np.float64) - ref) ** 2)))
    m_final = max(mses[-1] - b, 1e-30)
    k_eq = a / m_final
    t_pt = k_eq * pt_ms_per_pass * 1e-3
    log(f"nrc vs pt: NRC {frames} frames MSE {mses[-1]:.4g} in {ts[-1]:.2f} s = PT at {k_eq:.1f} spp, "
        f"{t_pt:.2f} s -> speed-up {t_pt / ts[-1]:.2f}x; per frame: " +
        ", ".join(f"{k} {v / frames:.2f}" for k, v in parts.items()))
    return {"frames": frames, "speedup": t_pt / ts[-1], "t_nrc_s": ts[-1], "t_pt_equal_rmse_s": t_pt,
            "pt_spp_equal_rmse": k_eq, "rmse_nrc": float(np.sqrt(m_final)),
            "nrc_frame_ms": {k: v / frames for k, v in parts.items()}, "cache_queries_last_frame": queries,
            "nrc_curve": {"frames": fr, "t_s": ts, "mse": mses}}


def adjoint_rate(hm, scene_path: Path, size: int = 512, spp: int = 16) -> dict:
    """The differentiable-appearance workload: c1 geometry at size^2, spp
    samples, per-strand sigma_a ~ U(0.01, 0.5) (seed 3) and the global
    roughness / tilt; gradient of an L2 loss against a target rendered with
    the scene's own sigma_a. Times the forward render and the path-replay
    adjoint."""
    cfg = json.loads(scene_path.read_text())
    cfg["integrator"].update({"width": size, "height": size, "spp": spp})
    with hm.Scene(json_text=cfg) as sc:
        target = sc.render_path_tracing(spp)[..., :3].astype(np.float64)
        ns = sc.num_strands
        sig = np.random.default_rng(3).uniform(0.01, 0.5, (ns, 3)).astype(np.float32)
        sc.set_strand_sigma_a(sig)
        sc.synchronize()
        t0 = time.perf_counter()
        img = sc.render_path_tracing(spp)[..., :3].astype(np.float64)
        t_fwd = time.perf_counter() - t0
        dl = 2.0 * (img - target) / img.size
        t0 = time.perf_counter()
        gs, gg = sc.adjoint(dl, spp)
        t_adj = time.perf_counter() - t0
    paths = size * size * spp
    log(f"adjoint: {size}x{size}, {spp} spp, {ns} strands: forward {t_fwd:.3f} s, adjoint {t_adj:.3f} s "
        f"({paths / t_adj / 1e6:.2f} M paths/s)")
    return {"workload": f"c1 geometry, {size}x{size}, {spp} spp, {ns}x3 per-strand sigma_a + beta_m, beta_n, alpha",
            "forward_s": t_fwd, "adjoint_s": t_adj, "adjoint_paths_per_s": paths / t_adj,
            "grad_global": [float(x) for x in gg], "grad_sigma_norm": float(np.linalg.norm(gs))}


def bsdf_rates(hm, n: int = 1 << 22, reps: int = 20) -> dict:
    """Batched hair-BSDF streams (hmsnn_hair_bsdf_eval_dev / _sample_dev) on n
    synthetic queries resident in HBM (wo in the hair frame's y = 0 plane,
    uniform wi, h, sigma_a; beta_m = beta_n = 0.3, alpha = 2 degrees), wall
    clock over reps calls after one warm call. Algorithmic bytes per query:
    eval reads the 40-byte query and writes f (12) + pdf (4); sample also
    reads 4 uniforms (16) and writes wi (12) + f (12) + pdf (4) + lobe (4)."""
    import ctypes as C
    lib = hm._lib.load()
    rng = np.random.default_rng(5)
    so = rng.uniform(-0.995, 0.995, n)
    q = np.zeros((n, 10), np.float32)
    q[:, 0], q[:, 2] = so, np.sqrt(1 - so * so)
    v = rng.normal(size=(n, 3))
    q[:, 3:6] = v / np.linalg.norm(v, axis=1, keepdims=True)
    q[:, 6] = rng.uniform(-0.999, 0.999, n)
    q[:, 7:10] = rng.uniform(0.01, 1.0, (n, 3))
    u = rng.random((n, 4), dtype=np.float32)
    dq, du = hm.DeviceBuffer.from_numpy(q), hm.DeviceBuffer.from_numpy(u)
    dwi, df, dp, dl = hm.DeviceBuffer(12 * n), hm.DeviceBuffer(12 * n), hm.DeviceBuffer(4 * n), hm.DeviceBuffer(4 * n)
    kernel_flops = json.loads(KERNEL_FLOPS_JSON.read_text()).get("kernels", {}) if KERNEL_FLOPS_JSON.exists() else {}
    prm = hm._hair_params(0.3, 0.3, float(np.radians(2.0)))
    calls = {
        "eval": (lambda: lib.hmsnn_hair_bsdf_eval_dev(C.byref(prm), dq.ptr, n, df.ptr, dp.ptr, None), 40 + 16),
        "sample": (lambda: lib.hmsnn_hair_bsdf_sample_dev(C.byref(prm), dq.ptr, du.ptr, n, dwi.ptr, df.ptr, dp.ptr,
                                                          dl.ptr, None), 40 + 16 + 32),
    }
    out = {}
    for name, (fn, nbytes) in calls.items():
        hm._lib.check(fn())
        hm._lib.check(lib.hmsnn_dev_synchronize())
        t0 = time.perf_counter()
        for _ in range(reps):
            hm._lib.check(fn())
        hm._lib.check(lib.hmsnn_dev_synchronize())
        dt = (time.perf_counter() - t0) / reps
        gbs = nbytes * n / dt / 1e9
        out[name] = {"queries": n, "ms": dt * 1e3, "Gqueries/s": n / dt / 1e9, "bytes_per_query": nbytes,
                     "GB/s": gbs, "peak_GB/s": 8000.0, "frac": gbs / 8000.0}
        fl = kernel_flops.get(f"k_bsdf_{name}", {}).get("valu_flops_per_launch")
        if fl:
            out[name]["valu_flops_per_query"] = fl / n
            out[name]["valu_TFLOPs"] = fl / dt / 1e12
            out[name]["valu_frac"] = fl / dt / 1e12 / VALU_FP32_PEAK_TFLOPS
    return out


def mlp_train_rate(hm, n: int = 1 << 16, reps: int = 20) -> dict:
    """One FP16 training step (forward, backward, weight gradient, Adam) on n
    seeded records: 124,032 FLOP per record (3 x the forward's 41,344), wall
    clock over reps steps queued back to back after one warm step."""
    import ctypes as C
    lib = hm._lib.load()
    rng = np.random.default_rng(6)
    rec = np.zeros((n, 12), np.float32)
    rec[:, 0:3] = rng.uniform(-0.5, 0.5, (n, 3))
    for a in (3, 6):
        v = rng.normal(size=(n, 3))
        rec[:, a:a + 3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    rec[:, 9:12] = rng.uniform(0.01, 0.5, (n, 3))
    tgt = rng.uniform(0.1, 2.0, (n, 3)).astype(np.float32)
    din, dt_ = hm.DeviceBuffer.from_numpy(rec), hm.DeviceBuffer.from_numpy(tgt)
    h = C.c_void_p()
    hm._lib.check(lib.hmsnn_mlp_create(0, 1, C.byref(h)))
    loss = C.c_float()
    hm._lib.check(lib.hmsnn_mlp_train_step_dev(h, din.ptr, dt_.ptr, n, 0, C.c_float(1e-3), C.byref(loss), None))
    hm._lib.check(lib.hmsnn_dev_synchronize())
    # steps queued back to back (no loss read-back, which would wait for each)
    t0 = time.perf_counter()
    for _ in range(reps):
        hm._lib.check(lib.hmsnn_mlp_train_step_dev(h, din.ptr, dt_.ptr, n, 0, C.c_float(1e-3), None, None))
    hm._lib.check(lib.hmsnn_dev_synchronize())
    dt = (time.perf_counter() - t0) / reps
    lib.hmsnn_mlp_destroy(h)
    tf = 124032.0 * n / dt / 1e12
    return {"records": n, "ms": dt * 1e3, "TFLOP/s": tf, "peak_TFLOP/s": 2500.0, "frac": tf / 2500.0}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-msnn", action="store_true", help="skip the HairMSNN vs PT time-to-equal-RMSE leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint (differentiable appearance) leg")
    ap.add_argument("--ref-spp", type=int, default=4096, help="passes of the RMSE reference image")
    ap.add_argument("--schedule", choices=("stream", "passes"), default="stream",
                    help="path-tracing schedule: stream (a pixel's next pass starts when its path ends) or passes "
                         "(one wavefront per pass, --pass-concurrency of them in flight)")
    ap.add_argument("--pass-concurrency", type=int, default=4, help="passes in flight with --schedule passes")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="instead of the benchmark, run 8 device-to-device copies of this many MiB (the known-size "
                         "kernel that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="instead of the benchmark, time only the batched BSDF streams and the MLP inference and "
                         "training kernels (a short program for rocprofv3 counter runs)")
    ap.add_argument("--scene", default=str(SCENE))
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (the path tracer would shard "
              "a frame by image tiles, one process per GPU). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.copy_mib > 0:
        import torch
        nbytes = args.copy_mib << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(8):
            dst.copy_(src)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / 8
        print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": dt * 1e3, "GBps_read_plus_write": 2 * nbytes / dt / 1e9}),
              flush=True)
        return

    import hairmsnn_amd as hm
    n, name = hm.device_info()
    if n < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    log(f"device: {name}")
    if args.kernels_only:
        print(json.dumps({"bsdf": bsdf_rates(hm), "mlp_inference": mlp_rates(hm), "mlp_training": mlp_train_rate(hm)}),
              flush=True)
        return
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
    sc.set_pt_schedule(args.schedule)
    sc.set_pass_concurrency(args.pass_concurrency)
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

    # per-kernel timing with HIP events on the stream each kernel runs on, every
    # launch running alone (streaming: one stream; passes: one pass in flight)
    sc.set_pass_concurrency(1)
    passes_t = max(2, min(16, args.steps))
    kt = sc.time_passes(passes_t)
    sc.set_pass_concurrency(args.pass_concurrency)
    launches = max(kt["extend_launches"], 1)
    launches_pass = launches / passes_t
    V, P, S = stats["closest_rays"], float(paths_per_step), stats["any_rays"]
    alg_pass = {
        "extend": stats["closest_nodes"] * NODE_BYTES + stats["closest_segs"] * SEG_BYTES + V * RAY_BYTES,
        "shade": (V * SHADE_VERTEX_BYTES + max(V - P, 0.0) * (SHADE_HIT_BYTES + SHADE_MOVE_BYTES)
                  + P * SHADE_DONE_BYTES + S * SHADE_SHADOW_BYTES),
        "shadow": stats["any_nodes"] * NODE_BYTES + stats["any_segs"] * SEG_BYTES + S * ANY_RAY_BYTES,
    }
    traffic_k = {}
    if TRAFFIC_JSON.exists():
        traffic_k = json.loads(TRAFFIC_JSON.read_text()).get("kernels", {})
    flops_k = {}
    if FLOPS_JSON.exists():
        flops_k = json.loads(FLOPS_JSON.read_text()).get("kernels", {})
    per_kernel = {}
    for k, kname in PASS_KERNELS.items():
        ms_pass = kt[f"{k}_ms"] / passes_t
        ms_launch = ms_pass / launches_pass
        b_launch = alg_pass[k] / launches_pass
        ach = b_launch / (ms_launch * 1e-3) / 1e9
        tr = traffic_k.get(kname, {}).get("traffic_bytes_per_launch")
        fl = flops_k.get(kname, {}).get("valu_flops_per_launch")
        per_kernel[k] = {"kernel": kname, "ms_per_pass": ms_pass, "ms_per_launch": ms_launch,
                         "bytes_per_launch": b_launch, "achieved": ach, "frac": ach / HBM_PEAK_GBS, "traffic": tr,
                         "traffic_GBps": tr / (ms_launch * 1e-3) / 1e9 if tr else None,
                         "valu_TFLOPs": fl / (ms_launch * 1e-3) / 1e12 if fl else None,
                         "valu_frac": fl / (ms_launch * 1e-3) / 1e12 / VALU_FP32_PEAK_TFLOPS if fl else None}
    dom = max(per_kernel, key=lambda k: per_kernel[k]["ms_per_pass"])
    for k, v in per_kernel.items():
        log(f"{k}: {v['ms_per_pass']:.2f} ms/pass, {v['ms_per_launch']:.3f} ms/launch, "
            f"{v['bytes_per_launch'] / 1e6:.1f} MB/launch -> {v['achieved']:.0f} GB/s ({100 * v['frac']:.1f}% of HBM)")
    log(f"dominant kernel by time: {dom}")
    d = per_kernel[dom]
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
            "width": W, "height": H, "paths_per_step": paths_per_step, "schedule": args.schedule,
            "segments": int(sc.info["num_segments"]), "bvh_depth": int(sc.info["bvh_depth"]),
        },
        "roofline": {
            "bound": "hbm", "achieved": d["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": d["frac"], "traffic": d["traffic"], "kernel": d["kernel"],
            "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/pt_c1_1024_hbm_traffic.md",
            "bytes_per_launch": d["bytes_per_launch"], "ms_per_launch": d["ms_per_launch"],
            "nodes_per_ray": stats["closest_nodes"] / max(stats["closest_rays"], 1),
            "segs_per_ray": stats["closest_segs"] / max(stats["closest_rays"], 1),
            "pass_kernels": per_kernel,
        },
        "kernels_ms_per_pass": {
            "extend": kt["extend_ms"] / passes_t, "shade": kt["shade_ms"] / passes_t,
            "shadow": kt["shadow_ms"] / passes_t, "total": kt["total_ms"] / passes_t,
        },
        "rays_per_pass": {"closest": stats["closest_rays"], "any": stats["any_rays"]},
        "rays_per_s": {"closest": stats["closest_rays"] * args.steps / wall, "any": stats["any_rays"] * args.steps / wall},
    }
    if not args.no_msnn:
        R = pt_reference(sc, args.ref_spp, 64)
        line["msnn_vs_pt"] = msnn_vs_pt(sc, wall / args.steps * 1e3, R, 64, beta=1)
        line["msnn_vs_pt"]["pt_fit"] = {"a": R["a"], "b_reference_noise": R["b"]}
        line["msnn_vs_pt"]["pt_curve"] = {"spp": R["ks"], "mse": R["mse"]}
        line["msnn_vs_pt"]["beta_sweep"] = {
            str(bt): {k: v for k, v in msnn_vs_pt(sc, wall / args.steps * 1e3, R, 64, beta=bt).items()
                      if k in ("speedup", "t_msnn_s", "pt_spp_equal_rmse", "rmse_msnn", "rmse_short_paths_only")}
            for bt in (2, 5, 10)}
        # the same comparison with the radiance cache trained in BF16 and
        # queried in FP8-E4M3, against the same reference image
        cfg8 = json.loads(Path(args.scene).read_text())
        cfg8.setdefault("hairmsnn", {})["precision"] = "fp8"
        with hm.Scene(json_text=cfg8, base_dir=str(Path(args.scene).resolve().parent)) as sc8:
            line["msnn_vs_pt"]["fp8"] = {
                k: v for k, v in msnn_vs_pt(sc8, wall / args.steps * 1e3, R, 64, beta=1).items()
                if k in ("speedup", "t_msnn_s", "pt_spp_equal_rmse", "rmse_msnn", "rmse_short_paths_only")}
        line["nrc_vs_pt"] = nrc_vs_pt(sc, wall / args.steps * 1e3, R, 64)
        fm = line["msnn_vs_pt"]["msnn_frame_ms"]
        train_records = int(json.loads(Path(args.scene).read_text()).get("hairmsnn", {}).get("train_records", 16384))
        line["mlp"] = {
            "frame_inference_TFLOP/s": 41344.0 * paths_per_step / (fm["infer_ms"] * 1e-3) / 1e12,
            "frame_training_TFLOP/s": 124032.0 * train_records / (fm["train_ms"] * 1e-3) / 1e12,
        }
        try:
            line["mlp"]["inference"] = mlp_rates(hm)
            line["mlp"]["training"] = mlp_train_rate(hm)
        except Exception as e:
            log(f"mlp timing failed: {e}")
    try:
        line["bsdf"] = bsdf_rates(hm)
    except Exception as e:
        log(f"bsdf timing failed: {e}")
    if not args.no_adjoint:
        try:
            line["adjoint"] = adjoint_rate(hm, Path(args.scene))
        except Exception as e:
            log(f"adjoint timing failed: {e}")
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
