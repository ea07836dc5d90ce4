"""Benchmark of the fused fusion frame on one MI355X.

A step is one fused frame of one session at 640x480 with a dense 512^3 TSDF
over 3 m (truncation 2.5 cm): preprocessing (passing + bilateral filter,
3-level pyramid), projective ICP with 10/5/4 iterations, integration of every
frame and raycast of the reference maps.  Synthetic frames (seeded orbit
scene) are uploaded to HBM before the timed region; the timed region is
bracketed by device synchronisation.

Prints one JSON line with the metric, the roofline of the dominant kernel
(timed with HIP events on the session stream) and the CPU oracle timed on a
bounded sample of the same workload.
"""
import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fused frames/s @640x480 (ICP+integrate+raycast); p99 frame latency @32 sessions"
HBM_PEAK_GBS = 8000.0
STORAGES = {"s16": 0, "f16": 1, "f32": 2}


def cpu_baseline(wl, frames):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    eng = O.OracleEngine(wl, prec=0, gate=0.0)
    depths = [wl.render(f)[0] for f in range(frames)]
    eng.run(depths[0])  # first frame: integrate only
    t0 = time.perf_counter()
    for d in depths[1:]:
        eng.run(d)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    return {"value": round((frames - 1) / dt, 3), "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": f"{frames - 1} fused frames (after frame 0) of the same 640x480 / 512^3 workload, "
                      f"float oracle with the reference's short2 volume, OpenMP"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--storage", choices=sorted(STORAGES), default="s16")
    ap.add_argument("--graph", type=int, default=0)
    ap.add_argument("--cpu-frames", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the fusion sessions shard by session with no collective, "
              f"and this build measures one GPU's share (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    from natureuserinterfacestudio_amd import nuif
    from natureuserinterfacestudio_amd import workloads as W

    wl = W.FULL
    nuif.init(0)
    info = nuif.device_info()
    n = args.warmup + args.steps
    if n > wl.frames:
        wl = dataclasses.replace(wl, frames=n)
    print(f"[bench] {info['name']} CUs={info['cu']}; rendering {n} frames of {wl.name}", file=sys.stderr, flush=True)
    frames = [nuif.DeviceArray.from_host(wl.render(f)[0]) for f in range(n)]
    storage = STORAGES[args.storage]
    eng = nuif.Engine(wl.kinfu_desc(storage=storage, gate=0.0, use_graph=bool(args.graph)))

    for f in range(args.warmup):
        eng.run(frames[f], sync=False)
    eng.synchronize()
    eng.set_stage_timing(not args.graph)
    eng.take_stage_times()
    upd0 = eng.updated_voxels()

    start, stop = nuif.Event(), nuif.Event()
    nuif.check(nuif.L().nuif_device_synchronize(), "sync")
    t0 = time.perf_counter()
    start.record(eng.stream)
    for f in range(args.warmup, n):
        eng.run(frames[f], sync=False)
    stop.record(eng.stream)
    stop.synchronize()
    nuif.check(nuif.L().nuif_device_synchronize(), "sync")
    wall = time.perf_counter() - t0
    gpu_ms = start.elapsed_ms(stop)
    ok = eng.synchronize()
    timed_frames, stage_ms = eng.take_stage_times()
    upd = eng.updated_voxels() - upd0

    ms_per_step = gpu_ms / args.steps
    value = args.steps / (gpu_ms / 1e3)
    stage_avg = [s / max(timed_frames, 1) for s in stage_ms]
    names = ["preprocess", "icp", "integrate", "raycast"]

    # dominant kernel: integration (one launch per frame).  Algorithmic bytes per
    # launch: read + write of every updated voxel (B bytes each) + the depth image.
    B = nuif.L().nuif_tsdf_voxel_bytes(storage)
    upd_per = upd / max(args.steps, 1)
    integ_bytes = 2 * B * upd_per + 4 * wl.width * wl.height
    roof = None
    if timed_frames:
        t_int = stage_avg[2] / 1e3
        achieved = integ_bytes / t_int / 1e9
        roof = {"kernel": "integrate_tile_kernel", "bound": "hbm", "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "algorithmic_bytes": int(integ_bytes), "avg_ms": round(stage_avg[2], 4),
                "updated_voxels_per_launch": int(upd_per)}

    cpu = None
    if not args.no_cpu:
        print(f"[bench] CPU baseline on {args.cpu_frames} frames", file=sys.stderr, flush=True)
        cpu = cpu_baseline(wl, args.cpu_frames)

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{wl.name}: 1 session, fused frame (preprocess + ICP 10/5/4 + integrate every "
                               f"frame + raycast)", "storage": args.storage, "graph": bool(args.graph),
                   "resolution": [wl.width, wl.height], "volume": [wl.volume_res] * 3, "volume_m": wl.volume_dim,
                   "tracking_ok": bool(ok)},
        "stage_ms": {k: round(v, 4) for k, v in zip(names, stage_avg)},
        "wall_s": round(wall, 4),
        "roofline": roof,
        "cpu_baseline": cpu,
        "device": info["name"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
