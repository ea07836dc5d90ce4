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
import ctypes as C
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
    eng.stage_seconds()
    t0 = time.perf_counter()
    for d in depths[1:]:
        eng.run(d)
    dt = time.perf_counter() - t0
    st = eng.stage_seconds() / (frames - 1) * 1e3
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round((frames - 1) / dt, 3), "unit": "frames/s", "cores": cores, "kind": "port",
            "cpu_model": model,
            "stage_ms": {k: round(float(v), 2) for k, v in zip(["preprocess", "icp", "integrate", "raycast"], st)},
            "sample": f"{frames - 1} fused frames (after frame 0) of the same 640x480 / 512^3 workload, "
                      f"float oracle with the reference's short2 volume, OpenMP"}


def fp64_check(nuif, wl, frames, storage, joint=False):
    """The server's session type (FP16 or FP32 volume, compensated ICP) on the GPU against
    the FP64 oracle on the same frames -- the oracle as the checker, on the host: pose
    drift per frame against the FP64 engine, the TSDF against an FP64 volume integrated at
    the GPU's poses (voxels observed by both with equal weights, in units of the
    truncation distance) and the raycast vertices against an FP64 raycast of that volume
    at the last pose (rays hit by both).  joint: the joint ICP order with colour frames
    against the FP64 instance of the joint order."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    rendered = [wl.render(f, color=joint)[:2] for f in range(frames)]
    depths = [r[0] for r in rendered]
    red = nuif.ICP_REDUCE_JOINT if joint else nuif.ICP_REDUCE_COMPENSATED
    eng = nuif.Engine(wl.kinfu_desc(storage=storage, gate=0.0, reduction=red, color=joint))
    for d, b in rendered:
        eng.run(nuif.DeviceArray.from_host(d), nuif.DeviceArray.from_host(b) if joint else None, sync=False)
    ok = eng.synchronize() == 1
    poses = eng.poses(frames)
    o64 = O.OracleEngine(wl, prec=1, gate=0.0, compensated=not joint, joint=joint, color=joint)
    vd = O.OracleVolumeD(wl)
    dt_max, ang_max = 0.0, 0.0
    drift = []
    for f, (d, b) in enumerate(rendered):
        _, p64, _ = o64.run(d, b)
        R, t = poses[f, :9].reshape(3, 3).astype(np.float64), poses[f, 9:12].astype(np.float64)
        R64, t64 = p64[:9].reshape(3, 3), p64[9:]
        dt_max = max(dt_max, float(np.linalg.norm(t - t64)))
        # the angle as atan2(|vee(skew part)|, (tr - 1) / 2): arccos of the trace would
        # turn the float32 rounding of R into ~0.01 deg near zero
        Rr = R.T @ R64
        w = 0.5 * np.array([Rr[2, 1] - Rr[1, 2], Rr[0, 2] - Rr[2, 0], Rr[1, 0] - Rr[0, 1]])
        ang_max = max(ang_max, float(np.degrees(np.arctan2(np.linalg.norm(w), (np.trace(Rr) - 1.0) / 2.0))))
        drift.append(round(float(np.linalg.norm(t - t64)) * 1e3, 4))
        vd.integrate(d, poses[f, :9].reshape(3, 3), poses[f, 9:12])
    o64.close()
    raw = eng.volume_raw()
    if storage == nuif.TSDF_F16:
        w32 = raw.view(np.uint32)
        t_g = (w32 & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float64)
        w_g = (w32 >> 16).astype(np.int64)
    else:
        f2 = raw.view(np.float32).reshape(-1, 2)
        t_g, w_g = f2[:, 0].astype(np.float64), f2[:, 1].astype(np.int64)
    t_d, w_d = vd.volume()
    same = (w_g > 0) & (w_d > 0) & (w_g == w_d)
    terr = np.abs(t_g[same] - t_d[same])
    vg, _ = eng.maps()
    vr, _ = vd.raycast(poses[-1, :9].reshape(3, 3), poses[-1, 9:12])
    vd.close()
    eng.close()
    hg, hd = ~np.isnan(vg[:, 2]), ~np.isnan(vr[:, 2])
    hit = hg & hd
    verr = np.linalg.norm(vg[hit].astype(np.float64) - vr[hit], axis=1)
    return {"session": f"{'FP16' if storage == nuif.TSDF_F16 else 'FP32'} volume, "
                       f"{'joint ICP (colour weight 0.1)' if joint else 'compensated ICP'}",
            "frames": frames, "tracking_ok": bool(ok), "pose_drift_mm_per_frame": drift,
            "pose_drift_max_mm": round(dt_max * 1e3, 5), "pose_drift_max_deg": round(ang_max, 6),
            "tsdf_err_over_mu": {"voxels": int(same.sum()), "p99": float(np.quantile(terr, 0.99)),
                                 "p999": float(np.quantile(terr, 0.999)), "max": float(terr.max()),
                                 "beyond_1e-3": round(float((terr > 1e-3).mean()), 6)},
            "raycast_vertex_err_mm": {"rays": int(hit.sum()), "p99": round(float(np.quantile(verr, 0.99)) * 1e3, 4),
                                      "p999": round(float(np.quantile(verr, 0.999)) * 1e3, 4),
                                      "beyond_0.5mm": round(float((verr > 5e-4).mean()), 6),
                                      "hit_pattern_differs": round(float((hg != hd).sum() / max(hd.sum(), 1)), 6)}}


def measured_traffic(key):
    """HBM bytes per launch of a kernel from the committed PMC summary (FETCH_SIZE and
    WRITE_SIZE scaled by a plain copy of known size), or None when not measured."""
    path = os.path.join(ROOT, "profiles", "hbm_bytes_full_640x480_512.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"traffic_bytes_{key}:"):
            return int(line.split(":")[1])
    return None


def run_single(nuif, wl, frames, storage, graph, warmup, steps, reduction=None, stage_timing=True):
    """One session, frames back to back on its stream.  Returns throughput, the
    per-stage device times (event-timed when graphs are off) and updated voxels."""
    eng = nuif.Engine(wl.kinfu_desc(storage=storage, gate=0.0, use_graph=bool(graph), reduction=reduction))
    for f in range(warmup):
        eng.run(frames[f], sync=False)
    eng.synchronize()
    eng.set_stage_timing(stage_timing and not graph)
    eng.take_stage_times()
    upd0 = eng.updated_voxels()
    start, stop = nuif.Event(), nuif.Event()
    nuif.check(nuif.L().nuif_device_synchronize(), "sync")
    t0 = time.perf_counter()
    start.record(eng.stream)
    for f in range(warmup, warmup + steps):
        eng.run(frames[f], sync=False)
    stop.record(eng.stream)
    stop.synchronize()
    nuif.check(nuif.L().nuif_device_synchronize(), "sync")
    wall = time.perf_counter() - t0
    gpu_ms = start.elapsed_ms(stop)
    ok = eng.synchronize()
    timed, stage_ms = eng.take_stage_times()
    upd = eng.updated_voxels() - upd0
    eng.close()
    return {"fps": steps / (gpu_ms / 1e3), "ms_per_frame": gpu_ms / steps, "wall_s": wall, "ok": bool(ok),
            "stage_avg": [x / max(timed, 1) for x in stage_ms] if timed else None,
            "updated_per_frame": upd / max(steps, 1)}


def run_server(nuif, wl, frames, sessions, storage, warmup, rounds, budget, round_graph=True):
    """`sessions` resident sessions; each round submits one frame per session and
    waits for all of them (closed loop).  Session s plays frames s, s+1, ..."""
    srv = nuif.Server(budget)
    srv.set_round_graph(round_graph)
    desc = wl.kinfu_desc(storage=storage, gate=0.0, use_graph=True)
    ids = []
    for _ in range(sessions):
        sid = srv.open(desc)
        if sid is None:
            break
        ids.append(sid)
    for r in range(warmup):  # first frames, graph capture
        srv.submit(ids, [frames[r + i] for i in range(len(ids))])
        srv.wait()
    # timed rounds, back to back (each round waits for the previous one)
    nuif.check(nuif.L().nuif_device_synchronize(), "sync")
    lat = []
    t0 = time.perf_counter()
    for r in range(warmup, warmup + rounds):
        srv.submit(ids, [frames[r + i] for i in range(len(ids))])
        lat.append(srv.wait())
    nuif.check(nuif.L().nuif_device_synchronize(), "sync")
    t_all = time.perf_counter() - t0
    ok = all(srv.sessions[i].synchronize() == 1 for i in ids)
    lat = np.concatenate(lat)
    out = {"sessions": len(ids), "fps": len(ids) * rounds / t_all, "round_ms": t_all / rounds * 1e3,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "max_ms": float(lat.max()), "bytes_in_use": int(srv.bytes_in_use), "ok": bool(ok)}
    out["stages"] = server_stages(nuif, srv, ids, wl, frames, warmup + rounds, min(rounds, 10), storage)
    srv.close()
    return out


def run_hash(nuif, wl, warmup, steps):
    """One voxel-hashing session (5 mm voxels in 8^3 FP16 + uint16 bricks, 2^20 buckets) at
    the workload's resolution, frames back to back with its stages event-timed: frames/s,
    stage ms and the integration stage against HBM peak (each visible brick read and
    written whole, 2 x 2 KiB + 16 B of entry per brick, plus the 4 W H bytes of depth)."""
    frames = [nuif.DeviceArray.from_host(wl.render(f)[0]) for f in range(warmup + steps)]
    eng = nuif.Engine(wl.kinfu_desc(storage=STORAGES["f16"], gate=0.0, bricks=1 << 18, buckets=1 << 20))
    for f in range(warmup):
        eng.run(frames[f], sync=False)
    eng.synchronize()
    eng.set_stage_timing(True)
    eng.take_stage_times()
    visible = []
    t0 = time.perf_counter()
    for f in range(warmup, warmup + steps):
        eng.run(frames[f], sync=False)
        visible.append(eng.hash_stats()[1])  # synchronises: the visible list of this frame
    wall = time.perf_counter() - t0
    ok = eng.synchronize()
    timed, ms = eng.take_stage_times()
    stats = eng.hash_stats()
    t0 = time.perf_counter()
    pts, idx = eng.mesh_shape()  # Volume2Mesh over the visible list: marching cubes + host copy
    mesh_ms = (time.perf_counter() - t0) * 1e3
    eng.close()
    avg = [x / max(timed, 1) for x in ms]
    names = ["preprocess", "icp", "integrate", "raycast"]
    device_ms = sum(avg)
    vis = float(np.mean(visible))
    int_bytes = vis * (2 * 2048 + 16) + 4 * wl.width * wl.height
    return {"workload": wl.name, "frames": steps, "fps": round(1e3 / device_ms, 3),
            "ms_per_frame": round(device_ms, 4), "wall_fps": round(steps / wall, 3), "tracking_ok": bool(ok),
            "stage_ms": {k: round(v, 4) for k, v in zip(names, avg)}, "visible_bricks_per_frame": round(vis, 1),
            "bricks_in_use": stats[0], "failed_insertions": stats[2], "buckets": 1 << 20,
            "mesh_wall_ms": round(mesh_ms, 2), "mesh_triangles": len(idx) // 3, "mesh_points": len(pts),
            "integrate_roofline": {"bound": "hbm", "algorithmic_bytes": int(int_bytes),
                                   "achieved": round(int_bytes / (avg[2] / 1e3) / 1e9, 2), "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": round(int_bytes / (avg[2] / 1e3) / 1e9 / HBM_PEAK_GBS, 5)}}


def server_stages(nuif, srv, ids, wl, frames, first, rounds, storage):
    """Further rounds with the batched round's stages event-timed on the server stream,
    and the algorithmic bytes of each stage from what the rounds did: updated voxels
    (integration: 2 B n_upd + 4 W H per session), iterations run per pyramid level (ICP:
    12 W_l H_l + 24 W H + 2 x 232 B per 512-pixel tile of partial sums, per iteration) and
    distinct voxels read (raycast: B U + 28 W H per session, U counted at the last pose)."""
    upd0 = [srv.sessions[i].updated_voxels() for i in ids]
    srv.set_stage_timing(True)
    srv.take_stage_times()
    icp_bytes = 0.0
    W, H = wl.width, wl.height
    for r in range(first, first + rounds):
        srv.submit(ids, [frames[r + i] for i in range(len(ids))])
        srv.wait()
        for i in ids:
            its = srv.sessions[i].state().level_iterations
            for lv in range(len(wl.iterations)):
                px = (W >> lv) * (H >> lv)
                icp_bytes += its[lv] * (12 * px + 24 * W * H + 2 * 232 * -(-px // 512))
    n, ms = srv.take_stage_times()
    srv.set_stage_timing(False)
    B = nuif.L().nuif_tsdf_voxel_bytes(storage)
    upd = sum(srv.sessions[i].updated_voxels() - u for i, u in zip(ids, upd0)) / max(n, 1)
    U = [srv.sessions[i].raycast_footprint() for i in ids]
    names = ["preprocess", "bounds", "icp", "integrate", "raycast"]
    avg = {k: v / max(n, 1) for k, v in zip(names, ms)}
    alg = {"icp": icp_bytes / max(n, 1), "integrate": 2 * B * upd + 4 * W * H * len(ids),
           "raycast": B * sum(U) + 28 * W * H * len(ids)}
    kernels = {"icp": "icp_tiles_kernel<1> + icp_finish_kernel<1> (all steps of the round)",
               "integrate": f"column_states_batch_kernel + integrate_tile_batch_kernel<{storage}>",
               "raycast": "raycast_batch_kernel"}
    roofs = []
    for k in ("icp", "integrate", "raycast"):
        gbs = alg[k] / (avg[k] / 1e3) / 1e9
        roofs.append({"stage": k, "kernel": kernels[k], "bound": "hbm", "achieved": round(gbs, 2),
                      "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5),
                      "traffic": measured_traffic(f"server_{k}"), "algorithmic_bytes": int(alg[k]),
                      "avg_ms": round(avg[k], 4)})
    return {"rounds": n, "avg_ms": {k: round(v, 4) for k, v in avg.items()},
            "updated_voxels_per_round": int(upd), "raycast_unique_voxels_per_session": int(np.mean(U)),
            "rooflines": roofs}


def run_c4(nuif, W, warmup, rounds, budget, pool_frames=16, unbatched=True):
    """The joint-ICP mix: 16 colour sessions (6 x 320x240 256^3, 6 x 640x480 and
    4 x 1280x720 512^3, FP16 volumes), closed-loop rounds of one frame per session with
    its BGRA colours.  Measured three ways: the joint order batched (all sessions' joint
    iterations in lockstep launches of one round), the same sessions in the geometric
    (compensated) order batched, and the joint sessions unbatched (each session's
    graph-captured frame pipeline on its own stream).  Frames come from per-resolution
    pools resident in HBM; session k of a resolution plays frames k, k+1, ..."""
    mix = W.joint_mix()
    pools = {}
    for wl, _ in mix:
        if wl.name not in pools:
            pools[wl.name] = [tuple(nuif.DeviceArray.from_host(a) for a in wl.render(f, color=True)[:2])
                              for f in range(pool_frames)]
    first = {}
    for k, (wl, _) in enumerate(mix):
        first.setdefault(wl.name, k)

    def frame(k, r):
        wl = mix[k][0]
        period = 2 * pool_frames - 2
        q = (r + k - first[wl.name]) % period
        return pools[wl.name][q if q < pool_frames else period - q]

    def leg(reduction, batching, stages=False):
        srv = nuif.Server(budget)
        srv.set_batching(batching)
        ids = []
        for wl, _ in mix:
            ids.append(srv.open(wl.kinfu_desc(storage=nuif.TSDF_F16, gate=0.0, color=True, use_graph=True,
                                              reduction=reduction)))
        assert None not in ids

        def submit(r):
            fr = [frame(k, r) for k in range(len(ids))]
            srv.submit(ids, [f[0] for f in fr], [f[1] for f in fr])

        for r in range(warmup):
            submit(r)
            srv.wait()
        nuif.check(nuif.L().nuif_device_synchronize(), "sync")
        lat = []
        t0 = time.perf_counter()
        for r in range(warmup, warmup + rounds):
            submit(r)
            lat.append(srv.wait())
        nuif.check(nuif.L().nuif_device_synchronize(), "sync")
        t_all = time.perf_counter() - t0
        ok = all((srv.sessions[i].synchronize() == 1) for i in ids)
        lat = np.concatenate(lat)
        # ICP iterations the last frame ran per level (mean over the sessions)
        its = np.mean([list(srv.sessions[i].state().level_iterations)[:3] for i in ids], axis=0)
        out = {"fps": round(len(ids) * rounds / t_all, 1), "round_ms": round(t_all / rounds * 1e3, 3),
               "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
               "tracking_ok": bool(ok), "icp_iterations_per_level": [round(float(v), 2) for v in its]}
        if stages:
            srv.set_stage_timing(True)
            srv.take_stage_times()
            n = min(rounds, 10)
            for r in range(warmup + rounds, warmup + rounds + n):
                submit(r)
                srv.wait()
            m, ms = srv.take_stage_times()
            srv.set_stage_timing(False)
            names = ["preprocess", "bounds", "icp", "integrate", "raycast"]
            out["stage_ms"] = {k: round(v / max(m, 1), 4) for k, v in zip(names, ms)}
        srv.close()
        return out

    joint = leg(nuif.ICP_REDUCE_JOINT, True, stages=True)
    geometric = leg(nuif.ICP_REDUCE_COMPENSATED, True, stages=True)
    unb = leg(nuif.ICP_REDUCE_JOINT, False) if unbatched else None
    return {"sessions": len(mix), "mix": "6 x 320x240 (256^3), 6 x 640x480 (512^3), 4 x 1280x720 (512^3); "
            "FP16 volumes with colour over 3 m; colour weight 0.1", "rounds": rounds,
            "joint_batched": joint, "geometric_batched": geometric, "joint_unbatched": unb,
            "joint_over_unbatched": round(joint["fps"] / unb["fps"], 3) if unb else None,
            "photometric_icp_cost": round(joint["stage_ms"]["icp"] / geometric["stage_ms"]["icp"], 3)}


C5_SEED = 0x4E554905  # "NUI\0" + 5


def c5_trace(sessions=64, rate=2.0, seed=C5_SEED):
    """The multi-session arrival trace: Poisson arrivals at `rate` per second, 200-600
    frames at 30 Hz per session, resolution uniform over 320x240 / 640x480 / 1280x720,
    scene uniform over dense 256^3, dense 512^3 (3 m, FP16 volumes) and hash 5 mm."""
    rng = np.random.default_rng(seed)
    arrive = np.cumsum(rng.exponential(1.0 / rate, sessions))
    frames = rng.integers(200, 601, sessions)
    res = rng.integers(0, 3, sessions)
    kind = rng.integers(0, 3, sessions)
    return arrive, frames, res, kind


def run_c5(nuif, W, sessions, rate, budget, pool_frames=40):
    """Virtual time in 30 Hz ticks: at each tick the sessions that have arrived are
    admitted while the device-memory budget allows (the rest wait, first come first
    served), then every active session submits its next frame as one server round.
    A failed frame resets its session (the reference's auto reset).  Frames come from
    per-(resolution, scene) pools played forwards and backwards, resident in HBM."""
    arrive, nfr, res, kind = c5_trace(sessions, rate)
    dense_wl = [dataclasses.replace(W.SMALL, volume_dim=3.0, trunc=0.025), W.FULL, W.HD]
    hash_wl = [W.HASH_SMALL, W.HASH, W.HASH_HD]
    pools = {}
    for r in range(3):
        for k in (0, 1):
            wl = dense_wl[r] if k == 0 else hash_wl[r]
            pools[(r, k)] = [nuif.DeviceArray.from_host(wl.render(f)[0]) for f in range(pool_frames)]

    def desc(i):
        r, k = int(res[i]), int(kind[i])
        if k == 2:
            return hash_wl[r].kinfu_desc(storage=STORAGES["f16"], gate=0.0)
        wl = dataclasses.replace(dense_wl[r], volume_res=256 if k == 0 else 512)
        return wl.kinfu_desc(storage=STORAGES["f16"], gate=0.0)

    def frame(i, f):
        period = 2 * pool_frames - 2
        q = f % period
        return pools[(int(res[i]), 1 if kind[i] == 2 else 0)][q if q < pool_frames else period - q]

    srv = nuif.Server(budget)
    srv.set_failure_policy(nuif.ON_FAILURE_RESET)
    tick = 1.0 / 30
    t, nxt, queue, active = 0.0, 0, [], {}
    lat, round_ms, peak, peak_bytes, refused, failures, done, rounds_over = [], [], 0, 0, 0, 0, 0, 0
    gpu_t = 0.0
    while nxt < sessions or queue or active:
        while nxt < sessions and arrive[nxt] <= t:
            queue.append(nxt)
            nxt += 1
        while queue:
            sid = srv.open(desc(queue[0]))
            if sid is None:
                refused += 1
                break
            active[sid] = [queue.pop(0), 0]
        if active:
            ids = list(active)
            fr = [frame(active[s][0], active[s][1]) for s in ids]
            t0 = time.perf_counter()
            srv.submit(ids, fr)
            lat.append(srv.wait())
            dt = time.perf_counter() - t0
            gpu_t += dt
            round_ms.append(dt * 1e3)
            rounds_over += dt > tick
            failures += int((srv.last_outcomes() == 0).sum())
            peak = max(peak, len(ids))
            peak_bytes = max(peak_bytes, srv.bytes_in_use)
            for s in ids:
                active[s][1] += 1
                if active[s][1] >= nfr[active[s][0]]:
                    srv.close_session(s)
                    del active[s]
                    done += 1
        t += tick
    srv.close()
    lat = np.concatenate(lat)
    rm = np.array(round_ms)
    return {"sessions": int(sessions), "arrival_rate_per_s": rate, "frames": int(lat.size),
            "trace_s": round(t, 2), "completed": done, "peak_concurrent": peak,
            "peak_bytes": int(peak_bytes), "admission_waits": refused, "tracking_failures": failures,
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "max_ms": round(float(lat.max()), 3), "round_p99_ms": round(float(np.percentile(rm, 99)), 3),
            "rounds_over_33ms": int(rounds_over), "aggregate_fps": round(lat.size / gpu_t, 1),
            "mix": "resolution uniform over 320x240/640x480/1280x720; dense 256^3 / dense 512^3 over 3 m "
                   "(FP16 volumes, compensated ICP) / hash 5 mm; failures reset the session"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60, help="timed frames (single) / rounds (server)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--storage", choices=sorted(STORAGES), default="s16", help="single-session volume storage")
    ap.add_argument("--graph", type=int, default=0, help="single session replayed from a hipGraph")
    ap.add_argument("--sessions", type=int, default=32, help="server sessions (0: single session only)")
    ap.add_argument("--server-storage", choices=sorted(STORAGES), default="f16")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="server device-memory budget")
    ap.add_argument("--cpu-frames", type=int, default=16)
    ap.add_argument("--c4-rounds", type=int, default=30, help="joint-ICP mix rounds (0: skip that leg)")
    ap.add_argument("--c4-batched-only", action="store_true", help="joint-ICP mix: skip the unbatched comparison")
    ap.add_argument("--no-round-graph", action="store_true", help="server: enqueue batched rounds launch by launch")
    ap.add_argument("--c5-sessions", type=int, default=64, help="arrival-trace sessions (0: skip that leg)")
    ap.add_argument("--c5-rate", type=float, default=2.0, help="arrivals per second")
    ap.add_argument("--hash-steps", type=int, default=20, help="frames of the 640x480 hash session (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--server-only", action="store_true",
                    help="skip the single-session and FP32 serial legs (profiles of the batched round)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy this many MiB device to device and exit (scales the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the fusion sessions shard by session with no collective, "
              f"and this build measures one GPU's share (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    if args.server_only and args.sessions <= 0:
        ap.error("--server-only needs --sessions > 0")
    from natureuserinterfacestudio_amd import nuif
    from natureuserinterfacestudio_amd import workloads as W

    nuif.init(0)
    if args.copy_probe:
        # a plain copy of known size: FETCH_SIZE / WRITE_SIZE per byte of this copy
        # convert the counters of the measured kernels to bytes
        n = args.copy_probe << 20
        a, b = nuif.DeviceArray((n,), np.uint8), nuif.DeviceArray((n,), np.uint8)
        for _ in range(3):
            nuif.check(nuif.L().nuif_memcpy_d2d(b.ptr, a.ptr, n, None), "copy")
        nuif.check(nuif.L().nuif_device_synchronize(), "sync")
        print(json.dumps({"copy_probe_bytes": n, "launches": 3}))
        return
    wl = W.FULL
    info = nuif.device_info()
    n = args.warmup + args.steps + max(args.sessions - 1, 0)
    if args.sessions > 0:
        n += min(args.steps, 10)  # the server's stage-timed rounds
    if n > wl.frames:
        wl = dataclasses.replace(wl, frames=n)
    print(f"[bench] {info['name']} CUs={info['cu']}; rendering {n} frames of {wl.name}", file=sys.stderr, flush=True)
    frames = [nuif.DeviceArray.from_host(wl.render(f)[0]) for f in range(n)]

    # 1) one session, reference-precision volume, stage-timed: the roofline of the integration kernel
    storage = STORAGES[args.storage]
    names = ["preprocess", "icp", "integrate", "raycast"]
    roof = single = None
    if not args.server_only:
        print(f"[bench] single session, storage {args.storage}", file=sys.stderr, flush=True)
        single = run_single(nuif, wl, frames, storage, args.graph, args.warmup, args.steps)
        B = nuif.L().nuif_tsdf_voxel_bytes(storage)
        integ_bytes = 2 * B * single["updated_per_frame"] + 4 * wl.width * wl.height
    if single and single["stage_avg"]:
        t_int = single["stage_avg"][2] / 1e3
        achieved = integ_bytes / t_int / 1e9
        roof = {"kernel": f"integrate_tile_kernel<{args.storage}>", "bound": "hbm", "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": measured_traffic(f"integrate_tile_kernel_{args.storage}"),
                "algorithmic_bytes": int(integ_bytes), "avg_ms": round(single["stage_avg"][2], 4),
                "updated_voxels_per_launch": int(single["updated_per_frame"])}

    server = fp32 = fp32c = None
    if args.sessions > 0:
        # 2) the plain FP32 pipeline, one session at a time (the server's reference point):
        #    FP32 volume, the reference's float ICP sums and LM update; beside it the same
        #    session with the server's compensated ICP
        if not args.server_only:
            print("[bench] FP32 single session, serial", file=sys.stderr, flush=True)
            fp32 = run_single(nuif, wl, frames, STORAGES["f32"], 0, args.warmup, args.steps,
                              reduction=nuif.ICP_REDUCE_REFERENCE, stage_timing=False)
            fp32c = run_single(nuif, wl, frames, STORAGES["f32"], 0, args.warmup, args.steps,
                               reduction=nuif.ICP_REDUCE_COMPENSATED, stage_timing=False)
        # 3) the multi-session server
        print(f"[bench] server: {args.sessions} sessions, storage {args.server_storage}", file=sys.stderr, flush=True)
        server = run_server(nuif, wl, frames, args.sessions, STORAGES[args.server_storage], args.warmup, args.steps,
                            args.budget_gb * 2 ** 30, round_graph=not args.no_round_graph)

    small = None
    if not args.server_only:
        # the small configuration: 320x240, dense FP32 volume of 128^3 over 2 m, the
        # reference's ICP order, its 30 frames (5 warm-up, 25 timed)
        print(f"[bench] small session: {W.SMALL.name}", file=sys.stderr, flush=True)
        sf = [nuif.DeviceArray.from_host(W.SMALL.render(f)[0]) for f in range(W.SMALL.frames)]
        r = run_single(nuif, W.SMALL, sf, STORAGES["f32"], 0, 5, W.SMALL.frames - 5,
                       reduction=nuif.ICP_REDUCE_REFERENCE)
        small = {"workload": W.SMALL.name, "storage": "f32", "icp": "reference order", "fps": round(r["fps"], 1),
                 "ms_per_frame": round(r["ms_per_frame"], 4), "tracking_ok": r["ok"],
                 "stage_ms": {k: round(v, 4) for k, v in zip(names, r["stage_avg"])} if r["stage_avg"] else None}

    hashed = None
    if args.hash_steps > 0:
        print(f"[bench] hash session: {W.HASH.name}", file=sys.stderr, flush=True)
        hashed = run_hash(nuif, W.HASH, min(args.warmup, 5), args.hash_steps)

    c4 = None
    if args.c4_rounds > 0:
        print(f"[bench] joint-ICP mix: 16 sessions, {args.c4_rounds} rounds", file=sys.stderr, flush=True)
        c4 = run_c4(nuif, W, 5, args.c4_rounds, args.budget_gb * 2 ** 30, unbatched=not args.c4_batched_only)
        print(f"[bench] joint-ICP mix done: {c4['joint_batched']['fps']} fps batched", file=sys.stderr, flush=True)
    c5 = None
    if args.c5_sessions > 0:
        print(f"[bench] arrival trace: {args.c5_sessions} sessions at {args.c5_rate}/s", file=sys.stderr, flush=True)
        c5 = run_c5(nuif, W, args.c5_sessions, args.c5_rate, args.budget_gb * 2 ** 30)
        print(f"[bench] arrival trace done: {c5['frames']} frames, peak {c5['peak_concurrent']} sessions",
              file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        print(f"[bench] CPU baseline on {args.cpu_frames} frames", file=sys.stderr, flush=True)
        cpu = cpu_baseline(wl, args.cpu_frames)
        # the same host leg checks the server's session type against the FP64 oracle
        print(f"[bench] FP64 check of the FP16 session on {args.cpu_frames} frames", file=sys.stderr, flush=True)
        cpu["fp64_check"] = fp64_check(nuif, wl, args.cpu_frames, STORAGES["f16"])
        print(f"[bench] FP64 check of the joint-order FP16 session on {args.cpu_frames} frames", file=sys.stderr,
              flush=True)
        cpu["fp64_check_joint"] = fp64_check(nuif, wl, args.cpu_frames, STORAGES["f16"], joint=True)

    sch = (C.c_int32 * 4)()
    nuif.check(nuif.L().nuif_schedule(sch), "schedule")
    schedule = dict(zip(["ray_batch", "walk_batch", "rmw_batch", "seg_voxels"], list(sch)))
    if server and server.get("stages"):
        # the roofline of the round's dominant stage (the server is the headline value)
        roofs = server["stages"]["rooflines"]
        single_roof = roof
        roof = dict(max(roofs, key=lambda r: r["avg_ms"]))
        roof["rooflines"] = roofs
        roof["single_session_integrate"] = single_roof
    if server:
        value, steps_ms = server["fps"], server["round_ms"]
        workload = (f"{wl.name}: {server['sessions']} resident sessions, one fused frame each per round "
                    f"(preprocess + ICP 10/5/4 + integrate every frame + raycast), {args.server_storage} volumes")
    else:
        value, steps_ms = single["fps"], single["ms_per_frame"]
        workload = f"{wl.name}: 1 session, fused frame (preprocess + ICP 10/5/4 + integrate every frame + raycast)"
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(steps_ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded scene renderer), depth frames resident in HBM before the timed region: the "
                "per-frame host upload (614 KB at 640x480) is not in value",
        "config": {"workload": workload, "resolution": [wl.width, wl.height], "volume": [wl.volume_res] * 3,
                   "volume_m": wl.volume_dim, "sessions": server["sessions"] if server else 1},
        "schedule": schedule,
        "p50_frame_latency_ms": round(server["p50_ms"], 3) if server else None,
        "p99_frame_latency_ms": round(server["p99_ms"], 3) if server else None,
        "server": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in server.items()} if server else None,
        "single_session": {"storage": args.storage, "graph": bool(args.graph), "fps": round(single["fps"], 3),
                           "ms_per_frame": round(single["ms_per_frame"], 4), "tracking_ok": single["ok"],
                           "stage_ms": {k: round(v, 4) for k, v in zip(names, single["stage_avg"])}
                           if single["stage_avg"] else None} if single else None,
        "fp32_serial_fps": round(fp32["fps"], 3) if fp32 else None,
        "server_vs_fp32_serial": round(server["fps"] / fp32["fps"], 3) if server and fp32 else None,
        "fp32_serial_compensated_fps": round(fp32c["fps"], 3) if fp32c else None,
        "roofline": roof,
        "small_session": small,
        "hash_session": hashed,
        "joint_icp_mix": c4,
        "arrival_trace": c5,
        "cpu_baseline": cpu,
        "device": info["name"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
