"""Schedule tuner of the dense integration and raycast.

The launch schedule (natureuserinterfacestudio_amd/csrc/schedule.h) sets how many raycast
samples, integration walk steps and integration column chunks each wave keeps in flight,
how many integration chunks a coarse cull test covers (compiled in) and how long a z
segment of the integration is (chosen per launch by the number of sessions sharing it).  The tuner builds one variant of the kernel library per
candidate of the compiled values and times each with bench.py on the GPU: the fused
server round of 32 sessions at 640x480 / 512^3 (FP16 volumes, per-stage device times)
and the single session (reference volume); then it forces each segment length at run
time for server rounds of 8 and 32 sessions and the single session.  The sweep table
goes to profiles/schedule_sweep_full_640x480_512.md; the best values are the defaults
and the segment rules of schedule.h.

    python -m natureuserinterfacestudio_amd.tune_schedules build   # here: variant libraries
    python -m natureuserinterfacestudio_amd.tune_schedules run     # on the GPU: sweep, table
    python -m natureuserinterfacestudio_amd.tune_schedules clean   # drop the variants
    python -m natureuserinterfacestudio_amd.tune_schedules xcd     # on the GPU: XCD numbering sweep

The XCD sweep forces each mask of XCD-aware block numbering (NUIF_XCD_REMAP: bit 0 ICP
tiles, bit 1 batched integration, bit 2 batched raycast) for the 32-session server round
and writes profiles/xcd_sweep_server_32x640x480_512.md.

The variants live in natureuserinterfacestudio_amd/lib/tune/<name>/ (all three
libraries, the kernel library rebuilt with the candidate's -D values); bench.py loads
them through NUIF_LIB_DIR.
"""
import json
import os
import shutil
import subprocess
import sys
import time

from . import _build

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
TUNE = os.path.join(PKG, "lib", "tune")
PROFILE = os.path.join(ROOT, "profiles", "schedule_sweep_full_640x480_512.md")

# candidate schedules: the defaults and one change at a time
CANDIDATES = {
    "default": {},
    "ray4": {"NUIF_RAY_BATCH": 4},
    "ray16": {"NUIF_RAY_BATCH": 16},
    "walk16": {"NUIF_WALK_BATCH": 16},
    "rmw4": {"NUIF_RMW_BATCH": 4},
    "cull2": {"NUIF_CULL_GROUP": 2},
    "cull8": {"NUIF_CULL_GROUP": 8},
    "cull16": {"NUIF_CULL_GROUP": 16},
}
# integration z-segment lengths, forced at run time (NUIF_SEG_VOXELS) for the single
# session and server rounds of SEG_SESSIONS sessions
SEGMENTS = (64, 128, 256, 512)
SEG_SESSIONS = (8, 32)
# XCD-aware numbering masks, forced at run time (NUIF_XCD_REMAP); each mask is run twice
# (interleaved) so that the spread between runs shows beside the differences
XCD_MASKS = (0, 1, 2, 4, 7)
XCD_PROFILE = os.path.join(ROOT, "profiles", "xcd_sweep_server_32x640x480_512.md")
XCD_ARGS = ["--server-only", "--steps", "20", "--warmup", "5", "--no-cpu", "--c5-sessions", "0", "--hash-steps", "0"]
# sources whose code depends on the schedule
SCHEDULED = ["tsdf.hip", "fusion.hip"]
BENCH_ARGS = ["--steps", "10", "--warmup", "3", "--no-cpu", "--c5-sessions", "0", "--hash-steps", "0"]


def build(names=None):
    """Variant libraries (CPU): the scheduled sources recompiled with the candidate's
    values and linked with the in-tree objects of everything else."""
    _build.build_all()
    csrc = os.path.join(PKG, "csrc")
    lib = os.path.join(PKG, "lib")
    others = [os.path.join(_build.OBJ, s.replace(".hip", ".o")) for s in _build.HIP_SOURCES if s not in SCHEDULED]
    for name in names or CANDIDATES:
        out = os.path.join(TUNE, name)
        os.makedirs(out, exist_ok=True)
        defs = [f"-D{k}={v}" for k, v in CANDIDATES[name].items()]
        objs = []
        for s in SCHEDULED:
            o = os.path.join(out, s.replace(".hip", ".o"))
            _build._run([_build.HIPCC, *_build.HIP_FLAGS, *defs, "-c", os.path.join(csrc, s), "-o", o])
            objs.append(o)
        _build._run([_build.HIPCC, f"--offload-arch={_build.ARCH}", "-shared", "-fPIC", "-o",
                     os.path.join(out, "libnuif.so"), *others, *objs])
        for o in objs:
            os.remove(o)
        for so in ("libnuif_kinfu.so", "libnuif_synth.so"):
            shutil.copy2(os.path.join(lib, so), os.path.join(out, so))
        print(f"[tune] built {name} {CANDIDATES[name]}", flush=True)


def _bench(env, args):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError(f"bench failed ({args}):\n{r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def run_segments():
    """The integration segment per launch size (GPU, the in-tree libraries)."""
    rows = []
    for sessions in SEG_SESSIONS:
        for seg in SEGMENTS:
            print(f"[tune] segment {seg}, {sessions} sessions", flush=True)
            b = _bench(dict(os.environ, NUIF_SEG_VOXELS=str(seg)),
                       ["--steps", "10", "--warmup", "3", "--no-cpu", "--c5-sessions", "0", "--hash-steps", "0",
                        "--sessions", str(sessions)])
            rows.append({"seg": seg, "sessions": sessions, "server_fps": b["value"],
                         "server_integrate_ms": b["server"]["stages"]["avg_ms"]["integrate"],
                         "single_integrate_ms": b["single_session"]["stage_ms"]["integrate"],
                         "fp32_serial_fps": b["fp32_serial_fps"]})
    return rows


def run(names=None):
    """The sweep (GPU): bench.py per variant and per segment length; writes the table."""
    rows = []
    for name in names or CANDIDATES:
        d = os.path.join(TUNE, name)
        if not os.path.exists(os.path.join(d, "libnuif.so")):
            raise FileNotFoundError(f"{d}: not built (run the build step first)")
        env = dict(os.environ, NUIF_LIB_DIR=d)
        print(f"[tune] {name}", flush=True)
        t0 = time.time()
        b = _bench(env, BENCH_ARGS)
        st = b["server"]["stages"]["avg_ms"]
        single = b["single_session"]["stage_ms"]
        rows.append({"name": name, "defines": CANDIDATES[name], "schedule": b.get("schedule"),
                     "server_fps": b["value"], "server_integrate_ms": st["integrate"],
                     "server_raycast_ms": st["raycast"], "single_integrate_ms": single["integrate"],
                     "single_raycast_ms": single["raycast"], "fp32_serial_fps": b["fp32_serial_fps"],
                     "seconds": round(time.time() - t0, 1)})
        print(f"[tune] {name}: {rows[-1]['server_fps']} frames/s", flush=True)
    segs = run_segments() if not names else []
    write_table(rows, segs)
    return rows


def write_table(rows, segs):
    best = max(rows, key=lambda r: r["server_fps"])
    lines = ["# Schedule sweep: 640x480 / 512^3", "",
             "Command: `python -m natureuserinterfacestudio_amd.tune_schedules build` (CPU), then on the GPU "
             "`python -m natureuserinterfacestudio_amd.tune_schedules run`; each variant runs "
             f"`bench.py {' '.join(BENCH_ARGS)}` with its libraries (NUIF_LIB_DIR).", "",
             "Server: 32 sessions, FP16 volumes, per-stage device time of the batched round (events on the "
             "server stream).  Single: one session, reference (short2) volume, stage events on its stream.", "",
             "| variant | -D values | server frames/s | server integrate ms | server raycast ms | single integrate ms "
             "| single raycast ms | FP32 serial frames/s |",
             "|---|---|---|---|---|---|---|---|"]
    for r in rows:
        defs = " ".join(f"{k}={v}" for k, v in r["defines"].items()) or "(defaults)"
        lines.append(f"| {r['name']} | {defs} | {r['server_fps']:.1f} | {r['server_integrate_ms']:.4f} | "
                     f"{r['server_raycast_ms']:.4f} | {r['single_integrate_ms']:.4f} | {r['single_raycast_ms']:.4f} "
                     f"| {r['fp32_serial_fps']:.1f} |")
    lines += ["", f"Best server round: **{best['name']}** ({best['server_fps']:.1f} frames/s)."]
    if segs:
        lines += ["", "## Integration z segment (NUIF_SEG_VOXELS), per launch size", "",
                  "| segment voxels | server sessions | server frames/s | server integrate ms | single integrate ms "
                  "| FP32 serial frames/s |", "|---|---|---|---|---|---|"]
        for r in segs:
            lines.append(f"| {r['seg']} | {r['sessions']} | {r['server_fps']:.1f} | {r['server_integrate_ms']:.4f} | "
                         f"{r['single_integrate_ms']:.4f} | {r['fp32_serial_fps']:.1f} |")
        single = min(segs, key=lambda r: r["single_integrate_ms"])
        lines += [""]
        for n in SEG_SESSIONS:
            b = min((r for r in segs if r["sessions"] == n), key=lambda r: r["server_integrate_ms"])
            lines.append(f"Best for {n} sessions: {b['seg']} voxels.")
        lines.append(f"Best for a single session: {single['seg']} voxels.")
    os.makedirs(os.path.dirname(PROFILE), exist_ok=True)
    with open(PROFILE, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"[tune] table written to {PROFILE}; best: {best['name']}", flush=True)


def run_xcd():
    """XCD-aware block numbering per stage (GPU, the in-tree libraries): the server round's
    stage times and frames/s for each mask, two interleaved passes."""
    rows = []
    for rep in range(2):
        for mask in XCD_MASKS:
            print(f"[tune] xcd mask {mask} (pass {rep + 1})", flush=True)
            b = _bench(dict(os.environ, NUIF_XCD_REMAP=str(mask)), XCD_ARGS)
            st = b["server"]["stages"]["avg_ms"]
            rows.append({"mask": mask, "pass": rep + 1, "fps": b["value"], "p99": b["p99_frame_latency_ms"],
                         "icp": st["icp"], "integrate": st["integrate"], "raycast": st["raycast"]})
    names = {1: "ICP tiles", 2: "integration", 4: "raycast"}
    lines = ["# XCD-aware block numbering: 32-session server round, 640x480 / 512^3", "",
             "Command (on the GPU): `python -m natureuserinterfacestudio_amd.tune_schedules xcd`; each mask runs "
             f"`bench.py {' '.join(XCD_ARGS)}` with NUIF_XCD_REMAP set, every mask twice in interleaved passes.", "",
             "Mask bits: 1 the ICP tile runs, 2 the batched integration, 4 the batched raycast (consecutive blocks "
             "of a session on one XCD); 0 the dispatcher's own round-robin order.  Stage times are device times of "
             "the batched round (events on the server stream).", "",
             "| mask | stages renumbered | pass | frames/s | p99 ms | ICP ms | integration ms | raycast ms |",
             "|---|---|---|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: (r["mask"], r["pass"])):
        on = ", ".join(v for k, v in names.items() if r["mask"] & k) or "none"
        lines.append(f"| {r['mask']} | {on} | {r['pass']} | {r['fps']:.1f} | {r['p99']:.3f} | {r['icp']:.4f} | "
                     f"{r['integrate']:.4f} | {r['raycast']:.4f} |")
    lines.append("")
    for stage, bit in (("icp", 1), ("integrate", 2), ("raycast", 4)):
        on = [r[stage] for r in rows if r["mask"] & bit]
        off = [r[stage] for r in rows if not r["mask"] & bit]
        lines.append(f"{stage}: renumbered {min(on):.4f}-{max(on):.4f} ms, dispatcher order "
                     f"{min(off):.4f}-{max(off):.4f} ms.")
    with open(XCD_PROFILE, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"[tune] table written to {XCD_PROFILE}", flush=True)
    return rows


def clean():
    shutil.rmtree(TUNE, ignore_errors=True)


def main(argv):
    if not argv or argv[0] not in ("build", "run", "clean", "xcd"):
        print(__doc__)
        return 2
    if argv[0] in ("clean", "xcd"):
        {"clean": clean, "xcd": run_xcd}[argv[0]]()
        return 0
    names = argv[1:] or None
    {"build": build, "run": run}[argv[0]](names)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
