"""GPU tests of the tracking engine's edge cases against the CPU oracle:

  * the integration gate at the reference's 0.15 (VO/NuiKinfuTrackingEngine.cpp:113-127):
    the per-frame integrate / skip pattern, poses and volume equal the oracle's;
  * a planar frame: the ICP normal equations lose three directions (rotation about the
    plane normal, translation within the plane), the system is singular and tracking
    fails (OCL/NuiKinfuOpenCLDepthTracker.cpp:465-472 for the reference's order, a non
    positive LDL^T pivot for the compensated one): status 1, pose kept, no integration;
  * frames without valid depth (all 0, or all beyond the far limit, so every pixel is
    NaN after the passing filter): no association, every level breaks at its first
    iteration, the frame tracks with the pose kept;
  * the pose history ring: with a capacity of a few rows the host copy still holds
    every frame's row in order.
"""
import sys

import numpy as np
import pytest

import oracle as O
from natureuserinterfacestudio_amd import nuif
from natureuserinterfacestudio_amd import workloads as W

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _dev(d):
    return nuif.DeviceArray.from_host(np.ascontiguousarray(d, np.uint16))


def _volume_equal(eng, ref, storage):
    if storage == nuif.TSDF_S16:
        raw = eng.volume_raw().view(np.int16).reshape(-1, 2)
        np.testing.assert_array_equal(raw, ref.volume_raw().reshape(-1, 2))
    else:
        np.testing.assert_array_equal(eng.volume_raw(), ref.volume_words())


def test_integration_gate_015_matches_oracle(gpu):
    announce("gate 0.15, 320x240 128^3, 16 frames vs float oracle")
    wl = W.SMALL
    n = 16
    eng = nuif.Engine(wl.kinfu_desc(gate=0.15))
    ref = O.OracleEngine(wl, prec=0, gate=0.15)
    flags, oposes = [], []
    for f in range(n):
        d = wl.render(f)[0]
        eng.run(_dev(d), sync=False)
        ok, pr, _ = ref.run(d)
        assert ok == 1
        flags.append(ref.integrated())
        oposes.append(pr.astype(np.float32))
    eng.synchronize()
    rows = eng.poses(n)
    assert rows.shape[0] == n
    # the reference's pattern at 2 deg and 4 cm a frame: some frames integrate, most skip
    assert 1 < sum(flags[1:]) < n - 1, flags
    np.testing.assert_array_equal(rows[:, 15] == 1.0, np.array(flags))
    np.testing.assert_array_equal(rows[:, 12], np.ones(n, np.float32))
    np.testing.assert_array_equal(rows[:, :12], np.array(oposes))
    _volume_equal(eng, ref, nuif.TSDF_S16)


@pytest.mark.parametrize("storage", [nuif.TSDF_S16, nuif.TSDF_F16], ids=["s16-reference", "f16-compensated"])
def test_planar_frame_fails(gpu, storage):
    announce(f"planar frames, storage {storage}")
    wl = W.SMALL
    plane = np.full((wl.height, wl.width), 1500, np.uint16)
    eng = nuif.Engine(wl.kinfu_desc(storage=storage))
    compensated = storage != nuif.TSDF_S16
    ref = O.OracleEngine(wl, prec=0, gate=0.0, compensated=compensated, storage=0 if storage == nuif.TSDF_S16 else 1)
    got = [eng.run(_dev(plane), sync=True) for _ in range(3)]
    want = [ref.run(plane)[0] for _ in range(3)]
    assert want == [1, 0, 0]
    assert got == want
    st = eng.state()
    assert st.status == 1
    rows = eng.poses(3)
    np.testing.assert_array_equal(rows[:, 12], np.array([1, 0, 0], np.float32))
    # the failed frames keep the camera pose (identity) and do not integrate
    for f in (1, 2):
        np.testing.assert_array_equal(rows[f, :12], rows[0, :12])
        assert rows[f, 15] == 0.0
    _volume_equal(eng, ref, storage)


@pytest.mark.parametrize("value", [0, 65535], ids=["zeros", "beyond-far"])
def test_frame_without_valid_depth(gpu, value):
    announce(f"frame of depth {value} between tracked frames")
    wl = W.SMALL
    bad = np.full((wl.height, wl.width), value, np.uint16)
    seq = [wl.render(0)[0], bad, wl.render(1)[0], wl.render(2)[0]]
    eng = nuif.Engine(wl.kinfu_desc())
    ref = O.OracleEngine(wl, prec=0, gate=0.0)
    oposes = []
    for d in seq:
        assert eng.run(_dev(d), sync=True) == 1
        ok, pr, _ = ref.run(d)
        assert ok == 1
        oposes.append(pr.astype(np.float32))
    rows = eng.poses(len(seq))
    np.testing.assert_array_equal(rows[:, :12], np.array(oposes))
    np.testing.assert_array_equal(rows[1, :12], rows[0, :12])  # nothing to track: pose kept
    assert rows[1, 14] == 0.0  # no pixel associated
    _volume_equal(eng, ref, nuif.TSDF_S16)


def test_pose_history_ring_wraps(gpu):
    announce("pose history: capacity 3 then 2 over 9 frames vs default capacity")
    wl = W.SMALL
    n = 9
    small = nuif.Engine(wl.kinfu_desc())
    small.set_history_capacity(3)
    full = nuif.Engine(wl.kinfu_desc())
    for f in range(n):
        if f == 5:
            small.set_history_capacity(2)  # mid-run: the rows so far are kept
        d = _dev(wl.render(f)[0])
        small.run(d, sync=False)
        full.run(d, sync=False)
    a, b = small.poses(n), full.poses(n)
    assert a.shape == (n, nuif.POSE_ROW)
    np.testing.assert_array_equal(a, b)
    with pytest.raises(nuif.NuifError):
        small.set_history_capacity(0)


def test_raycast_footprint_counts_the_voxels_read(gpu):
    announce("raycast footprint, 320x240 128^3 after 3 frames")
    wl = W.SMALL
    eng = nuif.Engine(wl.kinfu_desc())
    for f in range(3):
        eng.run(_dev(wl.render(f)[0]), sync=True)
    u1 = eng.raycast_footprint()
    u2 = eng.raycast_footprint()
    vg, _ = eng.maps()
    hits = int((~np.isnan(vg[:, 2])).sum())
    assert u1 == u2  # the same march at the same pose
    # every hit reads the 8 corners of at least 2 trilinear cells; the march reads at
    # most the whole volume
    assert hits > 10000 and 8 <= u1 < 128 ** 3
    assert u1 > hits // 8
