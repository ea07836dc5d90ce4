"""Tracking failures in the fusion server: every frame's outcome after a round, and the
per-session policies that replace the reference's tracking manager (reset the engine,
or stop: VO/NuiKinfuTrackingManager.cpp:94-106).  Session 1 sees a plane twice, which
makes its second frame fail (singular ICP system); sessions 0 and 2 track a moving
scene.  Batched rounds (the outcome written by the ICP end into pinned memory) and
per-session rounds (the outcome read from the session's state) are both covered."""
import sys

import numpy as np
import pytest

from natureuserinterfacestudio_amd import nuif
from natureuserinterfacestudio_amd import workloads as W

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _dev(d):
    return nuif.DeviceArray.from_host(np.ascontiguousarray(d, np.uint16))


def _setup(policy, batching):
    wl = W.SMALL
    srv = nuif.Server(8 << 30)
    srv.set_batching(batching)
    srv.set_failure_policy(policy)
    ids = [srv.open(wl.kinfu_desc()) for _ in range(3)]
    assert None not in ids
    scene = [_dev(wl.render(f)[0]) for f in range(4)]
    plane = _dev(np.full((wl.height, wl.width), 1500, np.uint16))
    return wl, srv, ids, scene, plane


def _round(srv, ids, frames):
    srv.submit(ids, frames)
    srv.wait()
    return list(srv.last_outcomes())


@pytest.mark.parametrize("batching", [True, False], ids=["batched", "per-session"])
def test_keep_policy_reports_failures_and_continues(gpu, batching):
    announce(f"server failures: keep, batching={batching}")
    wl, srv, ids, scene, plane = _setup(nuif.ON_FAILURE_KEEP, batching)
    assert _round(srv, ids, [scene[0], plane, scene[0]]) == [1, 1, 1]
    assert _round(srv, ids, [scene[1], plane, scene[1]]) == [1, 0, 1]
    if batching:
        assert srv.last_batch == 3
    assert _round(srv, ids, [scene[2], plane, scene[2]]) == [1, 0, 1]
    rows = srv.sessions[ids[1]].poses(8)
    np.testing.assert_array_equal(rows[:, 12], np.array([1, 0, 0], np.float32))
    # the failed session's frames did not disturb the others: they match a lone engine
    lone = nuif.Engine(wl.kinfu_desc())
    for f in range(3):
        lone.run(scene[f], sync=True)
    np.testing.assert_array_equal(srv.sessions[ids[0]].poses(8), lone.poses(8))
    np.testing.assert_array_equal(srv.sessions[ids[2]].volume_raw(), lone.volume_raw())
    srv.close()


@pytest.mark.parametrize("batching", [True, False], ids=["batched", "per-session"])
def test_reset_policy_starts_the_session_over(gpu, batching):
    announce(f"server failures: reset, batching={batching}")
    wl, srv, ids, scene, plane = _setup(nuif.ON_FAILURE_RESET, batching)
    _round(srv, ids, [scene[0], plane, scene[0]])
    assert _round(srv, ids, [scene[1], plane, scene[1]]) == [1, 0, 1]
    # session 1 starts over: its next frame is a first frame of an empty scene
    assert _round(srv, ids, [scene[2], scene[0], scene[2]]) == [1, 1, 1]
    assert _round(srv, ids, [scene[3], scene[1], scene[3]]) == [1, 1, 1]
    fresh = nuif.Engine(wl.kinfu_desc())
    fresh.run(scene[0], sync=True)
    fresh.run(scene[1], sync=True)
    np.testing.assert_array_equal(srv.sessions[ids[1]].poses(8), fresh.poses(8))
    np.testing.assert_array_equal(srv.sessions[ids[1]].volume_raw(), fresh.volume_raw())
    srv.close()


def test_evict_policy_closes_the_session(gpu):
    announce("server failures: evict")
    wl, srv, ids, scene, plane = _setup(nuif.ON_FAILURE_EVICT, True)
    _round(srv, ids, [scene[0], plane, scene[0]])
    before = srv.bytes_in_use
    assert _round(srv, ids, [scene[1], plane, scene[1]]) == [1, 0, 1]
    assert ids[1] not in srv.sessions
    assert srv.bytes_in_use < before
    with pytest.raises(nuif.NuifError):
        srv.submit([ids[1]], [scene[2]])
    assert _round(srv, [ids[0], ids[2]], [scene[2], scene[2]]) == [1, 1]
    srv.close()


def test_scene_failures_are_reported(gpu):
    announce("hash scene: dropped insertions are counted")
    wl = W.HASH_SMALL
    eng = nuif.Engine(wl.kinfu_desc(bricks=64))  # far too few bricks for the room
    eng.run(_dev(wl.render(0)[0]), sync=True)
    assert eng.scene_failures() > 0
    dense = nuif.Engine(W.SMALL.kinfu_desc())
    dense.run(_dev(W.SMALL.render(0)[0]), sync=True)
    assert dense.scene_failures() == 0
