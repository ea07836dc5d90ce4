"""The voxel-hashing session at its configured size: 640x480 depth of the 8 x 8 x 3 m
room with the seeded random-walk camera, 5 mm voxels in 8^3 bricks, a table of 2^20
buckets of 4 entries (OCL/NuiKinfuOpenCLHashScene.cpp:404-583 per frame,
OCL/NuiKinfuVoxelBlockHash.cpp:24-213 for the table), 30 frames through RunTracking
against the oracle's hash engine (float arithmetic of the reference kernels,
deterministic allocation).

Bars: every frame's pose bit for bit; after the last frame the allocated blocks, voxel
words, the reference maps (castRay) and the mesh (points and triangle indices of
Volume2Mesh) equal the oracle's; no insertion failed."""
import sys

import numpy as np
import pytest

import oracle as O
from natureuserinterfacestudio_amd import nuif
from natureuserinterfacestudio_amd import workloads as W

pytestmark = pytest.mark.gpu

WL = W.HASH
FRAMES = 30
BUCKETS = 1 << 20
BRICKS = 1 << 18


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def test_hash_session_full_size_matches_oracle_engine(gpu):
    announce(f"hash engine {WL.width}x{WL.height} 5 mm, 2^20 buckets, {FRAMES} frames vs oracle hash engine")
    desc = WL.kinfu_desc(bricks=BRICKS, buckets=BUCKETS)
    hp = nuif.hash_params_of_desc(desc)
    assert hp.buckets_log2 == 20
    depths = [WL.render(f)[0] for f in range(FRAMES)]
    eng = nuif.Engine(desc)
    for d in depths:
        eng.run(nuif.DeviceArray.from_host(d), sync=False)
    eng.synchronize()
    poses = eng.poses(FRAMES)
    stats = eng.hash_stats()
    assert stats[2] == 0 and eng.scene_failures() == 0
    ref = O.OracleHashEngine(WL, capacity=BRICKS)
    for f, d in enumerate(depths):
        ok, pr, _ = ref.run(d)
        assert ok == 1 and poses[f, 12] == 1.0, f
        np.testing.assert_array_equal(poses[f, :12], pr.astype(np.float32), err_msg=f"frame {f}")
    keys, _, w, words, _ = nuif.decode_hash_blocks(eng.volume_raw(), hp)
    okeys, osdf, ow, _ = ref.blocks()
    assert len(keys) == stats[0] and len(keys) > 20000
    np.testing.assert_array_equal(keys, okeys)
    np.testing.assert_array_equal((words & 0xFFFF).astype(np.uint16).view(np.int16), osdf)
    np.testing.assert_array_equal(w, ow.astype(np.int32))
    v, nm = eng.maps()
    ov, on = ref.refmaps()
    np.testing.assert_array_equal(v, ov)
    np.testing.assert_array_equal(nm, on)
    pts, idx = eng.mesh_shape()  # the visible list's triangles, indices 0, 1, 2, ...
    np.testing.assert_array_equal(idx, np.arange(len(idx), dtype=np.int32))
    tris = pts[idx].reshape(-1, 9)
    otris = ref.mesh().reshape(-1, 9)
    assert len(tris) == len(otris) > 10000

    def rows(a):  # the triangles as a sorted multiset of 36-byte rows
        return np.sort(np.ascontiguousarray(a).view(np.dtype((np.void, 36))).ravel())

    assert np.array_equal(rows(tris), rows(otris))
    announce(f"hash engine full size: {len(keys)} blocks, {len(tris)} triangles; poses / blocks / maps / mesh exact")
