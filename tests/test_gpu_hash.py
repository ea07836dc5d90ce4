"""GPU parity of the voxel-hashing scene against the CPU oracle (float
arithmetic of the reference kernels, deterministic allocation).

Bars: allocated and visible block sets exact; voxel words (int16 sdf * 32767,
weight) and colours exact; ray ranges exact; raycast vertex / normal maps bit
for bit; marching-cubes vertex count exact and the triangles equal as a
multiset (the emission order follows the table order of the entries).  FP16
bricks: sdf within 1e-3 of the int16 oracle, weights exact.  Forward render: vertex,
normal and colour maps bit for bit in both projection modes."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O
from natureuserinterfacestudio_amd import nuif
from natureuserinterfacestudio_amd import workloads as W

pytestmark = pytest.mark.gpu

WL = W.HASH_SMALL
FRAMES = [0, 1, 2]


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _depth(f):
    d, bgra, gt = WL.render(f, color=True)
    dep = np.empty(d.size, np.float32)
    O.lib().or_passing_filter_f(d.reshape(-1), dep, d.size, 0.5, 4.5)
    return dep, bgra, gt[:, :3].astype(np.float32).copy(), gt[:, 3].astype(np.float32).copy()


def _build(storage=0, color=False, frames=FRAMES):
    p = nuif.hash_params(buckets_log2=16, bricks=1 << 17, storage=storage, color=color)
    g = nuif.HashScene(p)
    o = O.OracleHashScene(capacity=1 << 17)
    cam = WL.camera_params()
    for f in frames:
        dep, bgra, R, t = _depth(f)
        dd = nuif.DeviceArray.from_host(dep)
        db = nuif.DeviceArray.from_host(bgra) if color else None
        pose = nuif.upload_struct(nuif.rigid(R, t))
        g.allocate(dd, cam, pose)
        g.integrate(dd, cam, pose, bgra=db)
        o.integrate(dep, WL, R, t, bgra=bgra if color else None)
    return g, o


def test_hash_allocation_visible_and_voxels_match_oracle(gpu):
    announce("hash 320x240 5 mm: 3 frames, blocks / visible list / voxels vs oracle")
    g, o = _build(color=True)
    st = g.stats()
    assert st["failed"] == 0
    keys, sdf, w, words, rgb = g.blocks()
    okeys, osdf, ow, orgb = o.blocks()
    assert len(okeys) > 5000
    np.testing.assert_array_equal(keys, okeys)  # allocated set: exact
    assert st["bricks"] == len(okeys)
    np.testing.assert_array_equal(g.visible_keys(), o.visible())  # visible set: exact
    raw_sdf = (words & 0xFFFF).astype(np.uint16).view(np.int16)
    np.testing.assert_array_equal(raw_sdf, osdf)
    np.testing.assert_array_equal(w, ow.astype(np.int32))
    np.testing.assert_array_equal(rgb, orgb)
    assert (w > 0).sum() > 100000


def test_hash_raycast_matches_oracle(gpu):
    announce("hash raycast after 3 frames at the next pose vs oracle")
    g, o = _build()
    _, _, R, t = _depth(3)
    pose = nuif.upload_struct(nuif.rigid(R, t))
    v, n, rg, _ = g.raycast(WL.camera_params(), pose)
    ov, on, org, _ = o.raycast(WL, R, t)
    np.testing.assert_array_equal(rg, org)
    hit = ~np.isnan(ov[:, 0])
    assert hit.mean() > 0.5
    np.testing.assert_array_equal(np.isnan(v[:, 0]), ~hit)
    np.testing.assert_array_equal(v[hit], ov[hit])
    np.testing.assert_array_equal(n[hit], on[hit])


@pytest.mark.parametrize("reference_transform", [False, True])
def test_hash_forward_render_matches_oracle(gpu, reference_transform):
    """Forward render of the frame-2 castRay maps (vertices and colours) at the frame-3
    pose: ranges, vertices, normals and colours equal the oracle's bit for bit."""
    announce(f"hash forward render (reference transform {reference_transform}) vs oracle")
    g, o = _build(color=True)
    _, _, R2, t2 = _depth(2)
    prev_v, _, _, prev_c = o.raycast(WL, R2, t2, color=True)
    _, _, R3, t3 = _depth(3)
    pose = nuif.upload_struct(nuif.rigid(R3, t3))
    v, n, rg, c = g.forward_render(WL.camera_params(), pose, prev_v, prev_c, reference_transform)
    ov, on, org, oc, k = o.forward_render(WL, R3, t3, prev_v, prev_c, reference_transform)
    assert k > 100
    np.testing.assert_array_equal(rg, org)
    np.testing.assert_array_equal(np.isnan(v), np.isnan(ov))
    hit = ~np.isnan(ov[:, 0])
    assert hit.mean() > 0.5
    np.testing.assert_array_equal(v[hit], ov[hit])
    np.testing.assert_array_equal(np.isnan(n), np.isnan(on))
    nh = ~np.isnan(on[:, 0])
    np.testing.assert_array_equal(n[nh], on[nh])
    np.testing.assert_array_equal(c, oc)


def test_hash_mesh_matches_oracle(gpu):
    announce("hash marching cubes over the visible list vs oracle")
    g, o = _build(frames=[0, 1])
    ref = o.mesh()
    verts, _, total = g.mesh()
    assert total == len(ref) and total > 3000
    def tris(a):
        t = a.reshape(-1, 9)
        return t[np.lexsort(t.T[::-1])]
    np.testing.assert_array_equal(tris(verts), tris(ref))


def test_hash_f16_bricks_close_to_oracle(gpu):
    announce("hash FP16 bricks vs int16 oracle")
    g, o = _build(storage=1)
    keys, sdf, w, _, _ = g.blocks()
    okeys, osdf, ow, _ = o.blocks()
    np.testing.assert_array_equal(keys, okeys)
    np.testing.assert_array_equal(w, ow.astype(np.int32))
    m = ow > 0
    assert np.abs(sdf[m] - osdf[m] / np.float32(32767.0)).max() <= 1e-3


@pytest.mark.parametrize("forward", [nuif.RENDER_CASTRAY, nuif.RENDER_FORWARD, nuif.RENDER_FORWARD_REFERENCE])
def test_hash_engine_matches_oracle_engine(gpu, forward):
    """RunTracking with the hash scene: preprocessing, ICP against the rendered maps
    (castRay, or the forward render in either projection), allocation and integration --
    poses and every brick equal the oracle's."""
    announce(f"hash engine 320x240, 5 frames, render mode {forward} vs oracle hash engine")
    n = 5
    desc = WL.kinfu_desc(bricks=1 << 17, forward=forward)
    eng = nuif.Engine(desc)
    for f in range(n):
        eng.run(nuif.DeviceArray.from_host(WL.render(f)[0]), sync=False)
    eng.synchronize()
    poses = eng.poses(n)
    ref2 = O.OracleHashEngine(WL, capacity=1 << 17, forward=forward)
    for f in range(n):
        ok, pr, _ = ref2.run(WL.render(f)[0])
        assert ok == 1 and poses[f, 12] == 1.0
        np.testing.assert_array_equal(poses[f, :12], pr.astype(np.float32))
    keys, _, w, words, _ = nuif.decode_hash_blocks(eng.volume_raw(), nuif.hash_params_of_desc(desc))
    okeys, osdf, ow, _ = ref2.blocks()
    np.testing.assert_array_equal(keys, okeys)
    np.testing.assert_array_equal((words & 0xFFFF).astype(np.uint16).view(np.int16), osdf)
    np.testing.assert_array_equal(w, ow.astype(np.int32))
    v, nm = eng.maps()
    ov, on = ref2.refmaps()
    np.testing.assert_array_equal(v, ov)
    np.testing.assert_array_equal(nm, on)
