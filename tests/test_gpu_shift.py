"""GPU parity of the shifting dense volume's kernels (K/tsdfVolume.cl:71,104,192-211,
K/raycast.cl:62,184, K/extract.cl:80-190).

Bars, all bit for bit: integration into a wrapped volume equals the plain
integration rolled by the wrap (tile kernel when w.z is a multiple of the voxels
per 16-byte vector, column kernel otherwise), and marks (at least) the blocks of
the stored tsdf <= 0 voxels; the raycast of the wrapped volume equals the plain
raycast (vertex, normal and colour maps); the slab reset equals its definition;
the slab fetch equals the oracle's restatement of fetchShiftVolumeKernel."""
import ctypes as C
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


def _depth(wl, f):
    d, bgra, _ = wl.render(f, color=True)
    dep = np.empty(d.size, np.float32)
    O.lib().or_passing_filter_f(d.reshape(-1), dep, d.size, 0.5, 4.5)
    O.lib().or_scale_depths_f(dep, wl.width, wl.height, wl.cam6())
    return dep, bgra


def _roll(arr, N, w, ch):
    v = arr.reshape(N, N, N, ch)
    return np.roll(v, shift=(w[1], w[0], w[2]), axis=(0, 1, 2)).reshape(arr.shape)


def _blocks_of_stored(vol_i16, N):
    v = vol_i16.reshape(N, N, N, 2)
    neg = (v[..., 1] > 0) & (v[..., 0] <= 0)
    nb = (N + 7) // 8
    blk = neg.reshape(nb, 8, nb, 8, nb, 8).any(axis=(1, 3, 5)).astype(np.uint8).reshape(-1)
    ns = (N + 31) // 32
    sup = blk.reshape(nb, nb, nb).reshape(ns, 4, ns, 4, ns, 4).max(axis=(1, 3, 5)).reshape(-1)
    return np.concatenate([blk, sup])


@pytest.mark.parametrize("wrap", [(0, 0, 0), (5, 7, 12), (3, 0, 15), (127, 64, 1)])
def test_wrapped_integration_and_raycast_match_the_rolled_volume(gpu, wrap):
    wl = W.SMALL
    N = 128
    announce(f"wrapped S16 {N}^3 integrate x3 + raycast, wrap {wrap}")
    L = gpu.L()
    p = nuif.tsdf_params(N, wl.volume_dim, wl.trunc)
    cam = wl.camera_params()
    w = np.array(wrap, np.int32)
    nb = L.nuif_tsdf_surface_blocks_bytes(C.byref(p))
    vols, cols, surfs = [], [], []
    for ww in (None, w):
        dv = nuif.DeviceArray((N ** 3 * 4,), np.uint8)
        dc = nuif.DeviceArray((N ** 3 * 4,), np.uint8)
        ds = nuif.DeviceArray((nb,), np.uint8)
        nuif.check(L.nuif_tsdf_reset(dv.ptr, dc.ptr, C.byref(p), 0, None))
        nuif.check(L.nuif_tsdf_surface_blocks_clear(ds.ptr, C.byref(p), None))
        vols.append(dv), cols.append(dc), surfs.append(ds)
    pose = None
    for k, f in enumerate([0, 4, 9]):
        dep, bgra = _depth(wl, f)
        a = np.deg2rad(-2.0 * f)
        R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]], np.float32).T.copy()
        t = np.array([0.01 * k, -0.005 * k, 0.02 * k], np.float32)
        pose = nuif.upload_struct(nuif.rigid(R, t))
        ddep, dbg = nuif.DeviceArray.from_host(dep), nuif.DeviceArray.from_host(bgra)
        for dv, dc, ds, ww in zip(vols, cols, surfs, (None, w)):
            nuif.check(L.nuif_tsdf_integrate_wrapped(ddep.ptr, dbg.ptr, dv.ptr, dc.ptr, C.byref(p), C.byref(cam),
                                                     pose.ptr, 0, 128, ww.ctypes.data if ww is not None else None,
                                                     None, None, None, ds.ptr, None, None))
    plain, wrapped = vols[0].download().view(np.int16), vols[1].download().view(np.int16)
    assert (plain[1::2] > 0).sum() > 10000
    np.testing.assert_array_equal(wrapped, _roll(plain, N, wrap, 2))
    np.testing.assert_array_equal(cols[1].download(), _roll(cols[0].download(), N, wrap, 4))
    # flags are never cleared, so they cover (at least) the blocks of the stored tsdf <= 0 voxels
    for ds, vol in ((surfs[0], plain), (surfs[1], wrapped)):
        got, need = ds.download(), _blocks_of_stored(vol, N)
        assert need.sum() > 100 and np.all(got >= need)
    n = wl.width * wl.height
    maps = []
    for dv, dc, ds, ww in zip(vols, cols, surfs, (None, w)):
        v, nm, cm = (nuif.DeviceArray((n, 3), np.float32), nuif.DeviceArray((n, 3), np.float32),
                     nuif.DeviceArray((n, 4), np.uint8))
        nuif.check(L.nuif_tsdf_raycast_wrapped(dv.ptr, dc.ptr, C.byref(p), C.byref(cam), pose.ptr, v.ptr, nm.ptr,
                                               cm.ptr, 0, ww.ctypes.data if ww is not None else None, None, ds.ptr,
                                               None))
        maps.append((v.download(), nm.download(), cm.download()))
    assert np.isfinite(maps[0][0]).all(axis=1).sum() > 10000
    for a, b in zip(maps[0], maps[1]):
        np.testing.assert_array_equal(a.view(np.uint8), b.view(np.uint8))


def test_wrapped_f32_tile_path_matches_rolled(gpu):
    wl = W.SMALL
    N = 128
    wrap = (9, 2, 6)  # w.z even: the F32 tile kernel (2 voxels per vector)
    announce(f"wrapped F32 {N}^3 integrate, wrap {wrap}")
    L = gpu.L()
    p = nuif.tsdf_params(N, wl.volume_dim, wl.trunc)
    cam = wl.camera_params()
    w = np.array(wrap, np.int32)
    dvs = []
    for ww in (None, w):
        dv = nuif.DeviceArray((N ** 3 * 8,), np.uint8)
        nuif.check(L.nuif_tsdf_reset(dv.ptr, None, C.byref(p), 2, None))
        dvs.append(dv)
    dep, _ = _depth(wl, 2)
    ddep = nuif.DeviceArray.from_host(dep)
    pose = nuif.upload_struct(nuif.rigid(np.eye(3, dtype=np.float32), np.zeros(3, np.float32)))
    for dv, ww in zip(dvs, (None, w)):
        nuif.check(L.nuif_tsdf_integrate_wrapped(ddep.ptr, None, dv.ptr, None, C.byref(p), C.byref(cam), pose.ptr, 2,
                                                 128, ww.ctypes.data if ww is not None else None, None, None, None,
                                                 None, None, None))
    plain, wrapped = dvs[0].download().view(np.uint32), dvs[1].download().view(np.uint32)
    np.testing.assert_array_equal(wrapped, _roll(plain, N, wrap, 2))


def test_slab_reset_and_fetch_match_definition_and_oracle(gpu):
    wl = W.SMALL
    N = 128
    announce(f"slab reset and fetch on a {N}^3 integrated volume")
    L = gpu.L()
    p = nuif.tsdf_params(N, wl.volume_dim, wl.trunc)
    cam = wl.camera_params()
    dv = nuif.DeviceArray((N ** 3 * 4,), np.uint8)
    dc = nuif.DeviceArray((N ** 3 * 4,), np.uint8)
    nuif.check(L.nuif_tsdf_reset(dv.ptr, dc.ptr, C.byref(p), 0, None))
    dep, bgra = _depth(wl, 1)
    ddep, dbg = nuif.DeviceArray.from_host(dep), nuif.DeviceArray.from_host(bgra)
    pose = nuif.upload_struct(nuif.rigid(np.eye(3, dtype=np.float32), np.zeros(3, np.float32)))
    nuif.check(L.nuif_tsdf_integrate(ddep.ptr, dbg.ptr, dv.ptr, dc.ptr, C.byref(p), C.byref(cam), pose.ptr, 0, 128,
                                     None, None, None, None, None))
    words = dv.download().view(np.uint32)
    cv = dc.download()
    cell = [p.cell_size[k] for k in range(3)]
    scratch = nuif.DeviceArray((L.nuif_points_dense_scratch_bytes(C.byref(p)),), np.uint8)
    tot = nuif.DeviceArray((1,), np.uint32)
    cases = [((0, 0, 0), (0, 0, 0), (N, N, N)), ((0, 0, 0), (200, -37, 15), (N, N, N)),
             ((100, 0, 0), (115, 0, 0), (15, N, N)), ((0, 120, 0), (0, -8, 0), (N, 15, N)),
             ((0, 0, 70), (0, 0, -58), (N, N, 15)), ((0, 0, N), (0, 0, -N), (N, N, 15))]
    for wrap, off, rng in cases:
        ref = O.points_dense_slice(words, cv, 0, (N, N, N), cell, wrap, off, rng)
        wa, oa, ra = (np.array(x, np.int32) for x in (wrap, off, rng))
        nuif.check(L.nuif_points_dense_slice(dv.ptr, dc.ptr, C.byref(p), 0, wa.ctypes.data, oa.ctypes.data,
                                             ra.ctypes.data, None, None, 0, tot.ptr, scratch.ptr, None))
        n = int(tot.download()[0])
        assert n == ref[2], (wrap, off, rng)
        pts = nuif.DeviceArray((max(n, 1), 3), np.float32)
        cls = nuif.DeviceArray((max(n, 1), 4), np.float32)
        nuif.check(L.nuif_points_dense_slice(dv.ptr, dc.ptr, C.byref(p), 0, wa.ctypes.data, oa.ctypes.data,
                                             ra.ctypes.data, pts.ptr, cls.ptr, n, tot.ptr, scratch.ptr, None))
        np.testing.assert_array_equal(pts.download()[:n].view(np.uint32), ref[0].view(np.uint32))
        np.testing.assert_array_equal(cls.download()[:n].view(np.uint32), ref[1].view(np.uint32))
    # slab reset: stored voxels of the slab back to (32767, 0) and colour 0, the rest unchanged
    wrap, rng = np.array([120, 0, 0], np.int32), np.array([15, N, N], np.int32)
    nuif.check(L.nuif_tsdf_reset_slice(dv.ptr, dc.ptr, C.byref(p), 0, wrap.ctypes.data, rng.ctypes.data, None))
    exp_v = words.view(np.int16).reshape(N, N, N, 2).copy()
    exp_c = cv.reshape(N, N, N, 4).copy()
    xs = [(120 + i) % N for i in range(15)]
    exp_v[:, xs, :, 0], exp_v[:, xs, :, 1], exp_c[:, xs] = 32767, 0, 0
    np.testing.assert_array_equal(dv.download().view(np.int16).reshape(N, N, N, 2), exp_v)
    np.testing.assert_array_equal(dc.download().reshape(N, N, N, 4), exp_c)


@pytest.mark.parametrize("shift,color", [(3, False), (4, True)])
def test_shift_engine_matches_oracle_engine(gpu, shift, color):
    """RunTracking with the shifting volume (OCL/NuiKinfuOpenCLShiftScene.cpp:289-364): global
    poses, voxel offsets, the stored volume and the vertex cache bit for bit against the
    oracle engine over 12 frames that shift the volume several times."""
    wl = W.SMALL
    n = 12
    announce(f"shift engine 320x240 128^3 voxel_shift {shift} colour {color}, {n} frames vs oracle")
    eng = nuif.Engine(wl.kinfu_desc(shift=shift, color=color))
    ref = O.OracleEngine(wl, prec=0, gate=0.0, color=color)
    ref.set_shift(shift)
    shifted = False
    for f in range(n):
        d, bgra, _ = wl.render(f, color=True)
        ok = eng.run(d, bgra if color else None)
        rok, pr, _ = ref.run(d, bgra if color else None)
        assert ok == 1 and rok == 1, f
        off, size = eng.shift_state()
        roff, rsize = ref.offset()
        np.testing.assert_array_equal(off, roff)
        np.testing.assert_array_equal(size, rsize.astype(np.float32))
        shifted |= bool(np.any(off != 0))
        poses = eng.poses(f + 1)
        expect = pr.astype(np.float32)
        expect[9:12] = expect[9:12] + rsize.astype(np.float32)  # global = local + offset (float)
        np.testing.assert_array_equal(poses[f, :12], expect, err_msg=f"frame {f}")
    assert shifted
    np.testing.assert_array_equal(eng.volume_raw().view(np.int16), ref.volume_raw())
    gp, gc = eng.shift_cache()
    rp, rc = ref.cache()
    assert len(rp) > 100
    np.testing.assert_array_equal(gp.view(np.uint32), rp.view(np.uint32))
    np.testing.assert_array_equal(gc.view(np.uint32), rc.view(np.uint32))


@pytest.mark.parametrize("N,storage,wrap", [(256, 0, (0, 0, 0)), (384, 0, (5, 7, 16)), (320, 2, (3, 250, 2))])
def test_walk_state_workspace_matches_the_replayed_walk(gpu, N, storage, wrap):
    """The per-column walk states computed once per z segment, with the z table that lets
    the integration skip chunks on estimated states, give the same volume, colour volume
    and updated-voxel count as every segment replaying the whole walk."""
    wl = W.SMALL
    announce(f"walk-state workspace vs replay: storage {storage}, {N}^3, wrap {wrap}, 3 frames")
    L = gpu.L()
    p = nuif.tsdf_params(N, wl.volume_dim * N / 128, wl.trunc)
    cam = wl.camera_params()
    w = np.array(wrap, np.int32)
    ws_bytes = L.nuif_tsdf_walk_states_bytes(C.byref(p))
    assert ws_bytes == ((N + 63) // 64 - 1) * (N * N + 1) * 8 + N * 8  # 64-voxel segments, then the z table
    ws = nuif.DeviceArray((ws_bytes,), np.uint8)
    vb = N ** 3 * L.nuif_tsdf_voxel_bytes(storage)
    color = storage == 0
    out = []
    for walk in (None, ws.ptr):
        dv = nuif.DeviceArray((vb,), np.uint8)
        dc = nuif.DeviceArray((N ** 3 * 4,), np.uint8) if color else None
        cnt = nuif.DeviceArray.from_host(np.zeros(1, np.uint64))
        nuif.check(L.nuif_tsdf_reset(dv.ptr, dc.ptr if color else None, C.byref(p), storage, None))
        for k, f in enumerate([0, 5, 11]):
            dep, bgra = _depth(wl, f)
            a = np.deg2rad(-1.5 * f)
            R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]], np.float32).T.copy()
            t = np.array([0.03 * k, -0.01 * k, -0.6 + 0.02 * k], np.float32)
            pose = nuif.upload_struct(nuif.rigid(R, t))
            ddep, dbg = nuif.DeviceArray.from_host(dep), nuif.DeviceArray.from_host(bgra)
            nuif.check(L.nuif_tsdf_integrate_wrapped(ddep.ptr, dbg.ptr if color else None, dv.ptr,
                                                     dc.ptr if color else None, C.byref(p), C.byref(cam), pose.ptr,
                                                     storage, 128, w.ctypes.data, None, cnt.ptr, None, None, walk,
                                                     None))
        nuif.check(L.nuif_stream_synchronize(None))
        out.append((dv.download(), dc.download() if color else None, int(cnt.download()[0])))
    assert out[0][2] == out[1][2] > 20000
    np.testing.assert_array_equal(out[0][0], out[1][0])
    if color:
        np.testing.assert_array_equal(out[0][1], out[1][1])
