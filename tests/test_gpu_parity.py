"""GPU parity of the HIP kernels against the CPU oracle (float "compat"
flavour, which follows the reference's float arithmetic operation by
operation) and, for end-to-end tracking, against the FP64 oracle.

Tolerances (stated per test):
  * reference-precision path (short2 volume, reference summation order):
    bit for bit -- filtered depth, vertices, normals, ICP sums, poses, raw
    int16 voxels and raycast maps equal the float oracle's;
  * poses against the FP64 oracle: <= 0.1 mm / 0.01 deg per frame, or no
    further than the float arithmetic of the reference itself is from FP64;
  * FP16 / FP32 volumes: tsdf within 1e-3 (FP16) / 2e-4 (FP32) of the short2
    oracle, weights exact.
"""
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


def ulp_close(a, b, ulps=4):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    nan_a, nan_b = np.isnan(a), np.isnan(b)
    assert np.array_equal(nan_a, nan_b), f"NaN patterns differ at {np.argwhere(nan_a != nan_b)[:5]}"
    m = ~nan_a
    tol = ulps * np.spacing(np.maximum(np.abs(a[m]), np.abs(b[m])))
    bad = np.abs(a[m] - b[m]) > tol
    assert not bad.any(), f"{bad.sum()} values beyond {ulps} ulp; max diff {np.abs(a[m] - b[m]).max()}"


def rot_y(deg):
    a = np.radians(deg)
    return np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]], np.float32)


def pose_err(Ra, ta, Rb, tb):
    dt = np.linalg.norm(np.asarray(ta, np.float64) - np.asarray(tb, np.float64))
    Rr = np.asarray(Ra, np.float64).T @ np.asarray(Rb, np.float64)
    # atan2 of the skew part and the trace: well conditioned near zero angle
    w = 0.5 * np.array([Rr[2, 1] - Rr[1, 2], Rr[0, 2] - Rr[2, 0], Rr[1, 0] - Rr[0, 1]])
    ang = np.degrees(np.arctan2(np.linalg.norm(w), (np.trace(Rr) - 1) / 2))
    return dt, ang


# ------------------------------------------------------------------ runtime

def test_malloc_fill_complete_before_stream_work(gpu):
    """nuif_malloc's zero fill has finished when it returns: a fill issued at once
    on a non-blocking library stream is not overwritten by it (512 MiB)."""
    announce("nuif_malloc 512 MiB + fill on a new stream")
    L = gpu.L()
    n = 1 << 27
    s = C.c_void_p()
    nuif.check(L.nuif_stream_create(C.byref(s)))
    try:
        d = nuif.DeviceArray((n,), np.uint32)
        nuif.check(L.nuif_fill_u32(d.ptr, 0xA5A5A5A5, n, s))
        nuif.check(L.nuif_stream_synchronize(s))
        out = d.download()
        assert (out == np.uint32(0xA5A5A5A5)).all(), int((out != np.uint32(0xA5A5A5A5)).sum())
        d.free()
    finally:
        L.nuif_stream_destroy(s)


# ------------------------------------------------------------------ preprocessing

def test_passing_filter_bit_exact(gpu):
    announce("passing filter 320x240 frame 3")
    d, _, _ = W.SMALL.render(3)
    raw = nuif.DeviceArray.from_host(d)
    out = nuif.DeviceArray(d.shape, np.float32)
    nuif.check(gpu.L().nuif_depth_passing_filter(raw.ptr, out.ptr, d.size, 0.5, 4.5, None))
    ref = np.empty(d.size, np.float32)
    O.lib().or_passing_filter_f(d.reshape(-1), ref, d.size, 0.5, 4.5)
    np.testing.assert_array_equal(out.download().reshape(-1), ref)


def test_bilateral_matches_oracle(gpu):
    announce("bilateral 320x240 frame 5, radius 3")
    d, _, _ = W.SMALL.render(5)
    dep = np.empty(d.size, np.float32)
    O.lib().or_passing_filter_f(d.reshape(-1), dep, d.size, 0.5, 4.5)
    g = np.empty(7, np.float32)
    O.lib().or_gaussian_table_f(3, 0.08, g)
    ref = np.empty_like(dep)
    O.lib().or_bilateral_f(dep, ref, 320, 240, g, 3, 50.0, 0.3)
    din = nuif.DeviceArray.from_host(dep)
    dout = nuif.DeviceArray(dep.shape, np.float32)
    nuif.check(gpu.L().nuif_bilateral_filter_depth(din.ptr, dout.ptr, 320, 240, g.ctypes.data, 3, 50.0, 0.3, None))
    # bit-exact: both sides evaluate the range weight with the reproducible exp
    np.testing.assert_array_equal(dout.download(), ref)


def test_fused_preprocess_matches_oracle_chain(gpu):
    for wl in (W.SMALL, W.FULL):
        announce(f"fused preprocess {wl.name}")
        Wd, Hd = wl.width, wl.height
        d, _, _ = wl.render(7)
        cam = wl.camera_params()
        g = np.empty(7, np.float32)
        O.lib().or_gaussian_table_f(3, 0.08, g)
        raw = nuif.DeviceArray.from_host(d)
        dout = nuif.DeviceArray((Hd, Wd), np.float32)
        dp = [nuif.DeviceArray(((Hd >> l) * (Wd >> l),), np.float32) for l in range(3)]
        vp = [nuif.DeviceArray(((Hd >> l) * (Wd >> l), 3), np.float32) for l in range(3)]
        dptr = (C.c_void_p * 3)(*[a.ptr for a in dp])
        vptr = (C.c_void_p * 3)(*[a.ptr for a in vp])
        nuif.check(gpu.L().nuif_frame_preprocess(raw.ptr, dout.ptr, dptr, vptr, 3, Wd, Hd, C.byref(cam),
                                                 g.ctypes.data, 3, 50.0, 0.3, 1, None))
        dep = np.empty(d.size, np.float32)
        O.lib().or_passing_filter_f(d.reshape(-1), dep, d.size, 0.5, 4.5)
        filt = np.empty_like(dep)
        O.lib().or_bilateral_f(dep, filt, Wd, Hd, g, 3, 50.0, 0.3)
        scaled = dep.copy()
        O.lib().or_scale_depths_f(scaled, Wd, Hd, wl.cam6())
        np.testing.assert_array_equal(dout.download().reshape(-1), scaled)
        gdep = [a.download() for a in dp]
        np.testing.assert_array_equal(gdep[0], filt)
        # downstream levels are checked from the GPU's own level-0 output, so the
        # comparison isolates half_sample + depth2vertex
        src = gdep[0]
        for l in range(3):
            wl_, hl = Wd >> l, Hd >> l
            if l > 0:
                h = np.empty(wl_ * hl, np.float32)
                O.lib().or_half_sample_f(np.ascontiguousarray(gdep[l - 1]), h, wl_, hl, 1, 0.3)
                ulp_close(gdep[l], h, ulps=1)
                src = gdep[l]
            v = np.empty(wl_ * hl * 3, np.float32)
            O.lib().or_depth2vertex_f(np.ascontiguousarray(src), v, wl_, hl, wl.cam6(), Wd, Hd, 1 << l)
            np.testing.assert_array_equal(vp[l].download().reshape(-1), v)


def test_normals_and_transform_bit_exact(gpu):
    announce("normals + transform 640x480 frame 2")
    wl = W.FULL
    d, _, _ = wl.render(2)
    dep = np.empty(d.size, np.float32)
    O.lib().or_passing_filter_f(d.reshape(-1), dep, d.size, 0.5, 4.5)
    v = np.empty(d.size * 3, np.float32)
    O.lib().or_depth2vertex_f(dep, v, wl.width, wl.height, wl.cam6(), wl.width, wl.height, 1)
    nref = np.empty_like(v)
    O.lib().or_normals_simple_f(v, nref, wl.width, wl.height, 0.3)
    dv = nuif.DeviceArray.from_host(v)
    dn = nuif.DeviceArray(v.shape, np.float32)
    nuif.check(gpu.L().nuif_estimate_normals_simple(dv.ptr, dn.ptr, wl.width, wl.height, 0.3, None))
    np.testing.assert_array_equal(dn.download(), nref)
    R = rot_y(12.0)
    t = np.array([0.1, -0.05, 0.2], np.float32)
    pose = nuif.upload_struct(nuif.rigid(R, t))
    vd = nuif.DeviceArray(v.shape, np.float32)
    nd = nuif.DeviceArray(v.shape, np.float32)
    nuif.check(gpu.L().nuif_transform_maps(dv.ptr, dn.ptr, vd.ptr, nd.ptr, wl.width, wl.height, pose.ptr, None))
    vr, nr = np.empty_like(v), np.empty_like(v)
    O.lib().or_transform_maps_f(v.reshape(-1, 3).copy(), nref.reshape(-1, 3).copy(), vr.reshape(-1, 3),
                                nr.reshape(-1, 3), d.size, R.reshape(-1).copy(), t)
    np.testing.assert_array_equal(vd.download(), vr)
    np.testing.assert_array_equal(nd.download(), nr)


# ------------------------------------------------------------------ ICP

def _oracle_state(wl, frames):
    """Oracle engine run over `frames` frames; returns it and its reference maps."""
    e = O.OracleEngine(wl, prec=0, gate=0.0)
    for f in range(frames):
        d, _, _ = wl.render(f)
        e.run(d)
    return e


def test_icp_iteration_sums(gpu):
    announce("ICP sums, 320x240, frame 2 vs oracle maps of frame 1, all levels")
    wl = W.SMALL
    e = _oracle_state(wl, 2)
    vref, nref = e.refmaps()
    d, _, gt = wl.render(2)
    # current vertex pyramid from the oracle chain
    dep = np.empty(d.size, np.float32)
    O.lib().or_passing_filter_f(d.reshape(-1), dep, d.size, 0.5, 4.5)
    g = np.empty(7, np.float32)
    O.lib().or_gaussian_table_f(3, 0.08, g)
    filt = np.empty_like(dep)
    O.lib().or_bilateral_f(dep, filt, 320, 240, g, 3, 50.0, 0.3)
    R = gt[:, :3].astype(np.float32).copy()
    t = gt[:, 3].astype(np.float32).copy()
    prevR = rot_y(-2.0)
    prevt = np.array([0.03, 0.0, 0.001], np.float32)
    cam = wl.camera_params()
    dvr, dnr = nuif.DeviceArray.from_host(vref), nuif.DeviceArray.from_host(nref)
    prev = nuif.upload_struct(nuif.rigid(prevR, prevt))
    st = nuif.DeviceArray((gpu.L().nuif_icp_state_bytes(),), np.uint8)
    part = nuif.DeviceArray((gpu.L().nuif_icp_partials_bytes(320, 240),), np.uint8)
    src = filt
    for level in range(3):
        wl_, hl = 320 >> level, 240 >> level
        if level > 0:
            h = np.empty(wl_ * hl, np.float32)
            O.lib().or_half_sample_f(np.ascontiguousarray(src), h, wl_, hl, 1, 0.3)
            src = h
        v = np.empty(wl_ * hl * 3, np.float32)
        O.lib().or_depth2vertex_f(np.ascontiguousarray(src), v, wl_, hl, wl.cam6(), 320, 240, 1 << level)
        dv = nuif.DeviceArray.from_host(v)
        start = nuif.upload_struct(nuif.rigid(R, t))
        sums = {}
        for mode in (nuif.ICP_REDUCE_REFERENCE, nuif.ICP_REDUCE_COMPENSATED):
            nuif.check(gpu.L().nuif_icp_state_init(st.ptr, start.ptr, None))
            nuif.check(gpu.L().nuif_icp_begin(st.ptr, start.ptr, None))
            nuif.check(gpu.L().nuif_icp_iteration(st.ptr, part.ptr, C.byref(cam), dv.ptr, level, dvr.ptr, dnr.ptr,
                                                  prev.ptr, 0.1333, 0, 0, mode, None))
            res = nuif.IcpResult()
            nuif.check(gpu.L().nuif_icp_state_read(st.ptr, C.byref(res), None))
            sums[mode] = np.array(res.sums)
        s_gpu = sums[nuif.ICP_REDUCE_COMPENSATED]
        s_f = np.empty(29)
        O.lib().or_icp_sums_f(wl.cam6(), 320, 240, v, level, R.reshape(-1).copy(), t, vref.reshape(-1).copy(),
                              nref.reshape(-1).copy(), prevR.reshape(-1).copy(), prevt, 0.1333, s_f)
        s_d = np.empty(29)
        O.lib().or_icp_sums_d(wl.cam6(), 320, 240, v.astype(np.float64), level, R.reshape(-1).astype(np.float64),
                              t.astype(np.float64), vref.reshape(-1).astype(np.float64),
                              nref.reshape(-1).astype(np.float64), prevR.reshape(-1).astype(np.float64),
                              prevt.astype(np.float64), 0.1333, s_d)
        assert s_gpu[28] > 100
        # reference order: the 29 float sums are the reference's, bit for bit
        np.testing.assert_array_equal(sums[nuif.ICP_REDUCE_REFERENCE], s_f)
        # correspondence decisions are float decisions identical to the reference's
        assert s_gpu[28] == s_f[28]
        # the compensated FP32 + FP64 sums are at least as close to the FP64 sums as the
        # reference's float tree
        scale = np.abs(s_d[:28]).max()
        assert np.abs(s_gpu[:28] - s_d[:28]).max() <= max(np.abs(s_f[:28] - s_d[:28]).max(), 1e-6 * scale) + 1e-9


# ------------------------------------------------------------------ dense TSDF

def _small_volume_inputs(wl, frame):
    """Ray-length depth of a frame (passing filter + scaleDepths of the oracle)."""
    d, _, _ = wl.render(frame)
    dep = np.empty(d.size, np.float32)
    O.lib().or_passing_filter_f(d.reshape(-1), dep, d.size, 0.5, 4.5)
    O.lib().or_scale_depths_f(dep, wl.width, wl.height, wl.cam6())
    return dep


@pytest.mark.parametrize("res,bounds", [(128, False), (128, True), (96, False)])  # 96^3: generic column kernel
def test_integrate_bit_exact_s16(gpu, res, bounds):
    """With depth bounds, runs of voxels behind every surface they project to are
    skipped wholesale; the volume must not change by a single bit."""
    announce(f"integrate S16 {res}^3 320x240, 3 frames with motion, depth bounds {bounds}")
    wl = W.SMALL
    N = res
    p = nuif.tsdf_params(N, wl.volume_dim, wl.trunc)
    vol = np.zeros((N ** 3, 2), np.int16)
    vol[:, 0] = 32767
    dvol = nuif.DeviceArray.from_host(vol)
    cam = wl.camera_params()
    for k, f in enumerate([0, 4, 9]):
        R = rot_y(-2.0 * f).T.copy()
        t = np.array([0.01 * k, -0.005 * k, 0.02 * k], np.float32)
        dep = _small_volume_inputs(wl, f)
        ddep = nuif.DeviceArray.from_host(dep)
        pose = nuif.upload_struct(nuif.rigid(R, t))
        db = None
        if bounds:
            db = nuif.DeviceArray((gpu.L().nuif_tsdf_depth_bounds_bytes(wl.width, wl.height),), np.uint8)
            nuif.check(gpu.L().nuif_tsdf_depth_bounds(ddep.ptr, wl.width, wl.height, db.ptr, None, None))
            # the coarsest level is the frame's largest valid depth
            mx = db.download().view(np.float32)[-1]
            assert mx == np.nanmax(dep)
        nuif.check(gpu.L().nuif_tsdf_integrate(ddep.ptr, None, dvol.ptr, None, C.byref(p), C.byref(cam), pose.ptr,
                                               0, 128, None, None, db.ptr if db else None, None, None))
        O.lib().or_integrate_s16_f(dep, None, vol.reshape(-1), None, N, wl.volume_dim, wl.trunc, wl.cam6(),
                                   wl.width, wl.height, R.reshape(-1).copy(), t, 128)
    got = dvol.download()
    same = np.all(got == vol, axis=1)
    assert (vol[:, 1] > 0).sum() > 10000
    assert same.mean() == 1.0, f"{(~same).sum()} voxels differ"


def _surface_blocks_of(vol, N):
    """Blocks of 8^3 voxels holding an observed voxel with tsdf <= 0 (y-major, z fastest)."""
    v = vol.reshape(N, N, N, 2)  # (y, x, z)
    neg = (v[..., 1] > 0) & (v[..., 0] <= 0)
    nb = (N + 7) // 8
    pad = np.zeros((nb * 8, nb * 8, nb * 8), bool)
    pad[:N, :N, :N] = neg
    return pad.reshape(nb, 8, nb, 8, nb, 8).any(axis=(1, 3, 5)).astype(np.uint8).reshape(-1)


def _with_superblocks(blocks, N):
    """The two-level surface buffer: blocks, then 32^3 superblocks (OR of 4^3 blocks)."""
    nb, ns = (N + 7) // 8, (N + 31) // 32
    pad = np.zeros((ns * 4, ns * 4, ns * 4), np.uint8)
    pad[:nb, :nb, :nb] = blocks.reshape(nb, nb, nb)
    sup = pad.reshape(ns, 4, ns, 4, ns, 4).max(axis=(1, 3, 5)).reshape(-1)
    return np.concatenate([blocks, sup]).astype(np.uint8)


@pytest.mark.parametrize("res", [128, 96])  # 96^3: generic column kernel
def test_surface_blocks_mark_stored_nonpositive_tsdf(gpu, res):
    """Three integrations of one frame at one pose: the set of voxels holding tsdf <= 0
    never changes, so the blocks the integration marks are exactly those holding one."""
    announce(f"surface blocks {res}^3 320x240, 3 integrations of one frame")
    wl = W.SMALL
    N = res
    p = nuif.tsdf_params(N, wl.volume_dim, wl.trunc)
    dvol = nuif.DeviceArray((N ** 3 * 4,), np.uint8)
    nuif.check(gpu.L().nuif_tsdf_reset(dvol.ptr, None, C.byref(p), 0, None))
    nb = gpu.L().nuif_tsdf_surface_blocks_bytes(C.byref(p))
    assert nb == ((N + 7) // 8) ** 3 + ((N + 31) // 32) ** 3
    surf = nuif.DeviceArray((nb,), np.uint8)
    nuif.check(gpu.L().nuif_fill_u8(surf.ptr, 7, nb, None))
    nuif.check(gpu.L().nuif_tsdf_surface_blocks_clear(surf.ptr, C.byref(p), None))
    cam = wl.camera_params()
    dep = _small_volume_inputs(wl, 2)
    ddep = nuif.DeviceArray.from_host(dep)
    pose = nuif.upload_struct(nuif.rigid(np.eye(3), np.zeros(3)))
    for _ in range(3):
        nuif.check(gpu.L().nuif_tsdf_integrate(ddep.ptr, None, dvol.ptr, None, C.byref(p), C.byref(cam), pose.ptr,
                                               0, 128, None, None, None, surf.ptr, None))
    expect = _with_superblocks(_surface_blocks_of(dvol.download().view(np.int16).reshape(-1, 2), N), N)
    got = surf.download()
    assert expect.sum() > 100
    np.testing.assert_array_equal(got, expect)


@pytest.mark.parametrize("storage", [1, 2])
def test_integrate_f16_f32_close_to_fp32_math(gpu, storage):
    announce(f"integrate storage {storage} 128^3 vs float oracle")
    wl = W.SMALL
    N = 128
    p = nuif.tsdf_params(N, wl.volume_dim, wl.trunc)
    vb = gpu.L().nuif_tsdf_voxel_bytes(storage)
    dvol = nuif.DeviceArray((N ** 3 * vb,), np.uint8)
    nuif.check(gpu.L().nuif_tsdf_reset(dvol.ptr, None, C.byref(p), storage, None))
    cam = wl.camera_params()
    vol = np.zeros((N ** 3, 2), np.int16)
    vol[:, 0] = 32767
    R = np.eye(3, dtype=np.float32)
    t = np.zeros(3, np.float32)
    dep = _small_volume_inputs(wl, 0)
    ddep = nuif.DeviceArray.from_host(dep)
    pose = nuif.upload_struct(nuif.rigid(R, t))
    for _ in range(3):
        nuif.check(gpu.L().nuif_tsdf_integrate(ddep.ptr, None, dvol.ptr, None, C.byref(p), C.byref(cam), pose.ptr,
                                               storage, 128, None, None, None, None, None))
        O.lib().or_integrate_s16_f(dep, None, vol.reshape(-1), None, N, wl.volume_dim, wl.trunc, wl.cam6(),
                                   wl.width, wl.height, R.reshape(-1).copy(), t, 128)
    tg, wg = nuif.decode_volume(dvol.download(), storage)
    tr, wr = vol[:, 0] / np.float32(32766), vol[:, 1]
    np.testing.assert_array_equal(wg, wr)
    m = wr > 0
    # FP16 storage: half-precision rounding of values in [-1, 1] (<= 2^-11); S16 RTZ bias <= 3/32766
    tol = 1e-3 if storage == 1 else 2e-4
    assert np.abs(tg[m] - tr[m]).max() <= tol


def test_raycast_matches_oracle(gpu):
    announce("raycast S16 128^3 320x240 after 4 integrations")
    wl = W.SMALL
    e = _oracle_state(wl, 4)
    vol = e.volume_raw()
    N = wl.volume_res
    p = nuif.tsdf_params(N, wl.volume_dim, wl.trunc)
    cam = wl.camera_params()
    _, _, gt = wl.render(4)
    R = gt[:, :3].astype(np.float32).copy()
    t = gt[:, 3].astype(np.float32).copy()
    dvol = nuif.DeviceArray.from_host(vol)
    pose = nuif.upload_struct(nuif.rigid(R, t))
    vm = nuif.DeviceArray((wl.width * wl.height, 3), np.float32)
    nm = nuif.DeviceArray((wl.width * wl.height, 3), np.float32)
    nuif.check(gpu.L().nuif_tsdf_raycast(dvol.ptr, None, C.byref(p), C.byref(cam), pose.ptr, vm.ptr, nm.ptr, None, 0,
                                         None))
    vr = np.empty(wl.width * wl.height * 3, np.float32)
    nr = np.empty_like(vr)
    O.lib().or_raycast_s16_f(vol, None, N, wl.volume_dim, wl.trunc, wl.cam6(), wl.width, wl.height,
                             R.reshape(-1).copy(), t, vr, nr, None)
    vg, ng = vm.download().reshape(-1), nm.download().reshape(-1)
    assert (~np.isnan(vr)).sum() > 3 * 5000
    np.testing.assert_array_equal(np.isnan(vg), np.isnan(vr))
    np.testing.assert_array_equal(vg, vr)
    np.testing.assert_array_equal(np.isnan(ng), np.isnan(nr))
    np.testing.assert_array_equal(ng, nr)


# ------------------------------------------------------------------ end to end


@pytest.mark.parametrize("extra", [0.0, 0.3])
def test_raycast_surface_blocks_identical(gpu, extra):
    """The raycast with surface blocks and superblocks (exact, or with 30 % of the other
    blocks marked too) gives the oracle's maps bit for bit."""
    announce(f"raycast S16 128^3 with surface blocks (extra {extra})")
    wl = W.SMALL
    e = _oracle_state(wl, 4)
    vol = e.volume_raw()
    N = wl.volume_res
    p = nuif.tsdf_params(N, wl.volume_dim, wl.trunc)
    cam = wl.camera_params()
    _, _, gt = wl.render(4)
    R = gt[:, :3].astype(np.float32).copy()
    t = gt[:, 3].astype(np.float32).copy()
    blocks = _surface_blocks_of(vol.reshape(-1, 2), N)
    rng = np.random.default_rng(4)
    blocks = blocks | (rng.random(blocks.size) < extra).astype(np.uint8)
    assert 0 < blocks.mean() < 0.5 or extra > 0
    blocks = _with_superblocks(blocks, N)
    assert gpu.L().nuif_tsdf_surface_blocks_bytes(C.byref(p)) == blocks.size
    dvol = nuif.DeviceArray.from_host(vol)
    dsurf = nuif.DeviceArray.from_host(blocks)
    pose = nuif.upload_struct(nuif.rigid(R, t))
    vm = nuif.DeviceArray((wl.width * wl.height, 3), np.float32)
    nm = nuif.DeviceArray((wl.width * wl.height, 3), np.float32)
    nuif.check(gpu.L().nuif_tsdf_raycast_gated(dvol.ptr, None, C.byref(p), C.byref(cam), pose.ptr, vm.ptr, nm.ptr,
                                               None, 0, None, dsurf.ptr, None))
    vr = np.empty(wl.width * wl.height * 3, np.float32)
    nr = np.empty_like(vr)
    O.lib().or_raycast_s16_f(vol, None, N, wl.volume_dim, wl.trunc, wl.cam6(), wl.width, wl.height,
                             R.reshape(-1).copy(), t, vr, nr, None)
    vg, ng = vm.download().reshape(-1), nm.download().reshape(-1)
    assert (~np.isnan(vr)).sum() > 3 * 5000
    np.testing.assert_array_equal(np.isnan(vg), np.isnan(vr))
    np.testing.assert_array_equal(vg[~np.isnan(vr)], vr[~np.isnan(vr)])
    np.testing.assert_array_equal(np.isnan(ng), np.isnan(nr))
    np.testing.assert_array_equal(ng[~np.isnan(nr)], nr[~np.isnan(nr)])

def _run_engine(wl, frames, storage=0, use_graph=False, gate=0.0):
    eng = nuif.Engine(wl.kinfu_desc(storage=storage, gate=gate, use_graph=use_graph))
    for f in range(frames):
        d, _, _ = wl.render(f)
        dd = nuif.DeviceArray.from_host(d)
        eng.run(dd, sync=False)
    eng.synchronize()
    return eng, eng.poses(frames)


def test_engine_small_matches_float_oracle(gpu):
    announce("engine 320x240 128^3, 12 frames vs float oracle")
    wl = W.SMALL
    n = 12
    eng, poses = _run_engine(wl, n)
    ref = O.OracleEngine(wl, prec=0, gate=0.0)
    for f in range(n):
        ok, pr, _ = ref.run(wl.render(f)[0])
        assert ok == 1 and poses[f, 12] == 1.0
        dt, ang = pose_err(poses[f, :9].reshape(3, 3), poses[f, 9:12], pr[:9].reshape(3, 3), pr[9:])
        assert dt < 1e-4 and ang < 0.01, (f, dt, ang)
        # the whole float path is reproducible: the poses are the oracle's, bit for bit
        np.testing.assert_array_equal(poses[f, :12], pr.astype(np.float32))
    raw = eng.volume_raw().view(np.int16).reshape(-1, 2)
    np.testing.assert_array_equal(raw, ref.volume_raw().reshape(-1, 2))


def test_engine_graph_replay_identical(gpu):
    announce("engine 320x240 graph vs stream, 8 frames")
    wl = W.SMALL
    _, p1 = _run_engine(wl, 8, use_graph=False)
    e2, p2 = _run_engine(wl, 8, use_graph=True)
    np.testing.assert_array_equal(p1, p2)


def test_engine_full_size_against_oracles(gpu):
    announce("engine 640x480 512^3, 4 frames vs float and FP64 oracles")
    wl = W.FULL
    n = 4
    eng, poses = _run_engine(wl, n)
    o32 = O.OracleEngine(wl, prec=0, gate=0.0)
    o64 = O.OracleEngine(wl, prec=1, gate=0.0)
    for f in range(n):
        d = wl.render(f)[0]
        _, p32, _ = o32.run(d)
        _, p64, _ = o64.run(d)
        dt, ang = pose_err(poses[f, :9].reshape(3, 3), poses[f, 9:12], p32[:9].reshape(3, 3), p32[9:])
        assert dt < 1e-4 and ang < 0.01, ("float oracle", f, dt, ang)
        np.testing.assert_array_equal(poses[f, :12], p32.astype(np.float32))  # bit for bit
        # against FP64: within 0.1 mm / 0.01 deg, or no further than the reference's own
        # float arithmetic is from FP64 (its accept/reject decisions amplify float rounding)
        dt, ang = pose_err(poses[f, :9].reshape(3, 3), poses[f, 9:12], p64[:9].reshape(3, 3), p64[9:])
        dtf, angf = pose_err(p32[:9].reshape(3, 3), p32[9:], p64[:9].reshape(3, 3), p64[9:])
        print(f"frame {f}: gpu-fp64 {dt * 1e3:.4f} mm {ang:.5f} deg; float-fp64 {dtf * 1e3:.4f} mm {angf:.5f} deg")
        assert dt <= max(1e-4, 1.5 * dtf + 2e-5) and ang <= max(0.01, 1.5 * angf + 2e-3), ("FP64 oracle", f, dt, ang)
    # TSDF against the float oracle's short2 volume: identical voxels
    raw = eng.volume_raw().view(np.int16).reshape(-1, 2)
    ref = o32.volume_raw().reshape(-1, 2)
    same = np.all(raw == ref, axis=1)
    print(f"voxels identical to the float oracle: {same.mean():.7f}")
    assert same.all(), f"{(~same).sum()} voxels differ"
    t_g, w_g = raw[:, 0] / 32766.0, raw[:, 1]
    t_r = ref[:, 0] / 32766.0
    m = (w_g > 0) & (ref[:, 1] > 0)
    assert np.abs(t_g[m] - t_r[m]).max() <= 1e-3
    # raycast (reference maps for the next frame): bit for bit, NaN pattern included,
    # as at 128^3 -- the superblock / surface-block skips and the walk states at full size
    vg, ng = eng.maps()
    vr, nr = o32.refmaps()
    assert (~np.isnan(vr[:, 2])).sum() > 50000
    for name, a, b in (("vmap", vg, vr), ("nmap", ng, nr)):
        a = np.asarray(a, np.float32)
        b = np.asarray(b, np.float32)
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b), err_msg=name)
        m = ~np.isnan(a)
        np.testing.assert_array_equal(a[m].view(np.uint32), b[m].view(np.uint32), err_msg=name)


# ------------------------------------------------------------------ multi-session server

def test_server_sessions_match_standalone_engine(gpu):
    """Three resident sessions advanced round by round on their own streams give
    the standalone engine's poses and volume bit for bit; admission refuses a
    session beyond the device-memory budget."""
    announce("server: 3 sessions 320x240 128^3, 6 rounds vs standalone engine")
    wl = W.SMALL
    n = 6
    frames = [nuif.DeviceArray.from_host(wl.render(f)[0]) for f in range(n + 2)]
    desc = wl.kinfu_desc(use_graph=True)
    per = nuif.Server.session_bytes(desc)
    assert per > wl.volume_res ** 3 * 4
    srv = nuif.Server(3 * per + per // 2)
    ids = [srv.open(desc) for _ in range(3)]
    assert None not in ids and len(set(ids)) == 3
    assert srv.open(desc) is None  # over the budget
    assert srv.bytes_in_use == 3 * per
    lats = []
    for r in range(n):
        srv.submit(ids, [frames[r + k] for k in range(3)])  # session k plays frames k, k+1, ...
        lats.append(srv.wait())
    lats = np.array(lats)
    assert (lats > 0).all() and np.isfinite(lats).all()
    for k, sid in enumerate(ids):
        ref = nuif.Engine(wl.kinfu_desc(use_graph=True))
        for r in range(n):
            ref.run(frames[r + k], sync=False)
        ref.synchronize()
        np.testing.assert_array_equal(srv.sessions[sid].poses(n), ref.poses(n))
        np.testing.assert_array_equal(srv.sessions[sid].volume_raw(), ref.volume_raw())
        ref.close()
    srv.close_session(ids[1])
    assert srv.bytes_in_use == 2 * per
    assert srv.open(desc) == ids[1]  # the freed id is reused
    srv.close()


def test_server_32_sessions_match_standalone_engines(gpu):
    """32 sessions in one batched round (the launch sizes of the server's workload: a whole
    column per integration segment, blocks numbered per XCD) give the standalone engines'
    poses, maps and volumes bit for bit (a session alone walks its columns in 64-voxel
    segments from precomputed walk states)."""
    announce("server: 32 F16 sessions 320x240 128^3, 4 rounds vs standalone engines")
    wl = W.SMALL
    S, n = 32, 4
    assert gpu.L().nuif_schedule_segment(S) >= wl.volume_res
    frames = [nuif.DeviceArray.from_host(wl.render(f)[0]) for f in range(n + S)]
    desc = wl.kinfu_desc(storage=nuif.TSDF_F16, use_graph=True)
    srv = nuif.Server(S * nuif.Server.session_bytes(desc) + (1 << 30))
    ids = [srv.open(desc) for _ in range(S)]
    assert None not in ids
    for r in range(n):
        srv.submit(ids, [frames[r + k] for k in range(S)])  # session k plays frames k, k+1, ...
        srv.wait()
        assert srv.last_batch == (0 if r == 0 else S)
    for k, sid in enumerate(ids):
        ref = nuif.Engine(desc)
        for r in range(n):
            ref.run(frames[r + k], sync=False)
        ref.synchronize()
        s = srv.sessions[sid]
        np.testing.assert_array_equal(s.poses(n).view(np.uint32), ref.poses(n).view(np.uint32), err_msg=str(k))
        for a, b in zip(s.maps(), ref.maps()):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=str(k))
        if k in (0, 13, 31):
            np.testing.assert_array_equal(s.volume_raw(), ref.volume_raw(), err_msg=str(k))
        ref.close()
    srv.close()


def test_server_mixed_sessions_match_standalone_engines(gpu):
    """One server round after another over sessions of every kind side by side
    (320x240 / 640x480 / 1280x720 frames; dense S16, F16 and F32 volumes of 128^3 and
    256^3, a hash scene and a shifting volume): each session's poses and maps, and
    the dense volumes, equal those of the same session run alone, bit for bit.  The
    dense sessions' tracked frames run as one batched round (nuif_fusion_round) beside
    the hash and shifting sessions on their own streams."""
    import dataclasses
    announce("server: 5 mixed sessions (320x240..1280x720, S16/F16/F32, hash, shift), 4 rounds vs standalone")
    full256 = dataclasses.replace(W.FULL, volume_res=256)
    hd = dataclasses.replace(W.FULL, name="hd-1280x720-256^3", width=1280, height=720, fx=1050.0, fy=1050.0,
                             cx=639.5, cy=359.5, volume_res=256, seed_index=3)
    cases = [(W.SMALL, W.SMALL.kinfu_desc(use_graph=True), True),
             (full256, full256.kinfu_desc(storage=nuif.TSDF_F16, use_graph=True), True),
             (hd, hd.kinfu_desc(storage=nuif.TSDF_F32, use_graph=True), True),
             (W.HASH_SMALL, W.HASH_SMALL.kinfu_desc(bricks=1 << 17), False),
             (W.SMALL, W.SMALL.kinfu_desc(shift=3), True)]
    n = 4
    frames = [[nuif.DeviceArray.from_host(wl.render(f)[0]) for f in range(n)] for wl, _, _ in cases]
    srv = nuif.Server(sum(nuif.Server.session_bytes(d) for _, d, _ in cases) + (1 << 30))
    ids = [srv.open(d) for _, d, _ in cases]
    assert None not in ids
    for r in range(n):
        srv.submit(ids, [frames[k][r] for k in range(len(cases))])
        lat = srv.wait()
        assert (lat > 0).all()
        # after their first frames the three plain dense sessions run as one batched round
        assert srv.last_batch == (0 if r == 0 else 3), (r, srv.last_batch)
    for k, ((wl, d, dense), sid) in enumerate(zip(cases, ids)):
        ref = nuif.Engine(d)
        for r in range(n):
            ref.run(frames[k][r], sync=False)
        ref.synchronize()
        s = srv.sessions[sid]
        assert s.synchronize() == 1, wl.name
        np.testing.assert_array_equal(s.poses(n).view(np.uint32), ref.poses(n).view(np.uint32), err_msg=wl.name)
        for a, b in zip(s.maps(), ref.maps()):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=wl.name)
        if dense:
            np.testing.assert_array_equal(s.volume_raw(), ref.volume_raw(), err_msg=wl.name)
        ref.close()
    srv.close()
