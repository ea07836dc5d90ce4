"""GPU parity of the photometric (colour) tracker against the CPU oracle.

Bars: pyramid and gradient bytes exact (both gradient encodings); the colour
difference sums exact (integers); one iteration's 29 sums equal to the oracle's
reference-order reduction bit for bit; the whole tracker (3 levels, 10/5/4) in both
decodes: the same iteration count, error, count and pose bit for bit."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O
from natureuserinterfacestudio_amd import nuif
from natureuserinterfacestudio_amd import workloads as W

pytestmark = pytest.mark.gpu

WL = W.SMALL


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _frame_model(f):
    d, bgra, gt = WL.render(f, color=True)
    dep = np.empty(d.size, np.float32)
    O.lib().or_passing_filter_f(d.reshape(-1), dep, d.size, 0.5, 4.5)
    cam = WL.cam6()
    ys, xs = np.divmod(np.arange(d.size), WL.width)
    v = np.stack([dep * ((xs - cam[2]) / cam[0]), dep * ((ys - cam[3]) / cam[1]), dep], -1)
    R, t = gt[:, :3].astype(np.float32), gt[:, 3].astype(np.float32)
    vw = (v @ R.T + t).astype(np.float32)
    cols = bgra.reshape(-1, 4).copy()
    cols[:, 3] = np.where(np.isnan(dep), 0, 255)
    return vw, cols, bgra, R, t


@pytest.mark.parametrize("signed", [False, True])
def test_color_pyramid_and_gradients_exact(gpu, signed):
    announce(f"colour pyramid 320x240, 3 levels, signed gradients {signed}")
    _, _, bgra, _, _ = _frame_model(2)
    tr = nuif.ColorTracker(WL.width, WL.height, 3, corrected=signed)
    tr.pyramid(nuif.DeviceArray.from_host(bgra))
    ref = bgra
    for l in range(3):
        w, h = WL.width >> l, WL.height >> l
        if l > 0:
            ref = O.color_half_sample(ref, w, h, 1)
        np.testing.assert_array_equal(tr.colors[l].download().reshape(h, w, 4), ref)
        gx, gy = O.color_gradient(ref, w, h, signed)
        np.testing.assert_array_equal(tr.gx[l].download().reshape(h, w, 4), gx)
        np.testing.assert_array_equal(tr.gy[l].download().reshape(h, w, 4), gy)


def test_color_difference_exact(gpu):
    announce("colour difference sums, levels 0-2")
    vw, cols, _, R0, t0 = _frame_model(3)
    _, _, frame, _, _ = _frame_model(4)
    tr = nuif.ColorTracker(WL.width, WL.height, 3)
    tr.pyramid(nuif.DeviceArray.from_host(frame))
    pose = nuif.upload_struct(nuif.rigid(R0.T, -(R0.T @ t0)))
    dv, dc = nuif.DeviceArray.from_host(vw), nuif.DeviceArray.from_host(cols)
    ref = frame
    for l in range(3):
        if l > 0:
            ref = O.color_half_sample(ref, WL.width >> l, WL.height >> l, 1)
        got = tr.difference(WL.camera_params(), l, dv, dc, pose)
        exp = O.color_difference(WL, l, vw, cols, ref, R0.T, -(R0.T @ t0))
        assert got == exp and exp[1] > 10000


@pytest.mark.parametrize("corrected", [False, True])
def test_color_icp_matches_oracle(gpu, corrected):
    announce(f"colour ICP 320x240, 10/5/4, corrected decode {corrected}")
    vw, cols, _, R0, t0 = _frame_model(4)
    _, _, frame, _, _ = _frame_model(5)
    tr = nuif.ColorTracker(WL.width, WL.height, 3, corrected=corrected)
    tr.pyramid(nuif.DeviceArray.from_host(frame))
    pose = nuif.upload_struct(nuif.rigid(R0, t0))
    r = tr.estimate(WL.camera_params(), nuif.DeviceArray.from_host(vw), nuif.DeviceArray.from_host(cols), pose,
                    [10, 5, 4])
    ok, R, t, its, err, cnt = O.color_icp(WL, vw, cols, frame, R0, t0, iterations=[10, 5, 4], corrected=corrected)
    assert r.ok == ok and r.iterations == its
    assert np.float32(r.error) == np.float32(err) and np.float32(r.count) == np.float32(cnt)
    g = nuif.NuiCLRigidTransform()
    nuif.check(nuif.L().nuif_memcpy_d2h(C.addressof(g), pose.ptr, C.sizeof(g), None), "d2h")
    nuif.check(nuif.L().nuif_stream_synchronize(None), "sync")
    np.testing.assert_array_equal(np.array(g.R[:], np.float32).reshape(3, 3), R)
    np.testing.assert_array_equal(np.array(g.t[:], np.float32), t)
    if corrected:
        assert ok == 1 and its > 3
    else:
        assert ok == 0 and its == 1  # the reference decode is not positive definite here


def test_color_icp_iteration_sums_match_reference_tree(gpu):
    announce("colour ICP one iteration's sums at 640x480 vs the oracle's tree")
    wl = W.FULL
    d, bgra, gt = wl.render(2, color=True)
    dep = np.empty(d.size, np.float32)
    O.lib().or_passing_filter_f(d.reshape(-1), dep, d.size, 0.5, 4.5)
    cam = wl.cam6()
    ys, xs = np.divmod(np.arange(d.size), wl.width)
    v = np.stack([dep * ((xs - cam[2]) / cam[0]), dep * ((ys - cam[3]) / cam[1]), dep], -1).astype(np.float32)
    cols = bgra.reshape(-1, 4).copy()
    cols[:, 3] = np.where(np.isnan(dep), 0, 255)
    _, frame, _ = wl.render(3, color=True)
    tr = nuif.ColorTracker(wl.width, wl.height, 1, corrected=True)
    tr.pyramid(nuif.DeviceArray.from_host(frame))
    pose = nuif.upload_struct(nuif.rigid(np.eye(3), np.zeros(3)))
    r = tr.estimate(wl.camera_params(), nuif.DeviceArray.from_host(v), nuif.DeviceArray.from_host(cols), pose, [1])
    gx, gy = O.color_gradient(frame, wl.width, wl.height, True)
    blocks = O.color_reduce(wl, 0, v, cols, frame, gx, gy, np.eye(3), np.zeros(3), signed=True)
    host = np.zeros(29, np.float32)
    for b in blocks:  # the host's sequential float sum
        host = (host + b).astype(np.float32)
    np.testing.assert_array_equal(np.array(r.sums[:], np.float32), host)
    assert host[28] > 100000


def test_intensity_icp_sums_match_oracle(gpu):
    """bgra_to_intensity, the derivatives and the joint kernel's 29 sums (levels 0 and 1,
    frame 5 against the world maps of frame 4 at a perturbed pose) bit for bit."""
    announce("joint depth + intensity ICP sums 320x240 vs oracle")
    def maps(f):
        d, bgra, gt = WL.render(f, color=True)
        dep = np.empty(d.size, np.float32)
        O.lib().or_passing_filter_f(d.reshape(-1), dep, d.size, 0.5, 4.5)
        cam = WL.cam6()
        ys, xs = np.divmod(np.arange(d.size), WL.width)
        v = np.stack([dep * ((xs - cam[2]) / cam[0]), dep * ((ys - cam[3]) / cam[1]), dep], -1).astype(np.float32)
        n = np.empty_like(v)
        O.lib().or_normals_simple_f(v.reshape(-1), n.reshape(-1), WL.width, WL.height, 0.3)
        return v, n, bgra, gt[:, :3].astype(np.float32), gt[:, 3].astype(np.float32)
    v5, n5, bgra5, R5, t5 = maps(5)
    v4, n4, bgra4, R4, t4 = maps(4)
    vw4 = (v4 @ R4.T + t4).astype(np.float32)
    nw4 = (n4 @ R4.T).astype(np.float32)
    I5 = nuif.DeviceArray((WL.width * WL.height,), np.float32)
    nuif.check(nuif.L().nuif_bgra_to_intensity(nuif.DeviceArray.from_host(bgra5).ptr, I5.ptr, WL.width * WL.height,
                                               None))
    np.testing.assert_array_equal(I5.download(), O.bgra_to_intensity(bgra5))
    I4h = O.bgra_to_intensity(bgra4)
    d4 = nuif.DeviceArray((WL.width * WL.height * 2,), np.float32)
    nuif.check(nuif.L().nuif_intensity_derivatives(nuif.DeviceArray.from_host(I4h).ptr, d4.ptr, WL.width,
                                                   WL.height, None))
    d4h = O.intensity_derivatives(I4h, WL.width, WL.height)
    np.testing.assert_array_equal(d4.download().reshape(d4h.shape), d4h)
    I5h = O.bgra_to_intensity(bgra5)
    Rp = (R5 @ np.array([[1, -0.01, 0], [0.01, 1, 0], [0, 0, 1]], np.float32)).astype(np.float32)
    tp = (t5 + np.float32([0.01, -0.005, 0.0])).astype(np.float32)
    euler = np.float32([0.01, -0.02, 0.005])
    thr = np.float32([0.1, 0.34, 0.2, 0.002])
    cam = WL.camera_params()
    for level in (0, 1):
        s = 1 << level
        def lv(a, k):
            return np.ascontiguousarray(a.reshape(WL.height, WL.width, k)[::s, ::s]).reshape(-1, k)
        args_h = [lv(v5, 3), lv(n5, 3), lv(I5h, 1), lv(vw4, 3), lv(nw4, 3), lv(I4h, 1), lv(d4h, 2)]
        exp = O.intensity_icp_sums(WL, level, *args_h, Rp, tp, R4, t4, euler, thr)
        dev = [nuif.DeviceArray.from_host(a) for a in args_h]
        ws = nuif.DeviceArray((nuif.L().nuif_intensity_icp_workspace_bytes(WL.width >> level, WL.height >> level),),
                              np.uint8)
        out = nuif.DeviceArray((29,), np.float32)
        pose = nuif.upload_struct(nuif.rigid(Rp, tp))
        prev = nuif.upload_struct(nuif.rigid(R4, t4))
        e3 = (C.c_float * 3)(*euler)
        t4c = (C.c_float * 4)(*thr)
        for _ in range(2):  # a second call re-arms the counters
            nuif.check(nuif.L().nuif_intensity_icp_sums(C.byref(cam), level, dev[0].ptr, dev[1].ptr, dev[2].ptr,
                                                        pose.ptr, dev[3].ptr, dev[4].ptr, dev[5].ptr, dev[6].ptr,
                                                        prev.ptr, e3, t4c, ws.ptr, out.ptr, None))
            np.testing.assert_array_equal(out.download(), exp)
        assert exp[28] > 1000
