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
