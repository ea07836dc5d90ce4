"""GPU side of the frame ingest: a recording in the reference's file format
(written by the oracle's restatement of Foundation/NuiFileIOUtilities.cpp:13-52),
loaded by nuif_frame_loader with upload, arrives in device memory byte for byte,
and a session fed from the loader's device copies tracks exactly as one fed the
same frames from host arrays (poses bit-exact)."""
import ctypes as C
import sys

import numpy as np

import frame_io as F
from natureuserinterfacestudio_amd import nuif
from natureuserinterfacestudio_amd import workloads as W

import pytest

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def test_loader_upload_feeds_the_engine_exactly(gpu, tmp_path):
    wl = W.SMALL
    n = 8
    announce(f"ingest: {n} recorded 320x240 frames, loader upload, engine on device copies vs host arrays")
    frames = []
    for i in range(n):
        d, bgra, _ = wl.render(i, color=True)
        F.write_recording_frame(str(tmp_path), i, {".depth": (i, d), ".color": (i, bgra)}, compressed=(i % 2 == 0),
                                camera=(wl.fx, wl.fy, wl.cx, wl.cy))
        frames.append((d, bgra))
    ld = nuif.FrameLoader(str(tmp_path), slots=3, threads=2, upload=True)
    e_dev = nuif.Engine(wl.kinfu_desc())
    e_host = nuif.Engine(wl.kinfu_desc())
    K = nuif.K()
    for i in range(n):
        f = ld.next_raw()
        assert f.frame_id == i and f.parts == nuif.FRAME_DEPTH | nuif.FRAME_COLOR | nuif.FRAME_CAMERA
        assert (f.width[0], f.height[0]) == (wl.width, wl.height)
        assert (f.fx, f.fy, f.cx, f.cy) == tuple(np.float32([wl.fx, wl.fy, wl.cx, wl.cy]))
        dd = np.empty((wl.height, wl.width), np.uint16)
        dc = np.empty((wl.height, wl.width, 4), np.uint8)
        nuif.check(nuif.L().nuif_memcpy_d2h(dd.ctypes.data, f.device_depth, dd.nbytes, None))
        nuif.check(nuif.L().nuif_memcpy_d2h(dc.ctypes.data, f.device_color, dc.nbytes, None))
        nuif.check(nuif.L().nuif_stream_synchronize(None))
        assert np.array_equal(dd, frames[i][0]) and np.array_equal(dc, frames[i][1])
        rc = K.nuifk_run(e_dev._h, f.device_depth, 1, None, 1)
        assert rc >= 0, K.nuifk_last_error()
        ld.release(f.slot)  # after the synchronous run: the device copy is no longer read
        e_host.run(frames[i][0])
    assert ld.next_raw() is None
    p_dev, p_host = e_dev.poses(n), e_host.poses(n)
    assert len(p_dev) == n
    np.testing.assert_array_equal(p_dev.view(np.uint32), p_host.view(np.uint32))
    ld.close()
