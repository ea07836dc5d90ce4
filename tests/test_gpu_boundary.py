"""GPU tests of the reference interfaces that the fusion path keeps for its callers:

  * NuiKinfuFrame::UpdateColorBuffers(ColorSpacePoint*, UINT*, UINT, const NuiColorImage&)
    (DS/NuiKinfuFrame.h:22, OCL/NuiKinfuOpenCLFrame.cpp:405-448): the C ABI kernel and the
    engine frame against the oracle's restatement, bit for bit;
  * NuiKinfuScene::Volume2Mesh(NuiMeshShape*) (DS/NuiKinfuScene.h:43, the SLAM
    controller's getMesh): the points and triangle indices equal the indexed mesh of
    nuif_mesh_dense on the same volume (itself bit-exact against the host mesher).
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


def _mapping(n, iw, ih, seed):
    rng = np.random.default_rng(seed)
    m = np.stack([rng.uniform(-3.0, iw + 3.0, n), rng.uniform(-3.0, ih + 3.0, n)], 1).astype(np.float32)
    lut = rng.integers(0, n + n // 20, n).astype(np.uint32)  # some entries past the frame
    image = rng.integers(0, 256, (ih, iw, 4), dtype=np.uint8)
    return m, lut, image


@pytest.mark.parametrize("with_lut", [False, True], ids=["no-table", "table"])
def test_map_depth_to_color_matches_oracle(gpu, with_lut):
    announce(f"depth-to-colour map, 320x240 depth, 640x480 colour, table={with_lut}")
    n, iw, ih = 320 * 240, 640, 480
    m, lut, image = _mapping(n, iw, ih, 11)
    want = O.map_depth_to_color(m, lut if with_lut else None, image)
    dm, dl, di = nuif.DeviceArray.from_host(m), nuif.DeviceArray.from_host(lut), nuif.DeviceArray.from_host(image)
    out = nuif.DeviceArray((n, 4), np.uint8)
    nuif.check(gpu.L().nuif_map_depth_to_color(dm.ptr, dl.ptr if with_lut else None, n, di.ptr, iw, ih, out.ptr,
                                               None))
    nuif.check(gpu.L().nuif_stream_synchronize(None))
    got = out.download()
    assert (want[:, 3] == 255).sum() > n // 2
    np.testing.assert_array_equal(got, want)


def test_engine_frame_colour_overload(gpu):
    announce("engine frame UpdateColorBuffers(ColorSpacePoint*, ...)")
    wl = W.SMALL
    n = wl.width * wl.height
    m, lut, image = _mapping(n, 512, 424, 12)
    eng = nuif.Engine(wl.kinfu_desc(color=True))
    eng.map_color(m, image, lut)
    np.testing.assert_array_equal(eng.frame_colors().reshape(-1, 4), O.map_depth_to_color(m, lut, image))


def test_volume2mesh_mesh_shape_matches_indexed_mesh(gpu):
    announce("Volume2Mesh(NuiMeshShape*) vs nuif_mesh_dense, 128^3 after 3 frames")
    wl = W.SMALL
    eng = nuif.Engine(wl.kinfu_desc())
    for f in range(3):
        eng.run(nuif.DeviceArray.from_host(wl.render(f)[0]), sync=True)
    pts, idx = eng.mesh_shape()
    assert len(idx) % 3 == 0 and len(idx) > 1000 and len(pts) > 100
    L = gpu.L()
    p = nuif.tsdf_params(wl.volume_res, wl.volume_dim, wl.trunc)
    vol = nuif.K().nuifk_volume_ptr(eng._h)
    cap = len(idx)
    scratch = nuif.DeviceArray((L.nuif_mesh_dense_scratch_bytes(C.byref(p), cap),), np.uint8)
    dp, di = nuif.DeviceArray((cap, 3), np.float32), nuif.DeviceArray((cap,), np.uint32)
    np_, ni = C.c_uint32(), C.c_uint32()
    nuif.check(L.nuif_mesh_dense(C.c_void_p(vol), C.byref(p), 0, cap, dp.ptr, di.ptr, C.byref(np_), C.byref(ni),
                                 scratch.ptr, None))
    assert (np_.value, ni.value) == (len(pts), len(idx))
    np.testing.assert_array_equal(pts, dp.download()[:np_.value])
    np.testing.assert_array_equal(idx, di.download()[:ni.value].astype(np.int32))
    assert idx.max() == len(pts) - 1 and idx.min() == 0
