"""CPU-side checks of the native libraries: they load, export every symbol the
public headers declare, refuse to run without a GPU, and the host logic
(configs, workload generator) behaves.  No kernel is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from natureuserinterfacestudio_amd import libkinfu, libnuif, libsynth, nuif
from natureuserinterfacestudio_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nuif[a-z]*_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = libnuif()
    names = declared("nuif.h")
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_engine_library_exports():
    lib = libkinfu()
    names = declared("nuif_kinfu.h")
    assert len(names) >= 29
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(nuif.declared_kinfu_symbols()), sorted(set(names) - set(nuif.declared_kinfu_symbols()))


def test_ingest_library_exports():
    lib = libkinfu()
    names = declared("nuif_io.h")
    assert len(names) == 11
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(nuif.declared_kinfu_symbols())


def test_python_binding_covers_header():
    # every header function has a ctypes signature in nuif.py (so tests call them typed)
    names = set(declared("nuif.h"))
    assert names <= set(nuif.declared_symbols()), sorted(names - set(nuif.declared_symbols()))


def test_struct_layouts_match_reference():
    # K/gpu_def.h:5-34
    assert C.sizeof(nuif.NuiCLCameraParams) == 40
    assert C.sizeof(nuif.NuiCLRigidTransform) == 84
    assert C.sizeof(nuif.TsdfParams) == 40


def test_calls_without_init_fail_loudly():
    lib = nuif.L()
    p = C.c_void_p()
    if nuif._initialised:
        pytest.skip("device already initialised in this process")
    assert lib.nuif_malloc(C.byref(p), 64) == -4  # NUIF_ERR_NODEV
    assert b"nuif_init" in lib.nuif_last_error()


def test_voxel_bytes_and_params():
    lib = nuif.L()
    assert [lib.nuif_tsdf_voxel_bytes(s) for s in (0, 1, 2, 7)] == [4, 4, 8, 0]
    p = nuif.TsdfParams()
    assert lib.nuif_tsdf_params(128, 128, 128, 2.0, 2.0, 2.0, 0.03, C.byref(p)) == 0
    assert np.isclose(p.cell_size[0], 2.0 / 128)
    assert np.isclose(p.tranc_dist, 2.1 * 2.0 / 128, rtol=1e-6)  # OCL/NuiKinfuOpenCLScene.cpp:41
    assert lib.nuif_tsdf_params(0, 1, 1, 1.0, 1.0, 1.0, 0.03, C.byref(p)) == -1


@pytest.mark.parametrize("shape,expect", [((128, 128, 64), 64 * 8), ((128, 128, 128), (128 * 128 + 1) * 8 + 128 * 8),
                                          ((512, 512, 512), 7 * (512 * 512 + 1) * 8 + 512 * 8),
                                          ((64, 96, 320), 4 * (64 * 96 + 1) * 8 + 320 * 8),
                                          ((8, 8, 129), 2 * (64 + 1) * 8 + 129 * 8)])
def test_walk_state_workspace_bytes(shape, expect):
    """8 B per column and z segment after the first, plus 8 B per such segment, for the
    shortest segment a launch may use (64 voxels), then the z table: 8 B per z."""
    lib = nuif.L()
    p = nuif.TsdfParams()
    assert lib.nuif_tsdf_params(*shape, 1.0, 1.0, 1.0, 0.03, C.byref(p)) == 0
    assert lib.nuif_tsdf_walk_states_bytes(C.byref(p)) == expect
    assert lib.nuif_tsdf_walk_states_bytes(None) == 0


def test_schedule_table():
    """The compiled launch schedule and the segment rule (short segments for a single
    session, long ones for a batched round), as the schedule sweep chose them."""
    lib = nuif.L()
    out = (C.c_int32 * 4)()
    assert lib.nuif_schedule(out) == 0
    assert list(out) == [8, 8, 2, 64]
    assert [lib.nuif_schedule_segment(n) for n in (1, 2, 7, 8, 16, 32, 64)] == [64, 128, 128, 512, 512, 512, 512]
    assert lib.nuif_schedule_segment(0) == -1
    from natureuserinterfacestudio_amd import tune_schedules as T
    assert T.CANDIDATES["default"] == {} and set(T.SEGMENTS) == {64, 128, 256, 512}


def test_synthetic_frames_are_deterministic():
    wl = W.tiny()
    d1, c1, p1 = wl.render(3, color=True)
    d2, c2, p2 = wl.render(3, color=True)
    assert np.array_equal(d1, d2) and np.array_equal(c1, c2) and np.array_equal(p1, p2)
    d0, _, p0 = wl.render(0)
    np.testing.assert_allclose(p0[:, :3], np.eye(3), atol=1e-7)  # frame 0 is the world frame
    np.testing.assert_allclose(p0[:, 3], 0, atol=1e-7)
    assert not np.array_equal(d0, d1)
    assert (d1 > 0).mean() > 0.2


def test_noise_free_depth_matches_geometry():
    import dataclasses
    wl = dataclasses.replace(W.tiny(), noise=0, hole_frac=0.0)
    d, _, _ = wl.render(0)
    # the floor plane y = 0.4 seen by the identity camera: depth = 0.4 * fy / (v - cy)
    H_, W_ = d.shape
    v = H_ - 1
    expect = 0.4 * wl.fy / (v - wl.cy)
    row = d[v].astype(np.float64) / 1000.0
    assert np.median(np.abs(row[row > 0] - expect)) < 1e-3 or (row > 0).sum() < W_ // 2


def test_tracker_config_file_roundtrip(tmp_path):
    # VO/NuiTrackerConfig.cpp:26-115 key=value format, parsed by the C++ NuiTrackerConfig
    lib = nuif.K()
    f = tmp_path / "tracker.cfg"
    f.write_text("icp_filter_radius=2\nicp_dist_threshold=0.05\nicp_iteration=7\nicp_iteration=0\n"
                 "icp_iteration=3\nunknown=1\n")
    d = W.SMALL.kinfu_desc()
    assert lib.nuifk_load_tracker_config(str(f).encode(), C.byref(d)) == 0
    assert d.filter_radius == 2 and np.isclose(d.dist_threshold, 0.05)
    assert d.levels == 2 and list(d.iterations)[:2] == [7, 3]  # zero iterations are dropped
    assert np.isclose(d.sigma_depth2_inv_half, 50.0)  # untouched keys keep the defaults
    out = tmp_path / "log.cfg"
    assert lib.nuifk_save_tracker_config(str(out).encode(), C.byref(d)) == 0
    d2 = W.SMALL.kinfu_desc()
    assert lib.nuifk_load_tracker_config(str(out).encode(), C.byref(d2)) == 0
    assert d2.levels == 2 and d2.filter_radius == 2 and np.isclose(d2.dist_threshold, 0.05)
    assert lib.nuifk_load_tracker_config(str(tmp_path / "missing.cfg").encode(), C.byref(d2)) == -1
