"""Synthetic workloads of the fusion path (seeded, reproducible frame by frame).

  small    320x240, fx=fy=262.5, dense 128^3 over 2 m, trunc 3 cm, ICP 10/5/4, 30 frames,
           orbit radius 1.2 m at 2 deg/frame
  full     640x480, fx=fy=525, dense 512^3 over 3 m, trunc 2.5 cm, 300 frames, orbit 1.5 m
  hash     640x480, 8x8x3 m room, random-walk camera, 5 mm voxels in 8^3 bricks, 500 frames
  server   mixed 320x240 / 640x480 / 1280x720 sessions

Frames are rendered by libnuif_synth (exact ray/primitive intersection of a
ground plane + 6 spheres + 4 boxes, Nguyen-style axial noise, mm quantisation,
2 % zero holes).
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import libsynth
from . import nuif


class SynthParams(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float),
                ("cy", C.c_float), ("scene_kind", C.c_int), ("orbit_radius", C.c_float),
                ("deg_per_frame", C.c_float), ("cube_size", C.c_float), ("jitter_pos", C.c_float),
                ("jitter_deg", C.c_float), ("hole_frac", C.c_float), ("noise", C.c_int), ("seed", C.c_ulonglong)]


SEED_BASE = 0x4E554900  # "NUI\0" + workload index


@dataclass
class Workload:
    name: str
    width: int
    height: int
    fx: float
    fy: float
    cx: float
    cy: float
    frames: int
    scene_kind: int = 0
    orbit_radius: float = 1.2
    deg_per_frame: float = 2.0
    volume_res: int = 128
    volume_dim: float = 2.0
    trunc: float = 0.03
    iterations: list = field(default_factory=lambda: [10, 5, 4])
    seed_index: int = 0
    noise: int = 1
    hole_frac: float = 0.02

    def synth_params(self):
        p = SynthParams()
        p.width, p.height = self.width, self.height
        p.fx, p.fy, p.cx, p.cy = self.fx, self.fy, self.cx, self.cy
        p.scene_kind = self.scene_kind
        p.orbit_radius = self.orbit_radius
        p.deg_per_frame = self.deg_per_frame
        p.cube_size = 2.0
        p.jitter_pos = 0.001
        p.jitter_deg = 0.1
        p.hole_frac = self.hole_frac
        p.noise = self.noise
        p.seed = SEED_BASE + self.seed_index
        return p

    def render(self, frame, color=False):
        """(depth uint16 HxW, bgra uint8 HxWx4 or None, gt pose 3x4 float32)."""
        lib = libsynth()
        p = self.synth_params()
        depth = np.empty((self.height, self.width), np.uint16)
        bgra = np.empty((self.height, self.width, 4), np.uint8) if color else None
        pose = np.empty(12, np.float32)
        rc = lib.nsynth_render(C.byref(p), C.c_int(frame), depth.ctypes.data_as(C.c_void_p),
                               bgra.ctypes.data_as(C.c_void_p) if color else None,
                               pose.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError("nsynth_render failed")
        return depth, bgra, np.concatenate([pose[:9].reshape(3, 3), pose[9:].reshape(3, 1)], axis=1)

    def cam6(self):
        return np.array([self.fx, self.fy, self.cx, self.cy, 0.5, 4.5], np.float32)

    def camera_params(self):
        return nuif.camera_params(self.width, self.height, self.fx, self.fy, self.cx, self.cy, 0.5, 4.5)

    def kinfu_desc(self, storage=nuif.TSDF_S16, gate=0.0, color=False, use_graph=False, iterations=None,
                   reduction=None, bricks=1 << 18, forward=0, shift=0, basis=(0.0, 0.0, 0.0), swapping=False,
                   buckets=0):
        d = nuif.KinfuDesc()
        d.width, d.height = self.width, self.height
        d.fx, d.fy, d.cx, d.cy = self.fx, self.fy, self.cx, self.cy
        d.depth_min, d.depth_max = 0.5, 4.5
        d.filter_radius = 3
        d.sigma_space2_inv_half = 0.08
        d.sigma_depth2_inv_half = 50.0
        d.depth_threshold = 0.3
        d.dist_threshold = 0.1
        its = iterations or self.iterations
        d.levels = len(its)
        for i, n in enumerate(its):
            d.iterations[i] = n
        d.integration_threshold = gate
        for k in range(3):
            d.resolution[k] = self.volume_res
            d.dimensions[k] = self.volume_dim
        d.tranc_dist = self.trunc
        d.color = 1 if color else 0
        d.storage = storage
        d.use_graph = 1 if use_graph else 0
        # the reference-precision volume pairs with the reference's summation order;
        # the FP16 / FP32 variants use the compensated sums
        if reduction is None:
            reduction = nuif.ICP_REDUCE_REFERENCE if storage == nuif.TSDF_S16 else nuif.ICP_REDUCE_COMPENSATED
        d.icp_reduction = reduction
        # voxel hashing (scene_kind 1 workloads): VO/NuiHashingSDFConfig.cpp:3-15 values
        d.scene = nuif.SCENE_HASH if self.volume_res == 0 else nuif.SCENE_DENSE
        d.hash_swapping = 1 if swapping else 0
        if shift:  # the shifting dense volume (OCL/NuiKinfuOpenCLShiftScene)
            d.scene = nuif.SCENE_SHIFT
            d.voxel_shift = shift
            for k in range(3):
                d.volume_basis[k] = basis[k]
        d.hash_voxel_size, d.hash_trunc_scale, d.hash_truncation = 0.005, self.trunc, 0.06
        d.hash_weight_max = 255
        d.hash_bricks = bricks
        d.hash_buckets = buckets
        d.hash_forward_render = forward
        return d


SMALL = Workload("small-320x240-128^3", 320, 240, 262.5, 262.5, 159.5, 119.5, frames=30, orbit_radius=1.2,
                 volume_res=128, volume_dim=2.0, trunc=0.03, iterations=[10, 5, 4], seed_index=0)
FULL = Workload("full-640x480-512^3", 640, 480, 525.0, 525.0, 319.5, 239.5, frames=300, orbit_radius=1.5,
                volume_res=512, volume_dim=3.0, trunc=0.025, iterations=[10, 5, 4], seed_index=1)
HASH = Workload("hash-640x480-5mm", 640, 480, 525.0, 525.0, 319.5, 239.5, frames=500, scene_kind=1,
                volume_res=0, volume_dim=0.0, trunc=0.02, iterations=[10, 5, 4], seed_index=2)


HASH_SMALL = Workload("hash-320x240-5mm", 320, 240, 262.5, 262.5, 159.5, 119.5, frames=30, scene_kind=1,
                      volume_res=0, volume_dim=0.0, trunc=0.02, iterations=[10, 5, 4], seed_index=2)

# the third resolution of the multi-session workloads (16:9, the same field of view across)
HD = Workload("hd-1280x720-512^3", 1280, 720, 1050.0, 1050.0, 639.5, 359.5, frames=300, orbit_radius=1.5,
              volume_res=512, volume_dim=3.0, trunc=0.025, iterations=[10, 5, 4], seed_index=4)
HASH_HD = Workload("hash-1280x720-5mm", 1280, 720, 1050.0, 1050.0, 639.5, 359.5, frames=500, scene_kind=1,
                   volume_res=0, volume_dim=0.0, trunc=0.02, iterations=[10, 5, 4], seed_index=4)


def joint_mix(volume_res=(256, 512, 512)):
    """The joint-ICP session mix: 6 x 320x240, 6 x 640x480 and 4 x 1280x720 sessions with
    colour, FP16 volumes over 3 m (volume_res^3 per resolution), the joint ICP order
    (geometric plus photometric rows, weight 0.1); each frame's pipeline captured as a
    graph when the session runs alone.  Returns [(workload, descriptor)] for the 16
    sessions."""
    import dataclasses
    out = []
    for (base, count), res in zip(((SMALL, 6), (FULL, 6), (HD, 4)), volume_res):
        wl = dataclasses.replace(base, volume_res=res, volume_dim=3.0, trunc=0.025)
        d = wl.kinfu_desc(storage=nuif.TSDF_F16, gate=0.0, color=True, use_graph=True,
                          reduction=nuif.ICP_REDUCE_JOINT)
        out += [(wl, d)] * count
    return out


def tiny(width=64, height=48, res=32, frames=4):
    """A small workload for quick checks (same scene family as SMALL, scaled intrinsics)."""
    s = width / 320.0
    return Workload(f"tiny-{width}x{height}-{res}^3", width, height, 262.5 * s, 262.5 * s, (width - 1) / 2.0,
                    (height - 1) / 2.0, frames=frames, orbit_radius=1.2, volume_res=res, volume_dim=2.0,
                    trunc=0.03, iterations=[10, 5, 4], seed_index=0)
