"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (liboracle.so).

Used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the
checker.  The product library never loads it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
i16p = np.ctypeslib.ndpointer(np.int16, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
vp = C.c_void_p
ci, cf, cd, cll = C.c_int, C.c_float, C.c_double, C.c_longlong


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise RuntimeError(f"{_PATH} missing: run `make -C oracle`")
        _lib = C.CDLL(_PATH)
        sig = {
            "or_passing_filter_f": (None, [u16p, f32p, cll, cf, cf]),
            "or_passing_filter_d": (None, [u16p, f64p, cll, cd, cd]),
            "or_gaussian_table_f": (None, [ci, cf, f32p]),
            "or_bilateral_f": (None, [f32p, f32p, ci, ci, f32p, ci, cf, cf]),
            "or_bilateral_d": (None, [f64p, f64p, ci, ci, f64p, ci, cd, cd]),
            "or_half_sample_f": (None, [f32p, f32p, ci, ci, ci, cf]),
            "or_half_sample_d": (None, [f64p, f64p, ci, ci, ci, cd]),
            "or_depth2vertex_f": (None, [f32p, f32p, ci, ci, f32p, ci, ci, ci]),
            "or_depth2vertex_d": (None, [f64p, f64p, ci, ci, f32p, ci, ci, ci]),
            "or_normals_simple_f": (None, [f32p, f32p, ci, ci, cf]),
            "or_normals_simple_d": (None, [f64p, f64p, ci, ci, cd]),
            "or_transform_maps_f": (None, [f32p, f32p, f32p, f32p, cll, f32p, f32p]),
            "or_scale_depths_f": (None, [f32p, ci, ci, f32p]),
            "or_scale_depths_d": (None, [f64p, ci, ci, f32p]),
            "or_icp_sums_f": (None, [f32p, ci, ci, f32p, ci, f32p, f32p, f32p, f32p, f32p, f32p, cf, f64p]),
            "or_icp_sums_d": (None, [f32p, ci, ci, f64p, ci, f64p, f64p, f64p, f64p, f64p, f64p, cd, f64p]),
            "or_cholesky6": (ci, [f64p, f64p, f64p, f64p]),
            "or_det6": (cd, [f64p]),
            "or_axis_rot_d": (None, [cd, ci, f64p]),
            "or_inverse3_d": (None, [f64p, f64p]),
            "or_rodrigues_norm_d": (cd, [f64p]),
            "or_engine_create": (vp, [ci, ci, ci, f32p, ci, cf, cf, cf, cf, ci, i32p, ci, cf, cf, cf, ci]),
            "or_engine_destroy": (None, [vp]),
            "or_engine_run": (ci, [vp, u16p, vp, f64p, f64p]),
            "or_engine_volume": (None, [vp, f32p, i32p]),
            "or_engine_volume_raw": (ci, [vp, i16p]),
            "or_engine_integrated": (ci, [vp]),
            "or_engine_refmaps": (None, [vp, f32p, f32p]),
            "or_integrate_s16_f": (cll, [f32p, vp, i16p, vp, ci, cf, cf, f32p, ci, ci, f32p, f32p, ci]),
            "or_raycast_s16_f": (None, [i16p, vp, ci, cf, cf, f32p, ci, ci, f32p, f32p, f32p, f32p, vp]),
            "or_effective_trunc": (cf, [ci, cf, cf]),
            "or_trilinear_s16_f": (cf, [i16p, ci, cf, cf, f32p]),
            "or_integrate_s16_wrap_f": (cll, [f32p, vp, i16p, vp, ci, cf, cf, f32p, ci, ci, f32p, f32p, ci, i32p]),
            "or_raycast_s16_wrap_f": (None, [i16p, vp, ci, cf, cf, f32p, ci, ci, f32p, f32p, f32p, f32p, vp, i32p]),
            "or_points_dense_slice": (cll, [vp, vp, ci, ci, ci, ci, f32p, i32p, i32p, i32p, vp, vp, cll]),
            "or_engine_set_shift": (None, [vp, ci, f32p]),
            "or_engine_set_icp_mode": (None, [vp, ci]),
            "or_engine_feedback_colors": (ci, [vp, vp]),
            "or_engine_set_photometric": (None, [vp, cf, cf, cf]),
            "or_engine_stage_seconds": (None, [vp, f64p]),
            "or_engine_photometric": (cll, [vp, ci, vp]),
            "or_intensity_half_sample": (None, [f32p, f32p, ci, ci]),
            "or_photometric_row_d": (ci, [C.c_double, f64p, f64p, f32p, ci, ci, f64p, f64p, C.c_double, C.c_double,
                                          C.c_double, f64p]),
            "or_engine_volume_words": (cll, [vp, vp]),
            "or_engine_refmaps_d": (None, [vp, f64p, f64p]),
            "or_f16_from_float": (C.c_uint16, [cf]),
            "or_se3_exp": (ci, [f64p, f64p, f64p, f64p]),
            "or_vold_create": (vp, [ci, cf, cf]),
            "or_vold_destroy": (None, [vp]),
            "or_vold_integrate": (cll, [vp, u16p, ci, ci, f32p, f32p, f32p]),
            "or_vold_volume": (None, [vp, f64p, i32p]),
            "or_vold_raycast": (None, [vp, ci, ci, f32p, f32p, f32p, f64p, f64p]),
            "or_se3_coefficients": (None, [cd, f64p]),
            "or_ldlt6_solve": (ci, [f64p, f64p, f64p, C.POINTER(cd)]),
            "or_gn_step": (ci, [f64p, cd, f64p, f64p]),
            "or_icp_sums_comp_f": (None, [f32p, ci, ci, f32p, ci, f32p, f32p, f32p, f32p, f32p, f32p, cf, f64p]),
            "or_f16_to_float": (cf, [C.c_uint16]),
            "or_engine_offset": (None, [vp, i32p, f64p]),
            "or_engine_cache": (cll, [vp, vp, vp, cll]),
            "or_mc_count_s16": (cll, [i16p, ci]),
            "or_mc_emit_s16": (cll, [i16p, vp, ci, cf, f32p, vp, cll]),
            "or_mc_mesh_dedup_s16": (cll, [i16p, ci, cf, f32p, cll, i32p, cll, C.POINTER(cll)]),
            "or_mc_table": (None, [np.ctypeslib.ndpointer(np.int8), np.ctypeslib.ndpointer(np.uint8)]),
            "or_hash_touched": (ci, [f32p, ci, ci, f32p, f32p, f32p, cf, cf, vp, cll]),
            "or_hash_index": (C.c_uint, [ci, ci, ci, C.c_uint]),
            "or_hash_engine_create": (vp, [ci, ci, f32p, ci, cf, cf, cf, cf, ci, i32p, cf, cf, cf, cf, ci, ci, ci]),
            "or_hash_engine_destroy": (None, [vp]),
            "or_hash_engine_run": (ci, [vp, u16p, vp, f64p, f64p]),
            "or_hash_engine_scene": (vp, [vp]),
            "or_hash_engine_refmaps": (None, [vp, vp, vp]),
            "or_hash_scene_create": (vp, [cf, cf, cf, ci, ci]),
            "or_hash_scene_destroy": (None, [vp]),
            "or_hash_scene_integrate": (cll, [vp, f32p, vp, ci, ci, f32p, f32p, f32p]),
            "or_hash_scene_raycast": (None, [vp, ci, ci, f32p, f32p, f32p, vp, vp, vp, vp]),
            "or_hash_scene_forward_render": (cll, [vp, ci, ci, f32p, f32p, f32p, vp, vp, vp, vp, ci]),
            "or_hash_scene_blocks": (ci, [vp, vp, vp, vp, vp, cll]),
            "or_hash_scene_visible": (ci, [vp, vp, cll]),
            "or_hash_scene_mesh": (cll, [vp, i32p, i32p, vp, cll]),
            "or_hash_scene_set_swapping": (None, [vp, ci]),
            "or_hash_engine_set_swapping": (None, [vp, ci]),
            "or_hash_scene_swap_states": (ci, [vp, vp, vp, cll]),
            "or_hash_scene_store": (ci, [vp, vp, vp, vp, vp, cll]),
            "or_color_half_sample": (None, [u8p, u8p, ci, ci, ci]),
            "or_color_gradient": (None, [u8p, u8p, u8p, ci, ci, ci]),
            "or_color_difference": (None, [ci, ci, f32p, ci, f32p, u8p, u8p, f32p, f32p,
                                           np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")]),
            "or_color_terms": (None, [ci, ci, f32p, ci, f32p, u8p, u8p, u8p, u8p, f32p, f32p, ci, ci, f32p]),
            "or_color_reduce": (ci, [ci, ci, f32p, ci, f32p, u8p, u8p, u8p, u8p, f32p, f32p, ci, ci, vp, ci]),
            "or_color_icp": (ci, [ci, ci, f32p, ci, i32p, i32p, cf, f32p, u8p, u8p, f32p, f32p, ci, f64p]),
            "or_bgra_to_intensity": (None, [u8p, f32p, cll]),
            "or_intensity_derivatives": (None, [f32p, f32p, ci, ci]),
            "or_eval_phis": (None, [f32p, f32p, f32p]),
            "or_intensity_icp_sums": (None, [ci, ci, f32p, ci, f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p,
                                             f32p, f32p, f32p]),
            "or_points_dense": (cll, [vp, vp, ci, ci, ci, ci, f32p, vp, vp, cll]),
            "or_points_hash": (cll, [ci, i32p, i32p, vp, vp, ci, cf, cf, vp, vp, cll]),
            "or_half_to_float": (cf, [C.c_uint16]),
            "or_expf_det": (cf, [cf]),
            "or_sincosf_det": (None, [cf, C.POINTER(cf), C.POINTER(cf)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


class OracleEngine:
    """RunTracking of the reference over a dense scene, on the CPU.

    prec 0: float arithmetic with the volume storage `storage` (0: the reference's
    short2, 1: FP16 tsdf + uint16 weight, 2: FP32 tsdf + FP32 weight); prec 1: FP64 oracle.
    compensated: the compensated ICP mode (Gauss-Newton, FP64 pose) instead of the
    reference's loop.
    """

    def __init__(self, wl, prec=0, gate=0.0, color=False, iterations=None, compensated=False, storage=0,
                 joint=False):
        its = np.array(list(iterations or wl.iterations) + [0] * 8, np.int32)[:8]
        self.wl = wl
        self.prec = prec
        self.storage = storage if prec == 0 else None
        code = 1 if prec == 1 else {0: 0, 1: 2, 2: 3}[storage]
        self.h = lib().or_engine_create(code, wl.width, wl.height, wl.cam6(), 3, 0.08, 50.0, 0.3, 0.1,
                                        len(iterations or wl.iterations), its, wl.volume_res, wl.volume_dim,
                                        wl.trunc, gate, 1 if color else 0)
        if joint:  # the joint order: photometric rows (weight 0.1) beside the compensated sums
            assert color, "the joint order needs colours"
            lib().or_engine_set_icp_mode(self.h, 2)
        elif compensated:
            lib().or_engine_set_icp_mode(self.h, 1)

    def run(self, depth, bgra=None):
        pose = np.zeros(12, np.float64)
        info = np.zeros(3, np.float64)
        ok = lib().or_engine_run(self.h, np.ascontiguousarray(depth, np.uint16),
                                 None if bgra is None else np.ascontiguousarray(bgra).ctypes.data, pose, info)
        return ok, pose, info

    def integrated(self):
        """Whether the last frame integrated (first frame, or the gate passed)."""
        return lib().or_engine_integrated(self.h) == 1

    def volume(self):
        n = self.wl.volume_res ** 3
        t = np.empty(n, np.float32)
        w = np.empty(n, np.int32)
        lib().or_engine_volume(self.h, t, w)
        return t, w

    def volume_words(self):
        """The float engine's volume in the GPU's storage words (uint8 bytes)."""
        n = lib().or_engine_volume_words(self.h, None)
        out = np.empty(max(n, 0), np.uint8)
        if n > 0:
            lib().or_engine_volume_words(self.h, out.ctypes.data)
        return out

    def stage_seconds(self):
        """Host seconds per stage (preprocess, ICP, integration, raycast) summed over the
        frames run since the last call."""
        out = np.zeros(4, np.float64)
        lib().or_engine_stage_seconds(self.h, out)
        return out

    def set_photometric(self, weight, max_difference=0.2, min_gradient=0.002):
        lib().or_engine_set_photometric(self.h, weight, max_difference, min_gradient)

    def feedback_colors(self):
        """The feedback colours (H*W x 4 BGRA) after the last frame, or None."""
        out = np.empty((self.wl.width * self.wl.height, 4), np.uint8)
        return out if lib().or_engine_feedback_colors(self.h, out.ctypes.data) else None

    def photometric(self, which):
        """The joint order's inputs of the last tracked frame: which = level l (0..3) of
        the intensity pyramid, -1 the feedback intensities, 4 their gradients."""
        n = lib().or_engine_photometric(self.h, which, None)
        out = np.empty(n, np.float32)
        if n:
            lib().or_engine_photometric(self.h, which, out.ctypes.data)
        return out

    def refmaps_d(self):
        n = self.wl.width * self.wl.height * 3
        v = np.empty(n, np.float64)
        nm = np.empty(n, np.float64)
        lib().or_engine_refmaps_d(self.h, v, nm)
        return v.reshape(-1, 3), nm.reshape(-1, 3)

    def volume_raw(self):
        n = self.wl.volume_res ** 3
        out = np.empty(2 * n, np.int16)
        lib().or_engine_volume_raw(self.h, out)
        return out

    def set_shift(self, voxel_shift, basis=(0.0, 0.0, 0.0)):
        """Shifting volume (OCL/NuiKinfuOpenCLShiftScene.cpp); before the first frame."""
        lib().or_engine_set_shift(self.h, voxel_shift, np.asarray(basis, np.float32))

    def offset(self):
        """(voxel offset int[3], its size in metres float64[3])."""
        o = np.zeros(3, np.int32)
        s = np.zeros(3, np.float64)
        lib().or_engine_offset(self.h, o, s)
        return o, s

    def cache(self):
        """The points evicted by the shifts so far: (points float32[n,3], colours float32[n,4])."""
        n = lib().or_engine_cache(self.h, None, None, 0)
        p = np.empty((max(n, 1), 3), np.float32)
        c = np.empty((max(n, 1), 4), np.float32)
        lib().or_engine_cache(self.h, p.ctypes.data, c.ctypes.data, n)
        return p[:n], c[:n]

    def refmaps(self):
        n = self.wl.width * self.wl.height * 3
        v = np.empty(n, np.float32)
        nm = np.empty(n, np.float32)
        lib().or_engine_refmaps(self.h, v, nm)
        return v.reshape(-1, 3), nm.reshape(-1, 3)

    def close(self):
        if self.h:
            lib().or_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class OracleVolumeD:
    """An FP64 dense volume integrated and raycast at given (float) poses: the precision
    reference of a GPU session's TSDF and raycast (the same frames and poses in double)."""

    def __init__(self, wl):
        self.wl = wl
        self.h = lib().or_vold_create(wl.volume_res, wl.volume_dim, wl.trunc)

    def integrate(self, depth, R, t):
        return lib().or_vold_integrate(self.h, np.ascontiguousarray(depth, np.uint16).reshape(-1), self.wl.width,
                                       self.wl.height, self.wl.cam6(), np.asarray(R, np.float32).reshape(-1).copy(),
                                       np.asarray(t, np.float32).reshape(-1).copy())

    def volume(self):
        n = self.wl.volume_res ** 3
        t = np.empty(n, np.float64)
        w = np.empty(n, np.int32)
        lib().or_vold_volume(self.h, t, w)
        return t, w

    def raycast(self, R, t):
        n = self.wl.width * self.wl.height * 3
        v = np.empty(n, np.float64)
        nm = np.empty(n, np.float64)
        lib().or_vold_raycast(self.h, self.wl.width, self.wl.height, self.wl.cam6(),
                              np.asarray(R, np.float32).reshape(-1).copy(), np.asarray(t, np.float32).reshape(-1).copy(),
                              v, nm)
        return v.reshape(-1, 3), nm.reshape(-1, 3)

    def close(self):
        if self.h:
            lib().or_vold_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mc_tables():
    """(edge_table int32[256], tri_table int32[256*16]) of the standard marching-cubes
    tables; the edge mask of a cube is the set of edges its triangles use."""
    tri = np.empty(256 * 16, np.int8)
    nv = np.empty(256, np.uint8)
    lib().or_mc_table(tri, nv)
    tri = tri.astype(np.int32)
    edge = np.zeros(256, np.int32)
    for c in range(256):
        for e in tri[c * 16:(c + 1) * 16]:
            if e < 0:
                break
            edge[c] |= 1 << int(e)
    return edge, tri


class _HashSceneView:
    """Read-back of a hash scene held by the oracle (blocks, visible list, mesh)."""

    def _scene(self):
        raise NotImplementedError

    def blocks(self, voxels=True):
        """(keys int16[n,3] sorted by (x,y,z), sdf int16[n,512], weight uint8[n,512], rgb uint8[n,512,3])."""
        sp = self._scene()
        n = lib().or_hash_scene_blocks(sp, None, None, None, None, 0)
        keys = np.empty((n, 3), np.int16)
        if not voxels:
            lib().or_hash_scene_blocks(sp, keys.ctypes.data, None, None, None, n)
            return keys
        sdf = np.empty((n, 512), np.int16)
        w = np.empty((n, 512), np.uint8)
        rgb = np.empty((n, 512, 3), np.uint8)
        lib().or_hash_scene_blocks(sp, keys.ctypes.data, sdf.ctypes.data, w.ctypes.data, rgb.ctypes.data, n)
        return keys, sdf, w, rgb

    def swap_states(self):
        """(keys int16[n,3] sorted by (x,y,z), swap state uint8[n]) of every block with an entry."""
        sp = self._scene()
        n = lib().or_hash_scene_swap_states(sp, None, None, 0)
        keys = np.empty((n, 3), np.int16)
        st = np.empty(n, np.uint8)
        lib().or_hash_scene_swap_states(sp, keys.ctypes.data, st.ctypes.data, n)
        return keys, st

    def store(self):
        """The host store of swapped-out blocks: (keys, sdf int16[n,512], weight uint8[n,512], rgb)."""
        sp = self._scene()
        n = lib().or_hash_scene_store(sp, None, None, None, None, 0)
        keys = np.empty((n, 3), np.int16)
        sdf = np.empty((n, 512), np.int16)
        w = np.empty((n, 512), np.uint8)
        rgb = np.empty((n, 512, 3), np.uint8)
        lib().or_hash_scene_store(sp, keys.ctypes.data, sdf.ctypes.data, w.ctypes.data, rgb.ctypes.data, n)
        return keys, sdf, w, rgb

    def visible(self):
        sp = self._scene()
        n = lib().or_hash_scene_visible(sp, None, 0)
        keys = np.empty((n, 3), np.int16)
        lib().or_hash_scene_visible(sp, keys.ctypes.data, n)
        return keys

    def mesh(self):
        edge, tri = mc_tables()
        sp = self._scene()
        n = lib().or_hash_scene_mesh(sp, edge, tri, None, 0)
        out = np.empty((max(n, 1), 3), np.float32)
        lib().or_hash_scene_mesh(sp, edge, tri, out.ctypes.data, n)
        return out[:n]


class OracleHashScene(_HashSceneView):
    """The reference's voxel-hashing scene on the CPU, driven with explicit poses."""

    def __init__(self, voxel=0.005, trunc_scale=0.02, truncation=0.06, wmax=255, capacity=1 << 18, swapping=False):
        self.h = lib().or_hash_scene_create(voxel, trunc_scale, truncation, wmax, capacity)
        if swapping:
            lib().or_hash_scene_set_swapping(self.h, 1)

    def _scene(self):
        return self.h

    def integrate(self, depth, wl, R, t, bgra=None):
        return lib().or_hash_scene_integrate(self.h, np.ascontiguousarray(depth, np.float32),
                                             None if bgra is None else np.ascontiguousarray(bgra).ctypes.data,
                                             wl.width, wl.height, wl.cam6(),
                                             np.ascontiguousarray(R, np.float32).reshape(-1),
                                             np.ascontiguousarray(t, np.float32))

    def raycast(self, wl, R, t, color=False):
        n = wl.width * wl.height
        v = np.empty((n, 3), np.float32)
        nm = np.empty((n, 3), np.float32)
        rg = np.empty((n, 2), np.float32)
        cm = np.empty((n, 4), np.uint8) if color else None
        lib().or_hash_scene_raycast(self.h, wl.width, wl.height, wl.cam6(),
                                    np.ascontiguousarray(R, np.float32).reshape(-1),
                                    np.ascontiguousarray(t, np.float32), v.ctypes.data, nm.ctypes.data,
                                    rg.ctypes.data, None if cm is None else cm.ctypes.data)
        return v, nm, rg, cm

    def forward_render(self, wl, R, t, vmap, cmap=None, reference_transform=False):
        """Forward render from the previous world vertex map `vmap` (and colours `cmap`)
        at pose (R, t).  Returns (vmap, nmap, range, cmap, pixels raycast)."""
        n = wl.width * wl.height
        v = np.array(vmap, np.float32).reshape(n, 3)
        nm = np.empty((n, 3), np.float32)
        rg = np.empty((n, 2), np.float32)
        cm = None if cmap is None else np.array(cmap, np.uint8).reshape(n, 4)
        k = lib().or_hash_scene_forward_render(self.h, wl.width, wl.height, wl.cam6(),
                                               np.ascontiguousarray(R, np.float32).reshape(-1),
                                               np.ascontiguousarray(t, np.float32), v.ctypes.data, nm.ctypes.data,
                                               rg.ctypes.data, None if cm is None else cm.ctypes.data,
                                               int(bool(reference_transform)))
        return v, nm, rg, cm, int(k)

    def close(self):
        if self.h:
            lib().or_hash_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class OracleHashEngine(_HashSceneView):
    """RunTracking of the reference over a voxel-hashing scene, float arithmetic,
    deterministic allocation (every touched block in the frame that touches it).
    forward: 0 renders with castRay, 1 with the forward render, 2 with the forward
    render and the reference's camera->world projection."""

    def __init__(self, wl, gate=0.0, voxel=0.005, trunc_scale=0.02, truncation=0.06, wmax=255, capacity=1 << 18,
                 iterations=None, forward=0, swapping=False):
        its = list(iterations or wl.iterations)
        self.wl = wl
        self.h = lib().or_hash_engine_create(wl.width, wl.height, wl.cam6(), 3, 0.08, 50.0, 0.3, 0.1, len(its),
                                             np.array(its + [0] * 8, np.int32)[:8], gate, voxel, trunc_scale,
                                             truncation, wmax, capacity, int(forward))
        if swapping:
            lib().or_hash_engine_set_swapping(self.h, 1)

    def _scene(self):
        return lib().or_hash_engine_scene(self.h)

    def run(self, depth, bgra=None):
        pose = np.zeros(12, np.float64)
        info = np.zeros(3, np.float64)
        ok = lib().or_hash_engine_run(self.h, np.ascontiguousarray(depth, np.uint16),
                                      None if bgra is None else np.ascontiguousarray(bgra).ctypes.data, pose, info)
        return ok, pose, info

    def refmaps(self):
        n = self.wl.width * self.wl.height
        v = np.empty((n, 3), np.float32)
        nm = np.empty((n, 3), np.float32)
        lib().or_hash_engine_refmaps(self.h, v.ctypes.data, nm.ctypes.data)
        return v, nm

    def close(self):
        if self.h:
            lib().or_hash_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- photometric (colour) tracker -------------------------------------------------

def _u8(a):
    return np.ascontiguousarray(a, np.uint8).reshape(-1)


def _f32(a):
    return np.ascontiguousarray(a, np.float32).reshape(-1)


def color_half_sample(src, w_dst, h_dst, radius=1):
    """half_sample_bgra_kernel: BGRA (2w x 2h) -> (w x h)."""
    dst = np.empty(w_dst * h_dst * 4, np.uint8)
    lib().or_color_half_sample(_u8(src), dst, w_dst, h_dst, radius)
    return dst.reshape(h_dst, w_dst, 4)


def color_gradient(bgra, w, h, signed=False):
    """bgra_gradient_kernel: (gx, gy) as uint8 (or int8 values stored in uint8)."""
    gx = np.empty(w * h * 4, np.uint8)
    gy = np.empty(w * h * 4, np.uint8)
    lib().or_color_gradient(_u8(bgra), gx, gy, w, h, int(bool(signed)))
    return gx.reshape(h, w, 4), gy.reshape(h, w, 4)


def color_difference(wl, level, vmap, colors, level_colors, R, t):
    """GetColorsDifference's exact sums: (sum |obs - known|^2, valid count)."""
    out = np.zeros(2, np.int64)
    lib().or_color_difference(wl.width, wl.height, wl.cam6(), level, _f32(vmap), _u8(colors), _u8(level_colors),
                              _f32(R), _f32(t), out)
    return int(out[0]), int(out[1])


def color_terms(wl, level, p, known, level_colors, gx, gy, R, t, rot_only=False, signed=False):
    out = np.empty(29, np.float32)
    lib().or_color_terms(wl.width, wl.height, wl.cam6(), level, _f32(p), _u8(known), _u8(level_colors), _u8(gx),
                         _u8(gy), _f32(R), _f32(t), int(bool(rot_only)), int(bool(signed)), out)
    return out


def color_reduce(wl, level, vmap, colors, level_colors, gx, gy, R, t, rot_only=False, signed=False):
    """One iteration's second-level block sums (nblocks x 29), in the reference's tree order."""
    args = (wl.width, wl.height, wl.cam6(), level, _f32(vmap), _u8(colors), _u8(level_colors), _u8(gx), _u8(gy),
            _f32(R), _f32(t), int(bool(rot_only)), int(bool(signed)))
    nb = lib().or_color_reduce(*args, None, 0)
    out = np.empty((nb, 29), np.float32)
    lib().or_color_reduce(*args, out.ctypes.data, nb)
    return out


def color_icp(wl, vmap, colors, frame_bgra, R, t, iterations=None, rot_only=None, dist_threshold=0.1,
              corrected=False):
    """ColorIterativeClosestPoint from pose (R, t): returns (ok, R, t, iterations, error, count)."""
    its = list(iterations or wl.iterations)
    ro = list(rot_only or [0] * len(its))
    Rn = np.array(R, np.float32).reshape(-1).copy()
    tn = np.array(t, np.float32).reshape(-1).copy()
    info = np.zeros(4, np.float64)
    ok = lib().or_color_icp(wl.width, wl.height, wl.cam6(), len(its), np.array(its, np.int32),
                            np.array(ro, np.int32), dist_threshold, _f32(vmap), _u8(colors), _u8(frame_bgra), Rn, tn,
                            int(bool(corrected)), info)
    return ok, Rn.reshape(3, 3), tn, int(info[1]), float(info[2]), float(info[3])


# ---- joint depth + intensity ICP terms ------------------------------------------

def intensity_half_sample(src, wd, hd):
    """Intensities one level down (mean of the valid ones of each 2 x 2 block)."""
    out = np.empty(wd * hd, np.float32)
    lib().or_intensity_half_sample(_f32(src), out, wd, hd)
    return out


def photometric_row_d(ic, pv, pp, px, py, Rinv, cam6, w, h, iref, gref, weight, pdiff, min_grad):
    """The joint order's photometric row of one pixel in double: (ok, c[7])."""
    at = np.array(list(pv) + list(pp) + [px, py], np.float64)
    out = np.zeros(7, np.float64)
    ok = lib().or_photometric_row_d(float(ic), at, np.ascontiguousarray(Rinv, np.float64).ravel(),
                                    np.ascontiguousarray(cam6, np.float32), w, h,
                                    np.ascontiguousarray(iref, np.float64).ravel(),
                                    np.ascontiguousarray(gref, np.float64).ravel(), float(np.sqrt(weight)),
                                    float(pdiff), float(min_grad) ** 2, out)
    return ok == 1, out

def bgra_to_intensity(bgra):
    b = _u8(bgra)
    out = np.empty(b.size // 4, np.float32)
    lib().or_bgra_to_intensity(b, out, out.size)
    return out


def intensity_derivatives(I, w, h):
    out = np.empty(w * h * 2, np.float32)
    lib().or_intensity_derivatives(_f32(I), out, w, h)
    return out.reshape(h, w, 2)


def eval_phis(euler, p):
    out = np.empty(9, np.float32)
    lib().or_eval_phis(_f32(euler), _f32(p), out)
    return out.reshape(3, 3)


def intensity_icp_sums(wl, level, v, n, I, vp, np_, Ip, dp, R, t, Rp, tp, euler, thr):
    """The joint kernel's 29 sums (the reference's reduction order) at `level`."""
    out = np.empty(29, np.float32)
    pose = np.concatenate([np.asarray(R, np.float32).reshape(-1), np.asarray(t, np.float32).reshape(-1)])
    prev = np.concatenate([np.asarray(Rp, np.float32).reshape(-1), np.asarray(tp, np.float32).reshape(-1)])
    lib().or_intensity_icp_sums(wl.width, wl.height, wl.cam6(), level, _f32(v), _f32(n), _f32(I), _f32(vp),
                                _f32(np_), _f32(Ip), _f32(dp), pose, prev, _f32(euler), _f32(thr), out)
    return out


# ---- point clouds -------------------------------------------------------------------

def points_dense(words, color_volume, storage, res, cell, cap=None):
    """fetchVolumeKernel: (points float32[n,3], colours float32[n,4], total) in (y, x, z)
    column order.  words: the raw volume (uint32 per voxel for storages 0/1, uint32 pairs
    for storage 2), res: (Nx, Ny, Nz), cell: (cx, cy, cz)."""
    words = np.ascontiguousarray(words)
    cv = None if color_volume is None else np.ascontiguousarray(color_volume, np.uint8)
    Nx, Ny, Nz = (int(r) for r in res)
    cell = np.asarray(cell, np.float32)
    total = lib().or_points_dense(words.ctypes.data, None if cv is None else cv.ctypes.data, storage, Nx, Ny, Nz,
                                  cell, None, None, 0)
    n = total if cap is None else min(cap, total)
    pts = np.empty((max(n, 1), 3), np.float32)
    cols = np.empty((max(n, 1), 4), np.float32)
    lib().or_points_dense(words.ctypes.data, None if cv is None else cv.ctypes.data, storage, Nx, Ny, Nz, cell,
                          pts.ctypes.data, cols.ctypes.data, n)
    return pts[:n], cols[:n], total


def points_hash(pos, ptr, bricks, colors, storage, voxel_size, trunc_scale, cap=None):
    """fetchHashScene_kernel over the visible blocks in list order: pos int32[n,3],
    ptr int32[n], bricks uint32[nbricks*512], colors uint8[nbricks*512*4] or None."""
    pos = np.ascontiguousarray(pos, np.int32).reshape(-1)
    ptr = np.ascontiguousarray(ptr, np.int32).reshape(-1)
    br = np.ascontiguousarray(bricks, np.uint32).reshape(-1)
    cv = None if colors is None else np.ascontiguousarray(colors, np.uint8).reshape(-1)
    args = (ptr.size, pos, ptr, br.ctypes.data, None if cv is None else cv.ctypes.data, storage, voxel_size,
            trunc_scale)
    total = lib().or_points_hash(*args, None, None, 0)
    n = total if cap is None else min(cap, total)
    pts = np.empty((max(n, 1), 3), np.float32)
    cols = np.empty((max(n, 1), 4), np.float32)
    lib().or_points_hash(*args, pts.ctypes.data, cols.ctypes.data, n)
    return pts[:n], cols[:n], total


def points_dense_slice(words, color_volume, storage, res, cell, wrap, offset, rng):
    """fetchShiftVolumeKernel (K/extract.cl:80-190): (points, colours, total) in (y, x, z) order."""
    words = np.ascontiguousarray(words)
    cv = None if color_volume is None else np.ascontiguousarray(color_volume, np.uint8)
    Nx, Ny, Nz = (int(r) for r in res)
    args = (words.ctypes.data, None if cv is None else cv.ctypes.data, storage, Nx, Ny, Nz,
            np.asarray(cell, np.float32), np.asarray(wrap, np.int32), np.asarray(offset, np.int32),
            np.asarray(rng, np.int32))
    n = lib().or_points_dense_slice(*args, None, None, 0)
    pts = np.empty((max(n, 1), 3), np.float32)
    cols = np.empty((max(n, 1), 4), np.float32)
    lib().or_points_dense_slice(*args, pts.ctypes.data, cols.ctypes.data, n)
    return pts[:n], cols[:n], n


def map_depth_to_color(map_xy, lut, image):
    """OCL/NuiKinfuOpenCLFrame.cpp:405-448 UpdateColorBuffers(ColorSpacePoint*, UINT*, UINT,
    const NuiColorImage&), restated in numpy: depth pixel i reads the map at j = lut[i] (i
    without a table); for j < n the colour pixel ((int)(x + 0.5f), (int)(y + 0.5f)) -- float
    addition, then truncation toward zero -- is taken with alpha 255 when it lies in the
    image, else the pixel is 0.  map_xy: (n, 2) float32, lut: (n,) uint32 or None,
    image: (ih, iw, 4) uint8 BGRA.  Returns (n, 4) uint8."""
    map_xy = np.asarray(map_xy, np.float32)
    n = map_xy.shape[0]
    ih, iw = image.shape[:2]
    j = np.arange(n, dtype=np.int64) if lut is None else np.asarray(lut, np.uint32).astype(np.int64)
    ok = j < n
    q = map_xy[np.where(ok, j, 0)]
    cx = np.trunc(q[:, 0] + np.float32(0.5)).astype(np.int64)
    cy = np.trunc(q[:, 1] + np.float32(0.5)).astype(np.int64)
    inside = ok & (cx >= 0) & (cx < iw) & (cy >= 0) & (cy < ih)
    out = np.zeros((n, 4), np.uint8)
    flat = image.reshape(-1, 4)
    out[inside] = flat[cy[inside] * iw + cx[inside]]
    out[inside, 3] = 255
    return out
