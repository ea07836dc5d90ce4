"""Pins of the voxel-hashing oracle (CPU): closed forms and invariants chosen so
that a dropped term, a wrong sign or index, or a transposed operand fails.

  * bucket hash: K/KinfuHashUtils.cl:9-16 on signed short coordinates;
  * touched blocks (K/KinfuHashScene.cl:52-136): every touched block meets the
    pixel's truncation-band segment, and every block the segment crosses for
    longer than the sample spacing is touched (brute force in FP64);
  * integration (K/KinfuHashScene.cl:387-470): a fronto-parallel plane gives
    sdf = RTZ(32767 * min(1, (D - z) / ts)) with z the voxel's own coordinate
    (no half-voxel offset), weight 1, exactly where the projective band holds;
    the weight cap freezes a voxel;
  * raycast (K/KinfuHashRaycast.cl:155-448): the plane is hit at depth D within
    2 mm (the step-scale quirk biases it by a fraction of a voxel) with normal
    (0, 0, -1);
  * marching cubes: the vertex count equals a brute-force count of the cube
    cases over the visible bricks, and every vertex lies on a voxel edge.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O

VS, TS = 0.005, 0.02
W_, H_ = 64, 48
CAM6 = np.array([60.0, 60.0, 31.5, 23.5, 0.5, 4.5], np.float32)


class _Wl:
    width, height = W_, H_
    fx, fy, cx, cy = 60.0, 60.0, 31.5, 23.5

    @staticmethod
    def cam6():
        return CAM6


I3 = np.eye(3, dtype=np.float32)
Z3 = np.zeros(3, np.float32)


def touched(depth, R=I3, t=Z3):
    n = O.lib().or_hash_touched(depth.reshape(-1), W_, H_, CAM6, R.reshape(-1).copy(), t, VS, TS, None, 0)
    keys = np.empty((n, 3), np.int16)
    O.lib().or_hash_touched(depth.reshape(-1), W_, H_, CAM6, R.reshape(-1).copy(), t, VS, TS, keys.ctypes.data, n)
    return keys


def test_bucket_hash_closed_form():
    P = (73856093, 19349669, 83492791)
    for (x, y, z) in [(0, 0, 0), (1, 2, 3), (-1, 5, -7), (-32768, 32767, 100), (123, -456, 789)]:
        h = ((x & 0xFFFFFFFF) * P[0] ^ (y & 0xFFFFFFFF) * P[1] ^ (z & 0xFFFFFFFF) * P[2]) & 0xFFFFFFFF
        for mask in (0xFFFFF, 0x14FFFF, 0xFFFF):
            assert O.lib().or_hash_index(x, y, z, mask) == (h & mask)


def test_touched_blocks_follow_the_band_segment():
    rng = np.random.default_rng(3)
    for trial in range(6):
        depth = np.full((H_, W_), np.nan, np.float32)
        u, v = int(rng.integers(0, W_)), int(rng.integers(0, H_))
        D = float(rng.uniform(0.6, 3.5))
        depth[v, u] = D
        a = np.radians(rng.uniform(-30, 30))
        R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]], np.float32)
        t = rng.uniform(-1, 1, 3).astype(np.float32)
        keys = touched(depth, R, t)
        # FP64 segment in block units
        P = np.array([(u - CAM6[2]) / CAM6[0] * D, (v - CAM6[3]) / CAM6[1] * D, D])
        eta = TS / np.linalg.norm(P)
        B = VS * 8
        a_ = (R.astype(np.float64) @ (P * (1 - eta)) + t) / B
        b_ = (R.astype(np.float64) @ (P * (1 + eta)) + t) / B
        L = np.linalg.norm(b_ - a_)
        n = int(np.ceil(2 * L))
        spacing = L / max(n - 1, 1)
        s = np.linspace(0, 1, 20001)[:, None]
        pts = a_ + s * (b_ - a_)
        blocks, cnt = np.unique(np.floor(pts).astype(np.int64), axis=0, return_counts=True)
        chord = cnt / 20000.0 * L
        kset = {tuple(k) for k in keys.astype(np.int64)}
        e = np.full(3, 1e-4)
        near = {tuple(b) for b in np.unique(np.floor(np.concatenate([pts + e, pts - e])).astype(np.int64), axis=0)}
        assert kset <= {tuple(b) for b in blocks} | near, (trial, kset)
        for b, c in zip(blocks, chord):
            if c > spacing + 1e-3:
                assert tuple(b) in kset, (trial, b, c, spacing)
        assert 1 <= len(kset) <= n


def plane_scene(D=1.2, frames=1, wmax=255, truncation=0.06):
    s = O.OracleHashScene(voxel=VS, trunc_scale=TS, truncation=truncation, wmax=wmax, capacity=1 << 16)
    depth = np.full(W_ * H_, D, np.float32)
    for _ in range(frames):
        s.integrate(depth, _Wl, I3, Z3)
    return s


def test_plane_integration_closed_form():
    D = 1.2
    s = plane_scene(D)
    keys, sdf, w, _ = s.blocks()
    assert len(keys) > 100
    np.testing.assert_array_equal(s.visible(), keys)  # first frame: visible == allocated
    lid = np.arange(512)
    lx, ly, lz = lid % 8, (lid // 8) % 8, lid // 64
    gx = keys[:, 0:1].astype(np.int64) * 8 + lx
    gy = keys[:, 1:2].astype(np.int64) * 8 + ly
    gz = keys[:, 2:3].astype(np.int64) * 8 + lz
    z = gz * VS
    u = gx * VS * CAM6[0] / np.where(z > 0, z, 1) + CAM6[2]
    v = gy * VS * CAM6[1] / np.where(z > 0, z, 1) + CAM6[3]
    inside = (z > 0) & (u + 0.5 >= 0.02) & (u + 0.5 < W_ - 0.02) & (v + 0.5 >= 0.02) & (v + 0.5 < H_ - 0.02)
    eta = D - z
    band = inside & (eta >= -TS + 1e-6)
    clear = inside & (eta < -TS - 1e-6)
    assert band.sum() > 10000
    assert (w[band] == 1).all() and (w[clear] == 0).all() and (sdf[clear] == 0).all()
    expect = np.trunc(np.minimum(1.0, eta / TS) * 32767.0)
    assert np.abs(sdf[band] - expect[band]).max() <= 1  # float vs FP64 rounding of eta / ts
    assert (sdf[band & (eta >= TS + 1e-6)] == 32767).all()


def test_weight_cap_freezes_voxels():
    s = plane_scene(frames=7, wmax=5)
    _, sdf5, w5, _ = s.blocks()
    assert w5.max() == 5 and (w5[w5 > 0] == 5).all()
    s2 = plane_scene(frames=9, wmax=5)
    _, sdf9, w9, _ = s2.blocks()
    np.testing.assert_array_equal(sdf9, sdf5)
    np.testing.assert_array_equal(w9, w5)


@pytest.mark.parametrize("truncation", [0.02, 0.06])
def test_plane_raycast_depth_and_normal(truncation):
    """The march stops less than one voxel (p0, along the ray) behind the plane;
    each of the two refinement steps moves sdf * stepScale voxels back, with
    stepScale = 0.6 * truncation / voxel while the field falls by dz * voxel / ts per
    voxel along a ray with direction z-component dz, so the hit stays behind the
    plane by (1 - k dz)^2 p0 with k = 0.6 * truncation / ts.  A wrong step scale,
    sign or trilinear weight breaks the bound."""
    D = 1.2
    s = plane_scene(D, frames=3, truncation=truncation)
    vmap, nmap, rg, _ = s.raycast(_Wl, I3, Z3)
    hit = ~np.isnan(vmap[:, 2])
    assert hit.mean() > 0.8
    uu, vv = np.meshgrid(np.arange(W_), np.arange(H_))
    rays = np.stack([(uu - CAM6[2]) / CAM6[0], (vv - CAM6[3]) / CAM6[1], np.ones_like(uu, dtype=np.float64)], -1)
    dz = (1.0 / np.linalg.norm(rays, axis=-1)).reshape(-1)[hit]
    k = 0.6 * truncation / TS
    # p0 < 1 voxel when the steps undershoot (k dz < 1); when they overshoot, one
    # step of stepScale * sdf <= 7.2 voxels from a voxel at least 4 voxels in front
    # of the plane passes it by at most 3.2 voxels
    p0 = np.where(k * dz < 1, 1.0, 0.6 * truncation / VS - TS / VS) * VS
    bound = (1 - k * dz) ** 2 * p0 * dz + 2e-5
    err = vmap[hit, 2] - D
    assert err.min() > -2e-5 and (err <= bound).all(), (err.max(), bound.max())
    assert err.max() > 0.2 * bound.max()  # the bias is real, not a lucky exact hit
    # normals: the gradient of the voxel field; near the frustum's sides it also
    # sees the never-integrated voxels (sdf 0), so only the central pixels are exact
    central = ((uu >= 8) & (uu < W_ - 8) & (vv >= 8) & (vv < H_ - 8)).reshape(-1) & hit
    assert central.sum() > 1000
    assert np.abs(nmap[central] - np.array([0, 0, -1], np.float32)).max() < 1e-3
    # ray ranges: near bound <= D <= far bound wherever the plane was hit
    assert (rg[hit, 0] <= D).all() and (rg[hit, 1] >= D - 0.05).all()


def test_mesh_count_matches_brute_force_and_vertices_on_edges():
    s = plane_scene(1.0, frames=2)
    verts = s.mesh()
    keys, sdf, w, _ = s.blocks()
    edge, tri = O.mc_tables()
    nv = np.array([(tri[c * 16:(c + 1) * 16] >= 0).sum() for c in range(256)])
    # voxel lookup over all allocated bricks
    table = {tuple(k): i for i, k in enumerate(keys.astype(np.int64))}
    vals = sdf.astype(np.float32) / np.float32(32767.0)

    def val(x, y, z):
        i = table.get((x >> 3, y >> 3, z >> 3))
        if i is None:
            return None
        return vals[i, (x & 7) + 8 * (y & 7) + 64 * (z & 7)]

    offs = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    total = 0
    for k in s.visible().astype(np.int64):
        for lz in range(8):
            for ly in range(8):
                for lx in range(8):
                    x, y, z = k[0] * 8 + lx, k[1] * 8 + ly, k[2] * 8 + lz
                    c = [val(x + a, y + b, z + d) for a, b, d in offs]
                    if any(q is None for q in c):
                        continue
                    ci = sum(1 << i for i, q in enumerate(c) if q < 0)
                    total += nv[ci]
    assert total == len(verts) and total > 100
    g = verts.astype(np.float64) / VS
    on_grid = np.abs(g - np.round(g)) < 1e-3
    assert (on_grid.sum(1) >= 2).all()


# ---- forward render (OCL/NuiKinfuOpenCLHashScene.cpp:991-1012) ----------------------

def _rot_y(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], np.float32)


def _normals(v):
    out = np.empty_like(v)
    O.lib().or_normals_simple_f(np.ascontiguousarray(v).reshape(-1), out.reshape(-1), W_, H_, 0.3)
    return out


def _pixel_point(u, v, z):
    """Camera-space point on the ray through pixel centre (u, v) at depth z."""
    return np.array([(u - CAM6[2]) / CAM6[0] * z, (v - CAM6[3]) / CAM6[1] * z, z], np.float64)


def test_forward_render_empty_scene_moves_points_to_their_pixels():
    """Empty scene: every ray range is empty, so nothing is raycast and the output is the
    forward projection alone.  Points built on the rays of known pixels (world = R c + t)
    must land on those pixels; of two points on one pixel the nearer wins; points behind
    the camera or left of the image are dropped; colours follow their points and pixels
    that receive none keep their own colour; the normals are those of the new map."""
    rng = np.random.default_rng(11)
    R, t = _rot_y(0.1), np.array([0.1, -0.05, 0.2], np.float32)
    n = W_ * H_
    vmap = np.full((n, 3), np.nan, np.float32)
    targets = {}  # dst pixel -> (depth, source)
    src = rng.permutation(n)[:400]
    for k, s in enumerate(src):
        u, v = int(rng.integers(0, W_)), int(rng.integers(0, H_))
        z = float(rng.uniform(0.8, 3.0))
        if k % 50 == 0:
            z = -1.0  # behind the camera
        if k % 50 == 1:
            u = -3    # left of the image
        c = _pixel_point(u, v, z)
        vmap[s] = (R.astype(np.float64) @ c + t).astype(np.float32)
        if z > 0 and u >= 0:
            d = v * W_ + u
            if d not in targets or z < targets[d][0]:
                targets[d] = (z, s)
    colors = rng.integers(0, 256, (n, 4), dtype=np.uint8)
    sc = O.OracleHashScene(voxel=VS, trunc_scale=TS)
    out_v, out_n, rg, out_c, raycast = sc.forward_render(_Wl, R, t, vmap, colors)
    assert raycast == 0 and (rg[:, 0] > rg[:, 1]).all()
    exp_v = np.full((n, 3), np.nan, np.float32)
    exp_c = colors.copy()
    for d, (_, s) in targets.items():
        exp_v[d] = vmap[s]
        exp_c[d] = colors[s]
    assert len(targets) > 300
    np.testing.assert_array_equal(out_v, exp_v)
    np.testing.assert_array_equal(out_c, exp_c)
    np.testing.assert_array_equal(out_n, _normals(exp_v))


def test_forward_render_reference_transform_is_camera_to_world():
    """The reference's forwardProject_kernel projects transform(p) = R p + t of a world
    point (scaled by the voxel size, a no-op at voxel size 1); with R = I that is the
    world->camera projection at translation -t, point for point.  Its pixel is the
    saturating uint conversion of p + 0.5: a point left of the image lands in column 0."""
    rng = np.random.default_rng(5)
    n = W_ * H_
    t = np.array([0.05, 0.02, -0.1], np.float32)
    vmap = np.full((n, 3), np.nan, np.float32)
    idx = rng.permutation(n)[:300]
    for s in idx:
        c = _pixel_point(int(rng.integers(0, W_)), int(rng.integers(0, H_)), float(rng.uniform(1.0, 3.0)))
        vmap[s] = (c - t).astype(np.float32)
    sc = O.OracleHashScene(voxel=1.0, trunc_scale=TS)
    ref_v, ref_n, _, _, _ = sc.forward_render(_Wl, I3, t, vmap, reference_transform=True)
    own_v, own_n, _, _, _ = sc.forward_render(_Wl, I3, -t, vmap, reference_transform=False)
    np.testing.assert_array_equal(ref_v, own_v)
    np.testing.assert_array_equal(ref_n, own_n)
    assert (~np.isnan(ref_v[:, 0])).sum() > 250
    # saturation: a point 5 pixels left of the image, row 7
    one = np.full((n, 3), np.nan, np.float32)
    one[0] = _pixel_point(-5, 7, 2.0).astype(np.float32)
    v2, _, _, _, _ = sc.forward_render(_Wl, I3, Z3, one, reference_transform=True)
    assert np.flatnonzero(~np.isnan(v2[:, 0])).tolist() == [7 * W_]
    v3, _, _, _, _ = sc.forward_render(_Wl, I3, Z3, one, reference_transform=False)
    assert np.isnan(v3).all()


def test_forward_render_fills_holes_with_castray():
    """Nothing to forward (all-NaN map): every pixel with a non-empty ray range is
    raycast, so vertices and colours equal castRay's there and stay NaN elsewhere."""
    s = plane_scene(1.2, frames=2)
    rv, _, rg, rc = s.raycast(_Wl, I3, Z3, color=True)
    n = W_ * H_
    fv, fn, frg, fc, k = s.forward_render(_Wl, I3, Z3, np.full((n, 3), np.nan, np.float32),
                                          np.zeros((n, 4), np.uint8))
    np.testing.assert_array_equal(frg, rg)
    rays = rg[:, 0] < rg[:, 1]
    assert k == rays.sum() and rays.mean() > 0.8
    np.testing.assert_array_equal(fv[rays], rv[rays])
    np.testing.assert_array_equal(fc[rays], rc[rays])
    assert np.isnan(fv[~rays]).all()
    np.testing.assert_array_equal(fn, _normals(fv))


def test_forward_render_of_a_raycast_at_its_own_pose_is_a_fixed_point():
    """Each raycast vertex lies on its own pixel's ray, so forwarding the map at the same
    pose puts (almost) every point back on its pixel and nothing is raycast again there."""
    s = plane_scene(1.2, frames=2)
    rv, _, _, _ = s.raycast(_Wl, I3, Z3)
    hit = ~np.isnan(rv[:, 0])
    fv, _, _, _, _ = s.forward_render(_Wl, I3, Z3, rv)
    same = (fv == rv).all(axis=1)
    assert same[hit].mean() > 0.99
