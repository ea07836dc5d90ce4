"""CPU pins of oracle functions that no closed form covered so far:

  * the dense colour running mean (K/tsdfVolume.cl:154-181): per channel
    (old * w + new) / (w + 1) in float, truncated toward zero, channels in the input's
    order, weight capped at the maximum and incremented as an unsigned char (it wraps);
  * the hash colour update (K/KinfuHashScene.cl:445-461): colour[0..2] <- the input's
    s2, s1, s0 (BGRA in, RGB stored), only for alpha 255, inside the 0.25 band of the
    normalised sdf, running mean over the voxel's weight, truncated;
  * the visible-list re-check (K/KinfuHashScene.cl:211-285): a block that was visible
    and that the new frame does not touch stays visible while any of its 8 corners
    projects into the image;
  * the trilinear interpolation (K/raycast.cl:36-120): its corner names follow an
    x-major layout over the y-major volume, so the x and y fractions weight swapped
    corners; a field linear in x is interpolated with the y fraction;
  * the depth-to-colour mapping (OCL/NuiKinfuOpenCLFrame.cpp:405-448): (int)(x + 0.5f)
    truncates toward zero, so x in (-1.5, -0.5) still lands on column 0.
Each is built so that the plausible slip (rounding to nearest, no channel swap, no
re-check, the textbook corner weights) fails.
"""
import numpy as np
import pytest

import oracle as O


def lib():
    return O.lib()


# ---------------------------------------------------------------- dense colour

def _colour_plane(frames, colours, N=32, dim=2.0, trunc=0.05, D=1.2, maxcw=128):
    W_, H_ = 160, 120
    cam = np.array([131.25, 131.25, 79.5, 59.5, 0.5, 4.5], np.float32)
    depth = np.full((H_, W_), D, np.float32)
    lib().or_scale_depths_f(depth, W_, H_, cam)
    vol = np.zeros((N ** 3, 2), np.int16)
    vol[:, 0] = 32767
    cvol = np.zeros((N ** 3, 4), np.uint8)
    I = np.eye(3, dtype=np.float32).reshape(-1)
    t0 = np.zeros(3, np.float32)
    for f in range(frames):
        bgra = np.empty((H_, W_, 4), np.uint8)
        bgra[...] = np.array(colours[f % len(colours)], np.uint8)
        lib().or_integrate_s16_f(depth, bgra.ctypes.data, vol.reshape(-1), cvol.ctypes.data, N, dim, trunc, cam, W_,
                                 H_, I, t0, maxcw)
    return vol, cvol


def test_dense_colour_running_mean_truncates_and_keeps_channel_order():
    A, B = (10, 100, 201, 255), (21, 51, 0, 255)
    _, c1 = _colour_plane(1, [A])
    hit = c1[:, 3] > 0
    assert hit.sum() > 500
    np.testing.assert_array_equal(c1[hit], np.tile([10, 100, 201, 1], (hit.sum(), 1)))
    _, c2 = _colour_plane(2, [A, B])
    np.testing.assert_array_equal(c2[:, 3] > 0, hit)
    # (10 + 21) / 2 = 15.5 -> 15, (100 + 51) / 2 = 75.5 -> 75, (201 + 0) / 2 = 100.5 -> 100
    np.testing.assert_array_equal(c2[hit], np.tile([15, 75, 100, 2], (hit.sum(), 1)))
    # a third colour: (15 * 2 + 40) / 3 = 23.33 -> 23, (75 * 2 + 0) / 3 = 50, (100 * 2 + 255) / 3 = 151.67 -> 151
    _, c3 = _colour_plane(3, [A, B, (40, 0, 255, 255)])
    np.testing.assert_array_equal(c3[hit], np.tile([23, 50, 151, 3], (hit.sum(), 1)))


def test_dense_colour_skips_transparent_pixels_and_colours_only_the_band():
    vol, c = _colour_plane(1, [(50, 60, 70, 0)])
    assert (c == 0).all()  # alpha 0: no colour update
    w = vol[:, 1]
    _, c1 = _colour_plane(1, [(50, 60, 70, 255)])
    hit = c1[:, 3] > 0
    # colour only where |sdf| < trunc: a strict subset of the integrated voxels (the
    # free space in front of the band integrates the tsdf 1.0 without colour)
    assert (w[hit] > 0).all() and hit.sum() < (w > 0).sum()


def test_dense_colour_weight_cap_and_unsigned_char_wrap():
    _, c = _colour_plane(9, [(10, 20, 30, 255)], maxcw=5)
    hit = c[:, 3] > 0
    assert set(np.unique(c[hit, 3])) == {5}
    # a maximum above 255: the weight is incremented as an unsigned char, 255 + 1 -> 0
    _, c255 = _colour_plane(255, [(10, 20, 30, 255)], maxcw=1000)
    assert set(np.unique(c255[hit, 3])) == {255}
    _, c256 = _colour_plane(256, [(10, 20, 30, 255)], maxcw=1000)
    assert set(np.unique(c256[hit, 3])) == {0}
    np.testing.assert_array_equal(c256[hit, :3], c255[hit, :3])


# ---------------------------------------------------------------- hash colour

VS, TS = 0.005, 0.02
W_, H_ = 64, 48
CAM6 = np.array([60.0, 60.0, 31.5, 23.5, 0.5, 4.5], np.float32)
I3 = np.eye(3, dtype=np.float32)
Z3 = np.zeros(3, np.float32)


class _Wl:
    width, height = W_, H_

    @staticmethod
    def cam6():
        return CAM6


def _hash_colour_scene(colours, D=1.2):
    s = O.OracleHashScene(voxel=VS, trunc_scale=TS, truncation=0.06, wmax=255, capacity=1 << 16)
    depth = np.full(W_ * H_, D, np.float32)
    for col in colours:
        bgra = np.empty((H_, W_, 4), np.uint8)
        bgra[...] = np.array(col, np.uint8)
        s.integrate(depth, _Wl, I3, Z3, bgra=bgra)
    return s


def _band_masks(keys, D=1.2):
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
    sdf = eta / TS
    colour = inside & (np.abs(sdf) < 0.25 - 1e-4) & (eta < TS)
    plain = inside & (eta >= -TS + 1e-6) & (np.abs(sdf) > 0.25 + 1e-4)
    return colour, plain


def test_hash_colour_swaps_channels_and_keeps_the_quarter_band():
    s = _hash_colour_scene([(10, 100, 200, 255)])
    keys, _, w, rgb = s.blocks()
    colour, plain = _band_masks(keys)
    assert colour.sum() > 1000 and plain.sum() > 1000
    # BGRA (10, 100, 200) stored as RGB (200, 100, 10)
    np.testing.assert_array_equal(rgb[colour], np.tile([200, 100, 10], (colour.sum(), 1)))
    assert (rgb[plain] == 0).all() and (w[plain] == 1).all()


def test_hash_colour_running_mean_and_alpha():
    s = _hash_colour_scene([(10, 100, 200, 255), (21, 51, 0, 255)])
    keys, _, _, rgb = s.blocks()
    colour, _ = _band_masks(keys)
    # R (200 + 0) / 2 = 100, G (100 + 51) / 2 = 75.5 -> 75, B (10 + 21) / 2 = 15.5 -> 15
    np.testing.assert_array_equal(rgb[colour], np.tile([100, 75, 15], (colour.sum(), 1)))
    # alpha below 255 leaves the colour as it was (the weight still grows)
    s2 = _hash_colour_scene([(10, 100, 200, 255), (21, 51, 0, 254)])
    _, _, w2, rgb2 = s2.blocks()
    np.testing.assert_array_equal(rgb2[colour], np.tile([200, 100, 10], (colour.sum(), 1)))
    assert (w2[colour] == 2).all()


# ---------------------------------------------------------------- visible-list re-check

def _touched(depth):
    n = lib().or_hash_touched(depth.reshape(-1), W_, H_, CAM6, I3.reshape(-1).copy(), Z3, VS, TS, None, 0)
    keys = np.empty((n, 3), np.int16)
    lib().or_hash_touched(depth.reshape(-1), W_, H_, CAM6, I3.reshape(-1).copy(), Z3, VS, TS, keys.ctypes.data, n)
    return {tuple(k) for k in keys.tolist()}


def _corner_in_view(key):
    """checkBlockVisibility: any of the block's 8 corners lands on a pixel (RTZ of
    screen + 0.5, no depth test), at the identity pose."""
    f = np.float32(8 * VS)
    p0 = np.array(key, np.float32) * f
    margin = False
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                x, y, z = p0[0] + dx * f, p0[1] + dy * f, p0[2] + dz * f
                if z == 0:
                    continue
                u = np.float32(x * CAM6[0] / z + CAM6[2]) + np.float32(0.5)
                v = np.float32(y * CAM6[1] / z + CAM6[3]) + np.float32(0.5)
                ui, vi = int(np.trunc(u)), int(np.trunc(v))
                for q, lim in ((u, W_), (v, H_)):
                    if min(abs(q - 0), abs(q - 1), abs(q - lim), abs(q + 1)) < 1e-3:
                        margin = True
                if 0 <= ui < W_ and 0 <= vi < H_:
                    return True, margin
    return False, margin


def test_visible_list_recheck_keeps_untouched_blocks_in_view():
    s = O.OracleHashScene(voxel=VS, trunc_scale=TS, truncation=0.06, wmax=255, capacity=1 << 16)
    full = np.full((H_, W_), 1.2, np.float32)
    s.integrate(full.reshape(-1), _Wl, I3, Z3)
    vis1 = {tuple(k) for k in s.visible().tolist()}
    # the second frame sees only the right half: the left half's blocks are not touched
    half = full.copy()
    half[:, : W_ // 2] = np.nan
    touched2 = _touched(half)
    s.integrate(half.reshape(-1), _Wl, I3, Z3)
    vis2 = {tuple(k) for k in s.visible().tolist()}
    kept = vis1 - touched2
    assert len(kept) > 50  # blocks the second frame does not touch
    expect = set(touched2)
    for k in kept:
        inview, margin = _corner_in_view(k)
        if margin:
            continue
        if inview:
            expect.add(k)
        else:
            assert k not in vis2, k
    assert expect <= vis2
    assert len(expect & kept) > 50  # the re-check keeps them (a touched-only list would not)


# ---------------------------------------------------------------- trilinear corner swap

def test_trilinear_weights_x_corners_with_the_y_fraction():
    N, dim, trunc = 16, 1.6, 0.3
    c = dim / N
    gy, gx, gz = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")  # y-major storage
    # a field linear in x only: tsdf = (gx - 8) / 10 (weight 1 everywhere)
    vol = np.zeros((N, N, N, 2), np.int16)
    vol[..., 0] = np.round((gx - 8) / 10.0 * 32766).astype(np.int16)
    vol[..., 1] = 1
    flat = vol.reshape(-1)
    # a point inside cell (g0, g1, g2) = (7, 7, 7) with x fraction 0.2 and y fraction 0.7
    a, b, cz = 0.2, 0.7, 0.4
    p = np.array([(7.5 + a) * c, (7.5 + b) * c, (7.5 + cz) * c], np.float32)
    got = lib().or_trilinear_s16_f(flat, N, dim, trunc, p)
    fx = lambda x: np.round((x - 8) / 10.0 * 32766) / 32766.0
    swapped = (1 - b) * fx(7) + b * fx(8)  # x corners weighted by the y fraction, as written
    textbook = (1 - a) * fx(7) + a * fx(8)
    assert abs(swapped - textbook) > 0.04
    assert abs(got - swapped) < 1e-5, (got, swapped, textbook)
    # and a field linear in y is interpolated with the x fraction
    vol[..., 0] = np.round((gy - 8) / 10.0 * 32766).astype(np.int16)
    got_y = lib().or_trilinear_s16_f(vol.reshape(-1), N, dim, trunc, p)
    assert abs(got_y - ((1 - a) * fx(7) + a * fx(8))) < 1e-5


# ---------------------------------------------------------------- depth-to-colour mapping

def test_depth_to_color_mapping_closed_forms():
    rng = np.random.default_rng(5)
    ih, iw = 6, 8
    image = rng.integers(0, 255, (ih, iw, 4), dtype=np.uint8)
    ys, xs = np.meshgrid(np.arange(ih), np.arange(iw), indexing="ij")
    ident = np.stack([xs.ravel(), ys.ravel()], 1).astype(np.float32)
    out = O.map_depth_to_color(ident, None, image)
    want = image.reshape(-1, 4).copy()
    want[:, 3] = 255
    np.testing.assert_array_equal(out, want)
    # +0.49 stays on the pixel; the reversal table reverses the output
    np.testing.assert_array_equal(O.map_depth_to_color(ident + np.float32(0.49), None, image), want)
    rev = np.arange(ih * iw)[::-1].astype(np.uint32)
    np.testing.assert_array_equal(O.map_depth_to_color(ident, rev, image), want[::-1])
    # truncation toward zero: x in (-1.5, -0.5) still lands on column 0, -1.5 does not
    m = np.array([[-0.7, 0.0], [-1.0, 0.0], [-1.49, 2.0], [-1.5, 0.0], [7.49, 5.0], [7.5, 0.0]], np.float32)
    small = O.map_depth_to_color(m, None, image)
    np.testing.assert_array_equal(small[:3, :3], image[[0, 0, 2], [0, 0, 0], :3])
    assert (small[:3, 3] == 255).all()
    assert (small[3] == 0).all() and (small[5] == 0).all()
    np.testing.assert_array_equal(small[4, :3], image[5, 7, :3])
    # a table entry beyond the frame gives 0
    bad = np.zeros(6, np.uint32)
    bad[2] = 6
    out2 = O.map_depth_to_color(m, bad, image)
    assert (out2[2] == 0).all() and (out2[0] == small[0]).all()
