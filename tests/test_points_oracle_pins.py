"""Pins of the oracle's point-cloud extraction (K/extract.cl:4-78 fetchVolumeKernel,
K/KinfuHashVisualization.cl:5-52 fetchHashScene_kernel) against closed forms.

Dense: the reference's inverse-|tsdf| weighting of the two voxels of a sign change
is the linear zero crossing between them, so a ramp with a known root pins the
position (a swapped weight or a z+1 neighbour moves it by a visible amount); an
exact zero gives the voxel centre; unobserved voxels break the pair; the colour
channels of the two branches are swapped relative to each other.  Hash: every
in-band voxel with weight becomes its centre ((8 * block + local) + 0.5) * voxel
size, computed exactly here for a power-of-two voxel size."""
import numpy as np
import pytest

import oracle as O


def s16_volume(q, w):
    """raw words of a short2 volume from int16 tsdf codes q and weights w (same shape)."""
    out = np.empty(q.size * 2, np.int16)
    out[0::2] = q.reshape(-1)
    out[1::2] = w.reshape(-1)
    return out.view(np.uint32)


def test_half_to_float_matches_numpy_everywhere():
    h = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    ref = h.view(np.float16).astype(np.float32)
    got = np.array([O.lib().or_half_to_float(int(v)) for v in h[::1]], np.float32)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), ref[~nan].view(np.uint32))


@pytest.mark.parametrize("qa,qb", [(-300, 900), (900, -300), (-1, 32000), (5000, -5000)])
def test_dense_crossing_is_the_linear_root(qa, qb):
    Nx, Ny, Nz = 1, 1, 16
    q = np.full(Nz, 20000, np.int16) * np.sign(qb).astype(np.int16)
    q[:8] = 20000 * np.sign(qa)
    q[7], q[8] = qa, qb
    vol = s16_volume(q, np.ones(Nz, np.int16))
    cell = (0.5, 0.25, 0.125)
    pts, cols, n = O.points_dense(vol, None, 0, (Nx, Ny, Nz), cell)
    assert n == 1
    ta, tb = qa / 32766.0, qb / 32766.0
    root = 7 + abs(ta) / (abs(ta) + abs(tb))  # linear interpolation between z = 7 and z = 8
    assert abs(pts[0, 2] - (root + 0.5) * 0.125) < 1e-6
    assert pts[0, 0] == (0 + 0.5 - 0) * 0.5 and pts[0, 1] == (-0 - 0.5 + 0) * 0.25
    assert cols[0].tolist() == [0.0, 0.0, 0.0, 1.0]


def test_dense_exact_zero_centre_and_colour_channels():
    Nx, Ny, Nz = 3, 2, 8
    q = np.full((Ny, Nx, Nz), 1000, np.int16)
    w = np.ones_like(q)
    q[1, 2, 3] = 0          # exact zero: voxel centre, colour bytes in order
    q[0, 1, 5:] = -1000     # crossing at z = 5: colour bytes reversed
    cv = np.zeros((Ny, Nx, Nz, 4), np.uint8)
    cv[1, 2, 3] = (10, 20, 30, 7)
    cv[0, 1, 5] = (40, 50, 60, 9)
    pts, cols, n = O.points_dense(s16_volume(q, w), cv, 0, (Nx, Ny, Nz), (1.0, 1.0, 1.0))
    assert n == 2
    # (y, x, z) order: column (y=0, x=1) first
    hx, hy = float(Nx >> 1), float(Ny >> 1)
    assert pts[0].tolist() == [1 + 0.5 - hx, -0 - 0.5 + hy, pytest.approx(4.5 + 0.5, abs=1e-6)]
    assert np.array_equal(cols[0], np.float32([60, 50, 40, 255]) / np.float32(255) * np.float32([1, 1, 1, 0])
                          + np.float32([0, 0, 0, 1]))
    assert pts[1].tolist() == [2 + 0.5 - hx, -1 - 0.5 + hy, 3 + 0.5]
    assert np.array_equal(cols[1][:3], np.float32([10, 20, 30]) / np.float32(255)) and cols[1][3] == 1


def test_dense_unobserved_and_zero_neighbours_break_pairs():
    Nz = 12
    q = np.array([500, -500, 700, 0, -700, 300, -300, 800, -800, 100, -100, 100], np.int16)
    w = np.ones(Nz, np.int16)
    w[5] = 0    # (5, 6): prev unobserved -> no crossing at 6
    w[8] = -3   # negative weights count as unobserved: no crossing at 8 nor at 9
    pts, _, n = O.points_dense(s16_volume(q, w), None, 0, (1, 1, Nz), (1, 1, 1))
    # points: crossing at z=1 (500 -> -500), z=2 (-500 -> 700), exact zero at 3,
    # none at 4 (prev 0), z=7 (prev 6 = -300 -> 800), z=10 (100 -> -100), z=11
    zs = [1, 2, 3, 7, 10, 11]
    assert n == len(zs)
    roots = [0.5, 1 + 500 / 1200, 3.0, 6.0 + 300 / 1100, 9.5, 10.5]
    for k, (z, r) in enumerate(zip(zs, roots)):
        assert abs(pts[k, 2] - (r + 0.5)) < 1e-5, (k, z)


def test_dense_order_cap_and_storages_agree():
    rng = np.random.default_rng(3)
    Nx, Ny, Nz = 5, 4, 70
    q = rng.integers(-2000, 2000, (Ny, Nx, Nz)).astype(np.int16)
    q[rng.random(q.shape) < 0.05] = 0
    w = (rng.random(q.shape) < 0.85).astype(np.int16)
    cell = (0.01, 0.02, 0.03)
    pts, cols, n = O.points_dense(s16_volume(q, w), None, 0, (Nx, Ny, Nz), cell)
    assert n > 100
    # (y, x) column order, z ascending inside a column
    cx = np.round((pts[:, 0] / cell[0]) - 0.5 + (Nx >> 1)).astype(int)
    cy = np.round(-(pts[:, 1] / cell[1]) - 0.5 + (Ny >> 1)).astype(int)
    key = cy * Nx + cx
    assert np.all(np.diff(key) >= 0)
    same = np.diff(key) == 0
    assert np.all(np.diff(pts[:, 2])[same] > 0)
    # the cap keeps the total and writes the prefix
    p2, _, n2 = O.points_dense(s16_volume(q, w), None, 0, (Nx, Ny, Nz), cell, cap=37)
    assert n2 == n and np.array_equal(p2, pts[:37])
    # FP32 storage holding the decoded values gives the same points bit for bit
    f = np.empty((q.size, 2), np.float32)
    f[:, 0] = q.reshape(-1).astype(np.float32) / np.float32(32766)
    f[:, 1] = w.reshape(-1)
    p3, _, n3 = O.points_dense(f.view(np.uint32).reshape(-1), None, 2, (Nx, Ny, Nz), cell)
    assert n3 == n and np.array_equal(p3.view(np.uint32), pts.view(np.uint32))
    # FP16 storage of small integers: the same points as FP32 storage of the same values
    qi = rng.integers(-8, 9, (Ny, Nx, Nz)).astype(np.int16)
    h = np.empty((qi.size, 2), np.uint16)
    h[:, 0] = qi.reshape(-1).astype(np.float16).view(np.uint16)
    h[:, 1] = w.reshape(-1).astype(np.uint16)
    p4, _, n4 = O.points_dense(h.view(np.uint32).reshape(-1), None, 1, (Nx, Ny, Nz), cell)
    f[:, 0] = qi.reshape(-1)
    p5, _, n5 = O.points_dense(f.view(np.uint32).reshape(-1), None, 2, (Nx, Ny, Nz), cell)
    assert n4 == n5 > 50 and np.array_equal(p4.view(np.uint32), p5.view(np.uint32))


def _brick_words(sdf_codes, weights):
    return (sdf_codes.astype(np.int16).view(np.uint16).astype(np.uint32)
            | (weights.astype(np.uint32) << 16))


def test_hash_points_are_voxel_centres_in_band():
    vs, ts = 0.25, 0.02
    band = np.float32(32767) * np.float32(ts)   # 655.34
    codes = np.zeros((2, 512), np.int32)
    wts = np.zeros((2, 512), np.int32)
    codes[0, 5], wts[0, 5] = 655, 3      # in band
    codes[0, 6], wts[0, 6] = -655, 1     # in band
    codes[0, 7], wts[0, 7] = 656, 3      # outside
    codes[0, 300], wts[0, 300] = 0, 0    # no weight
    codes[1, :], wts[1, :] = 10, 1       # every voxel of brick 1
    words = _brick_words(codes, wts).reshape(-1)
    cols = np.zeros((2 * 512, 4), np.uint8)
    cols[5] = (1, 2, 3, 4)
    pos = np.array([[1, -2, 3], [0, 0, 0], [-4, 5, 6]], np.int32)
    ptr = np.array([0, -1, 1], np.int32)  # the middle block holds no brick
    pts, cc, n = O.points_hash(pos, ptr, words, cols, 0, vs, ts)
    assert 655 < band < 656
    assert n == 2 + 512
    def centre(b, i):
        loc = np.array([i % 8, (i % 64) // 8, i // 64])
        return ((pos[b] * 8 + loc) + 0.5) * vs
    assert np.array_equal(pts[0], centre(0, 5).astype(np.float32))
    assert np.array_equal(pts[1], centre(0, 6).astype(np.float32))
    assert np.array_equal(cc[0], np.float32([1, 2, 3, 0]) / np.float32(255) + np.float32([0, 0, 0, 1]))
    for i in (0, 7, 64, 511):
        assert np.array_equal(pts[2 + i], centre(2, i).astype(np.float32))
    # FP16 storage: band |sdf| < trunc_scale
    h = np.zeros(512, np.float16)
    h[[1, 2, 3]] = [0.0195, -0.0195, 0.0205]
    words16 = h.view(np.uint16).astype(np.uint32) | (np.uint32(2) << 16)
    pts16, _, n16 = O.points_hash(pos[:1], ptr[:1], words16, None, 1, vs, ts)
    assert n16 == 511 and np.array_equal(pts16[2], centre(0, 2).astype(np.float32))
    assert np.array_equal(pts16[3], centre(0, 4).astype(np.float32))  # voxel 3 is outside
    # the cap keeps the total
    _, _, nc = O.points_hash(pos, ptr, words, cols, 0, vs, ts, cap=3)
    assert nc == n
