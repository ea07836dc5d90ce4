"""Pins of the photometric (colour) tracker oracle (CPU), chosen so that a dropped
term, a wrong sign or index, or a transposed operand fails:

  * half_sample_bgra_kernel: a constant image stays constant; a ramp c(x) = x
    halves to c'(x) = 2x inside and to the clipped window's integer mean on the
    border; radius 0 picks the even pixels;
  * bgra_gradient_kernel: a ramp of slope a gives gx = a, gy = 0 (and the
    transposed ramp the reverse); a falling ramp saturates to 0 in uint8 and is
    -a in int8; the border is all zero (alpha 0);
  * color_difference_kernel: a model on its own pixels with every colour off by
    (d0, d1, d2) gives sum = count * (d0^2 + d1^2 + d2^2);
  * color_icp_block_kernel: each parameter's photometric Jacobian equals the
    gradient image times a central finite difference (FP64) of the projection
    under the parameter's motion (translation or rotation about an axis), and the
    29 terms sit in the kernel's slots;
  * the reductions sum every point once (FP64 sum of the per-point terms);
  * the corrected tracker recovers a known offset to within a level-0 pixel and
    halves the pose error between two frames of the textured orbit; the
    reference's own decode (rows of the depth tracker over the kernel's lower-row
    slots) is not positive definite on the same data and fails the track.
"""
import numpy as np
import pytest

import oracle as O
from natureuserinterfacestudio_amd import workloads as W

WL = W.SMALL


class _Wl:
    def __init__(self, w, h, fx=60.0, cx=None, cy=None):
        self.width, self.height = w, h
        self._cam = np.array([fx, fx, (w - 1) / 2 if cx is None else cx, (h - 1) / 2 if cy is None else cy, 0.5, 4.5],
                             np.float32)

    def cam6(self):
        return self._cam


def _img(w, h, f):
    ys, xs = np.mgrid[0:h, 0:w]
    out = np.zeros((h, w, 4), np.uint8)
    for c in range(3):
        out[..., c] = f(xs, ys, c)
    out[..., 3] = 255
    return out


def test_half_sample_closed_forms():
    const = _img(16, 12, lambda x, y, c: np.full_like(x, 40 + 50 * c))
    h = O.color_half_sample(const, 8, 6)
    assert (h[..., :3] == const[:6, :8, :3]).all() and (h[..., 3] == 255).all()
    ramp = _img(16, 12, lambda x, y, c: x + c)
    h = O.color_half_sample(ramp, 8, 6)
    xs = np.arange(8)
    for c in range(3):
        inner = h[:, 1:, c]
        np.testing.assert_array_equal(inner, np.broadcast_to(2 * xs[1:] + c, inner.shape))
        np.testing.assert_array_equal(h[:, 0, c], np.full(6, (0 + 1 + 2 * c) // 2))  # columns 0, 1 only
    h0 = O.color_half_sample(ramp, 8, 6, radius=0)
    np.testing.assert_array_equal(h0[..., :3], ramp[::2, ::2, :3])


@pytest.mark.parametrize("signed", [False, True])
def test_gradient_closed_forms(signed):
    a = 3
    up = _img(10, 8, lambda x, y, c: 20 + a * x)
    gx, gy = O.color_gradient(up, 10, 8, signed)
    inner = (slice(1, 7), slice(1, 9))
    assert (gx[inner][..., :3] == a).all() and (gy[inner][..., :3] == 0).all()
    assert (gx[inner][..., 3] == 255).all()
    border = np.ones((8, 10), bool)
    border[inner] = False
    assert (gx[border] == 0).all() and (gy[border] == 0).all()
    down = _img(10, 8, lambda x, y, c: 200 - a * y)
    gx, gy = O.color_gradient(down, 10, 8, signed)
    expect = (256 - a) if signed else 0  # -a as int8 stored in a byte, or saturated
    assert (gy[inner][..., :3] == expect).all() and (gx[inner][..., :3] == 0).all()


def _model_on_pixels(wl, depth, colors):
    """World points (camera = world) on the centres of every pixel."""
    cam = wl.cam6()
    ys, xs = np.mgrid[0:wl.height, 0:wl.width]
    z = np.full(xs.shape, depth, np.float64)
    v = np.stack([(xs - cam[2]) / cam[0] * z, (ys - cam[3]) / cam[1] * z, z], -1).reshape(-1, 3)
    return v.astype(np.float32), colors.reshape(-1, 4)


def test_color_difference_closed_form():
    wl = _Wl(24, 16)
    frame = _img(24, 16, lambda x, y, c: 50 + 3 * x + 5 * y + c)
    known = frame.copy()
    known[..., :3] -= np.array([2, 3, 4], np.uint8)
    known[0, :, 3] = 0  # invalid model colours drop
    v, k = _model_on_pixels(wl, 1.5, known)
    v[30] = np.nan  # a NaN model point in a valid row drops too
    dist, count = O.color_difference(wl, 0, v, k, frame, np.eye(3), np.zeros(3))
    n = 24 * 16 - 24 - 1
    assert count == n and dist == n * (4 + 9 + 16)


def _proj(wl, q):
    cam = wl.cam6().astype(np.float64)
    return np.array([q[0] * cam[0] / q[2] + cam[2], q[1] * cam[1] / q[2] + cam[3]])


@pytest.mark.parametrize("signed", [False, True])
def test_color_terms_are_the_photometric_jacobian(signed):
    rng = np.random.default_rng(7)
    wl = _Wl(32, 24)
    frame = rng.integers(0, 256, (24, 32, 4), dtype=np.uint8)
    frame[..., 3] = 255
    gx, gy = O.color_gradient(frame, 32, 24, signed)
    R = np.eye(3, dtype=np.float32)
    t = np.zeros(3, np.float32)
    for trial in range(20):
        q = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.2, 0.2), rng.uniform(0.8, 2.0)], np.float32)
        px = np.round(_proj(wl, q)).astype(int)
        if not (1 <= px[0] < 31 and 1 <= px[1] < 23):
            continue
        known = np.array([rng.integers(0, 256), rng.integers(0, 256), rng.integers(0, 256), 255], np.uint8)
        out = O.color_terms(wl, 0, q, known, frame, gx, gy, R, t, signed=signed)
        li = px[1] * 32 + px[0]
        g = (lambda b: b.view(np.int8).astype(np.float64)) if signed else (lambda b: b.astype(np.float64))
        GX, GY = g(gx.reshape(-1, 4)[li, :3]), g(gy.reshape(-1, 4)[li, :3])
        diff = frame.reshape(-1, 4)[li, :3].astype(np.float64) - known[:3]
        qd = q.astype(np.float64)
        motions = [np.eye(3)[0], np.eye(3)[1], np.eye(3)[2]]
        J = []
        for p in range(6):
            h = 1e-6
            move = (lambda e: motions[p] * e) if p < 3 else (lambda e, a=motions[p - 3]: e * np.cross(a, qd))
            du = (_proj(wl, qd + move(h)) - _proj(wl, qd + move(-h))) / (2 * h)
            J.append(du[0] * GX + du[1] * GY)
        for p in range(6):
            for c in range(p + 1):
                e = J[p] @ J[c]
                assert abs(out[p * (p + 3) // 2 + c] - e) <= 1e-4 * max(1.0, abs(e)), (trial, p, c)
            e = J[p] @ diff
            assert abs(out[p * (p + 3) // 2 + p + 1] - e) <= 1e-4 * max(1.0, abs(e)), (trial, p)
        assert abs(out[27] - np.linalg.norm(diff)) <= 1e-5 * max(1.0, np.linalg.norm(diff))
        assert out[28] == 1.0
        # rotation only: the three rotation parameters in the first slots
        ro = O.color_terms(wl, 0, q, known, frame, gx, gy, R, t, rot_only=True, signed=signed)
        for p in range(3):
            for c in range(p + 1):
                e = J[3 + p] @ J[3 + c]
                assert abs(ro[p * (p + 3) // 2 + c] - e) <= 1e-4 * max(1.0, abs(e))
        # |diff| and the count follow the nine terms (the reference host reads slots 27, 28)
        assert ro[9] == out[27] and ro[10] == 1.0 and (ro[11:] == 0).all()


def _frame_model(f):
    d, bgra, gt = WL.render(f, color=True)
    dep = np.empty(d.size, np.float32)
    O.lib().or_passing_filter_f(d.reshape(-1), dep, d.size, 0.5, 4.5)
    cam = WL.cam6()
    ys, xs = np.divmod(np.arange(d.size), WL.width)
    v = np.stack([dep * ((xs - cam[2]) / cam[0]), dep * ((ys - cam[3]) / cam[1]), dep], -1)
    R, t = gt[:, :3], gt[:, 3]
    vw = (v @ R.T + t).astype(np.float32)
    cols = bgra.reshape(-1, 4).copy()
    cols[:, 3] = np.where(np.isnan(dep), 0, 255)
    return vw, cols, bgra, R, t


def test_reductions_sum_every_point_once():
    vw, cols, bgra, R, t = _frame_model(3)
    _, _, frame, _, _ = _frame_model(4)
    gx, gy = O.color_gradient(frame, WL.width, WL.height, True)
    Rc = R.T.astype(np.float32)
    tc = (-Rc @ t).astype(np.float32)
    blocks = O.color_reduce(WL, 0, vw, cols, frame, gx, gy, Rc, tc, signed=True)
    total = blocks.astype(np.float64).sum(0)
    idx = np.random.default_rng(1).choice(len(vw), 3000, replace=False)
    per = np.array([O.color_terms(WL, 0, vw[i], cols[i], frame, gx, gy, Rc, tc, signed=True) for i in idx])
    assert total[28] > 20000  # valid points
    # the count is exact; the sampled mean of every term matches the full mean
    full = np.array([O.color_terms(WL, 0, vw[i], cols[i], frame, gx, gy, Rc, tc, signed=True)
                     for i in range(len(vw))], np.float64).sum(0)
    assert total[28] == full[28]
    np.testing.assert_allclose(total, full, rtol=2e-4, atol=1e-2 * np.abs(full).max() * 1e-4)
    assert per.shape == (3000, 29)


def _pose_err(R, t, R1, t1):
    ang = np.degrees(np.arccos(np.clip((np.trace(R.T @ R1) - 1) / 2, -1, 1)))
    return np.linalg.norm(t - t1) * 1000, ang


def test_corrected_tracker_recovers_an_offset_within_a_pixel():
    vw, cols, bgra, R0, t0 = _frame_model(4)
    ok, R, t, its, _, _ = O.color_icp(WL, vw, cols, bgra, R0, t0 + np.float32([0.01, 0, 0]),
                                      iterations=[10, 10, 10], corrected=True)
    terr, rerr = _pose_err(R, t, R0, t0)
    # one level-0 pixel at 1.2 m with fx = 262.5 is 4.6 mm
    assert ok == 1 and terr < 4.6 and rerr < 0.2, (terr, rerr)


def test_corrected_tracker_halves_the_error_between_frames_and_reference_decode_fails():
    vw, cols, _, R0, t0 = _frame_model(4)
    _, _, frame, R1, t1 = _frame_model(5)
    t_start, r_start = _pose_err(R0, t0, R1, t1)
    ok, R, t, _, _, cnt = O.color_icp(WL, vw, cols, frame, R0, t0, iterations=[10, 5, 4], corrected=True)
    terr, rerr = _pose_err(R, t, R1, t1)
    assert ok == 1 and cnt > 20000
    assert terr < 0.75 * t_start and rerr < 0.6 * r_start, (terr, t_start, rerr, r_start)
    ok_ref, Rr, tr, its, _, _ = O.color_icp(WL, vw, cols, frame, R0, t0, iterations=[10, 5, 4], corrected=False)
    assert ok_ref == 0 and its == 1
    np.testing.assert_array_equal(Rr, R0.astype(np.float32))


# ---- joint depth + intensity ICP terms (K/icp.cl:328-480) ---------------------------

def test_intensity_and_derivative_closed_forms():
    rng = np.random.default_rng(2)
    px = rng.integers(0, 256, (200, 4), dtype=np.uint8)
    px[:20, 3] = 0
    I = O.bgra_to_intensity(px)
    assert np.isnan(I[:20]).all()
    ref = 0.299 * px[20:, 2] / 255.0 + 0.587 * px[20:, 1] / 255.0 + 0.114 * px[20:, 0] / 255.0
    assert np.abs(I[20:] - ref).max() < 3e-7
    assert abs(O.bgra_to_intensity(np.array([[255, 255, 255, 255]], np.uint8))[0] - 1.0) < 1.2e-7
    # a plane of intensity a x + b y (dyadic values: exact) has derivative (a, b) inside
    w, h = 12, 9
    ys, xs = np.mgrid[0:h, 0:w]
    a, b = 0.0625, -0.125
    plane = (0.25 + a * xs + b * ys).astype(np.float32)
    d = O.intensity_derivatives(plane, w, h)
    assert (d[1:-1, 1:-1, 0] == a).all() and (d[1:-1, 1:-1, 1] == b).all()
    border = np.ones((h, w), bool)
    border[1:-1, 1:-1] = False
    assert np.isnan(d[border]).all()
    holed = plane.copy()
    holed[4, 5] = np.nan
    dh = O.intensity_derivatives(holed, w, h)
    near = np.zeros((h, w), bool)
    near[3:6, 4:7] = True
    assert np.isnan(dh[near]).all() and (dh[~near & ~border] == d[~near & ~border]).all()


def test_euler_derivatives_at_zero_are_the_rotation_generators():
    rng = np.random.default_rng(3)
    for _ in range(10):
        p = rng.normal(size=3).astype(np.float32)
        phis = O.eval_phis(np.zeros(3, np.float32), p)
        x, y, z = p
        np.testing.assert_array_equal(phis[0], np.array([-y, x, 0], np.float32))   # about z
        np.testing.assert_array_equal(phis[1], np.array([z, 0, -x], np.float32))   # about y
        np.testing.assert_array_equal(phis[2], np.array([0, -z, y], np.float32))   # about x


def _camera_maps(f):
    d, bgra, gt = WL.render(f, color=True)
    dep = np.empty(d.size, np.float32)
    O.lib().or_passing_filter_f(d.reshape(-1), dep, d.size, 0.5, 4.5)
    cam = WL.cam6()
    ys, xs = np.divmod(np.arange(d.size), WL.width)
    v = np.stack([dep * ((xs - cam[2]) / cam[0]), dep * ((ys - cam[3]) / cam[1]), dep], -1).astype(np.float32)
    n = np.empty_like(v)
    O.lib().or_normals_simple_f(v.reshape(-1), n.reshape(-1), WL.width, WL.height, 0.3)
    return v, n, O.bgra_to_intensity(bgra), gt[:, :3].astype(np.float32), gt[:, 3].astype(np.float32)


def test_intensity_icp_identical_frames_closed_form():
    """Current = previous maps at the same pose, no intensity term: every valid point
    matches itself (distance 0), so its terms are the products of (n x p, n, 0) (the
    generators at zero Euler angles against the negated normal) and the count slot,
    which needs a distance above 1 % of the threshold, stays 0."""
    v, n, I, R, t = _camera_maps(3)
    Rw, tw = np.eye(3, dtype=np.float32), np.zeros(3, np.float32)
    d = O.intensity_derivatives(I, WL.width, WL.height)
    thr = np.array([0.1, 0.34, 0.0, 0.0], np.float32)  # colour threshold 0: no intensity term
    out = O.intensity_icp_sums(WL, 0, v, n, I, v, n, I, d, Rw, tw, Rw, tw, np.zeros(3, np.float32), thr)
    assert out[27] == 0 and out[28] == 0
    ok = ~np.isnan(v).any(1) & ~np.isnan(n).any(1)
    p, nn = v[ok].astype(np.float64), n[ok].astype(np.float64)
    # a normal's cross product with itself is zero: every valid point passes the angle test;
    # c[:, 0:3] = -n . (generator) for the generators about z, y, x, then n, then n . 0
    c = np.concatenate([np.zeros((len(p), 3)), nn, np.zeros((len(p), 1))], 1)
    c[:, 0] = -(nn[:, 0] * -p[:, 1] + nn[:, 1] * p[:, 0])
    c[:, 1] = -(nn[:, 0] * p[:, 2] - nn[:, 2] * p[:, 0])
    c[:, 2] = -(-nn[:, 1] * p[:, 2] + nn[:, 2] * p[:, 1])
    expect = []
    for i in range(6):
        for j in range(i, 7):
            expect.append((c[:, i] * c[:, j]).sum())
    expect = np.array(expect)
    scale = np.abs(expect).max()
    assert np.abs(out[:27] - expect).max() <= 1e-4 * scale
