"""The joint ICP order's parts in the oracle, pinned by what the mathematics fixes (the
order has no counterpart in the reference: its photometric terms, K/icp.cl:352-480,
have no host loop).

  * intensity half-sample: constant images stay constant, a block's mean lies within
    its values, NaN exactly where all four are NaN, and it equals numpy's nanmean to
    the rounding of four float additions;
  * photometric row: for a linear and a quadratic intensity image and small pose
    increments x = (w, t), the row's c[0:6] . x / sqrt(weight) is the change of the
    model's intensity at the projection of the moved point p + t - w x p (the same
    motion under which the geometric row (n x p, n) gives n . dp), and c[6] /
    sqrt(weight) is I - I_ref; rows are refused for NaN, a large difference or a small
    gradient;
  * engines: the joint order with weight 0 runs the compensated order's poses bit for
    bit (float and FP64 instances); with weight 0.1 it differs from it and follows the
    same motion within 2 mm and 0.2 degrees at 640x480 / 512^3.
"""
import numpy as np
import pytest

import oracle as O
from natureuserinterfacestudio_amd import workloads as W


def test_intensity_half_sample_properties():
    rng = np.random.default_rng(3)
    wd, hd = 37, 23
    src = rng.uniform(0.0, 1.0, (2 * hd, 2 * wd)).astype(np.float32)
    src[rng.uniform(size=src.shape) < 0.3] = np.nan
    src[0:2, 0:2] = np.nan  # a block without a valid value
    out = O.intensity_half_sample(src, wd, hd).reshape(hd, wd)
    blocks = src.reshape(hd, 2, wd, 2).transpose(0, 2, 1, 3).reshape(hd, wd, 4)
    allnan = np.isnan(blocks).all(axis=2)
    assert np.array_equal(np.isnan(out), allnan) and allnan[0, 0]
    with np.errstate(invalid="ignore"), pytest.warns(RuntimeWarning):
        mean = np.nanmean(blocks.astype(np.float64), axis=2)
    ok = ~allnan
    np.testing.assert_allclose(out[ok], mean[ok], rtol=4e-7, atol=0)
    lo = np.nanmin(blocks[ok], axis=1)
    hi = np.nanmax(blocks[ok], axis=1)
    assert (out[ok] >= lo).all() and (out[ok] <= hi).all()
    const = np.full((2 * hd, 2 * wd), 0.375, np.float32)
    assert (O.intensity_half_sample(const, wd, hd) == np.float32(0.375)).all()


def _rot(axis, ang):
    a = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


def _sobel8(I):
    g = np.full(I.shape + (2,), np.nan)
    g[1:-1, 1:-1, 0] = ((I[:-2, 2:] - I[:-2, :-2]) + 2 * (I[1:-1, 2:] - I[1:-1, :-2]) + (I[2:, 2:] - I[2:, :-2])) / 8
    g[1:-1, 1:-1, 1] = ((I[2:, :-2] - I[:-2, :-2]) + 2 * (I[2:, 1:-1] - I[:-2, 1:-1]) + (I[2:, 2:] - I[:-2, 2:])) / 8
    return g


@pytest.mark.parametrize("kind", ["linear", "quadratic"])
def test_photometric_row_is_the_intensity_derivative(kind):
    w, h = 160, 120
    fx = fy = 131.25
    cx, cy = 79.5, 59.5
    cam6 = np.array([fx, fy, cx, cy, 0.5, 4.5], np.float32)
    v, u = np.mgrid[0:h, 0:w].astype(np.float64)
    if kind == "linear":
        f = lambda uu, vv: 0.2 + 0.003 * uu - 0.002 * vv  # noqa: E731
    else:
        f = lambda uu, vv: 0.3 + 0.006 * uu + 0.001 * vv + 1e-5 * (uu - 80) ** 2 - 2e-5 * (uu - 80) * (vv - 60)  # noqa: E731
    I = f(u, v)
    G = _sobel8(I)
    # previous camera: rotated and moved; pp = Ri (p - tp)
    Rprev = _rot([0.3, 1.0, -0.2], 0.25)
    tp = np.array([0.1, -0.05, 0.2])
    Ri = Rprev.T
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(200):
        # a point in front of the previous camera, away from the border
        uv = rng.uniform([20, 20], [w - 20, h - 20])
        z = rng.uniform(1.0, 3.0)
        pc = np.array([(uv[0] - cx) / fx * z, (uv[1] - cy) / fy * z, z])
        p = Rprev @ pc + tp
        pp = Ri @ (p - tp)
        px, py = pp[0] * fx / pp[2] + cx, pp[1] * fy / pp[2] + cy
        ic = f(px, py) + 0.01
        ok, c = O.photometric_row_d(ic, p, pp, px, py, Ri, cam6, w, h, I, G, 0.1, 0.2, 0.002)
        assert ok
        psw = np.sqrt(0.1)

        def model(q):  # the model's intensity at the projection of world point q
            qq = Ri @ (q - tp)
            return f(qq[0] * fx / qq[2] + cx, qq[1] * fy / qq[2] + cy)

        assert abs(c[6] / psw - (ic - model(p))) < 2e-3  # bilinear vs exact image
        for _ in range(3):
            wv = rng.normal(size=3) * 1e-4
            tv = rng.normal(size=3) * 1e-4
            dp = tv - np.cross(wv, p)
            fd = model(p + dp) - model(p - dp)
            lin = 2 * (c[:3] @ wv + c[3:6] @ tv) / psw
            tol = 1e-9 if kind == "linear" else 5e-7 + 0.02 * abs(fd)
            assert abs(lin - fd) <= tol + 1e-3 * abs(fd), (kind, lin, fd)
        checked += 1
    assert checked == 200
    # refusals: NaN intensity, a large difference, a small gradient
    uv = np.array([80.3, 60.7])
    pc = np.array([(uv[0] - cx) / fx * 2.0, (uv[1] - cy) / fy * 2.0, 2.0])
    p = Rprev @ pc + tp
    pp = Ri @ (p - tp)
    px, py = pp[0] * fx / pp[2] + cx, pp[1] * fy / pp[2] + cy
    base = f(px, py)
    assert not O.photometric_row_d(np.nan, p, pp, px, py, Ri, cam6, w, h, I, G, 0.1, 0.2, 0.002)[0]
    assert not O.photometric_row_d(base + 0.25, p, pp, px, py, Ri, cam6, w, h, I, G, 0.1, 0.2, 0.002)[0]
    assert O.photometric_row_d(base + 0.15, p, pp, px, py, Ri, cam6, w, h, I, G, 0.1, 0.2, 0.002)[0]
    assert not O.photometric_row_d(base, p, pp, px, py, Ri, cam6, w, h, I, G, 0.1, 0.2, 0.5)[0]
    flat = np.full_like(I, 0.5)
    assert not O.photometric_row_d(0.5, p, pp, px, py, Ri, cam6, w, h, flat, _sobel8(flat), 0.1, 0.2, 0.002)[0]


def _poses(eng, frames, wl):
    out = []
    for f in range(frames):
        d, bgra, _ = wl.render(f, color=True)
        ok, pose, _ = eng.run(d, bgra)
        assert ok == 1, f
        out.append(pose.copy())
    return np.array(out)


def test_joint_engine_weight_zero_is_the_compensated_order():
    wl = W.SMALL
    frames = 5
    comp = O.OracleEngine(wl, prec=0, color=True, compensated=True, storage=2)
    joint = O.OracleEngine(wl, prec=0, color=True, joint=True, storage=2)
    joint.set_photometric(0.0)
    np.testing.assert_array_equal(_poses(joint, frames, wl), _poses(comp, frames, wl))


def test_joint_engine_tracks_the_orbit():
    # at 640x480 / 512^3 (at 320x240 / 128^3 with 1.6 cm voxels even the float and FP64
    # instances of one order differ by millimetres, so the orders are compared here)
    wl = W.FULL
    frames = 5
    comp = O.OracleEngine(wl, prec=0, color=True, compensated=True, storage=1)
    joint = O.OracleEngine(wl, prec=0, color=True, joint=True, storage=1)
    pj, pc = _poses(joint, frames, wl), _poses(comp, frames, wl)
    assert not np.array_equal(pj, pc)  # the photometric rows take part
    # the same motion as the geometric order's (the orbit moves 5 cm and 2 deg a frame)
    for f in range(frames):
        assert np.linalg.norm(pj[f, 9:] - pc[f, 9:]) < 2e-3, f
        Rr = pj[f, :9].reshape(3, 3).T @ pc[f, :9].reshape(3, 3)
        w = 0.5 * np.array([Rr[2, 1] - Rr[1, 2], Rr[0, 2] - Rr[2, 0], Rr[1, 0] - Rr[0, 1]])
        assert np.degrees(np.arctan2(np.linalg.norm(w), (np.trace(Rr) - 1) / 2)) < 0.2, f
    assert len(joint.photometric(0)) == wl.width * wl.height
    assert len(joint.photometric(1)) == (wl.width // 2) * (wl.height // 2)
    assert len(joint.photometric(4)) == 2 * wl.width * wl.height
    joint.close()
    comp.close()


def test_fp64_joint_engine_weight_zero_is_the_fp64_compensated_order():
    """The FP64 instance of the joint order (intensities, half-samples and derivatives
    in double) reduces to the FP64 compensated order when the photometric weight is 0."""
    wl = W.SMALL
    frames = 4
    comp = O.OracleEngine(wl, prec=1, color=True, compensated=True)
    joint = O.OracleEngine(wl, prec=1, color=True, joint=True)
    joint.set_photometric(0.0)
    np.testing.assert_array_equal(_poses(joint, frames, wl), _poses(comp, frames, wl))
    joint.close()
    comp.close()
