"""Pins of the oracle's host swapping of hash bricks (OCL/NuiKinfuOpenCLHashGlobalCache.cpp,
K/KinfuHashSwapping.cl) on the CPU.

Blocks that leave the view go to the store with exactly the content they had; while the
camera does not move nothing leaves; a block that comes back is re-integrated into a fresh
brick and then weight-averaged with its stored copy, which agrees with never having
swapped it out up to the int16 quantisation of the fresh brick (weights exactly)."""
import numpy as np

import oracle as O
from natureuserinterfacestudio_amd import workloads as W

WL = W.HASH_SMALL


def _frame(f):
    d, bgra, gt = WL.render(f, color=True)
    dep = np.empty(d.size, np.float32)
    O.lib().or_passing_filter_f(d.reshape(-1), dep, d.size, 0.5, 4.5)
    return dep, bgra, gt[:, :3].astype(np.float32).copy(), gt[:, 3].astype(np.float32).copy()


def _away(R):
    a = np.deg2rad(150.0)
    Ry = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]], np.float32)
    return (R @ Ry).astype(np.float32)


def _by_key(keys, *arrs):
    return {tuple(k): tuple(a[i] for a in arrs) for i, k in enumerate(keys.tolist())}


def test_static_camera_swaps_nothing_out():
    s = O.OracleHashScene(capacity=1 << 17, swapping=True)
    dep, bgra, R, t = _frame(0)
    for _ in range(3):
        s.integrate(dep, WL, R, t)
    keys, st = s.swap_states()
    assert len(keys) > 100 and np.all(st == 2)
    assert len(s.store()[0]) == 0


def test_blocks_leaving_the_view_are_stored_as_they_were_and_merge_back():
    sw = O.OracleHashScene(capacity=1 << 17, swapping=True)
    plain = O.OracleHashScene(capacity=1 << 17)
    dep, bgra, R, t = _frame(2)
    for _ in range(2):
        sw.integrate(dep, WL, R, t, bgra=bgra)
        plain.integrate(dep, WL, R, t, bgra=bgra)
    before = _by_key(*sw.blocks())
    # look away: the blocks of the first view are not visible any more
    sw.integrate(dep, WL, _away(R), t, bgra=bgra)
    skeys, ssdf, sw_w, srgb = sw.store()
    assert len(skeys) > 100
    for i, k in enumerate(skeys.tolist()):
        b = before[tuple(k)]
        assert np.array_equal(ssdf[i], b[0]) and np.array_equal(sw_w[i], b[1]) and np.array_equal(srgb[i], b[2])
    # only the blocks of the current view hold bricks
    held = sw.blocks(voxels=False)
    assert not (set(map(tuple, held.tolist())) & set(map(tuple, skeys.tolist())))
    # come back: re-integrated, then merged with the stored copy
    sw.integrate(dep, WL, R, t, bgra=bgra)
    plain.integrate(dep, WL, R, t, bgra=bgra)
    got = _by_key(*sw.blocks())
    ref = _by_key(*plain.blocks())
    back = [k for k in map(tuple, skeys.tolist()) if k in got]
    assert len(back) > 100
    worst = 0
    for k in back:
        g, r = got[k], ref[k]
        assert np.array_equal(g[1], r[1])                      # weights exactly
        worst = max(worst, int(np.abs(g[0].astype(int) - r[0].astype(int)).max()))
    assert worst <= 2, worst                                    # sdf codes: quantisation only
    # the store keeps the last copy; the blocks came back, so they are synchronised again
    keys, st = sw.swap_states()
    stm = dict(zip(map(tuple, keys.tolist()), st.tolist()))
    assert all(stm[k] == 2 for k in back)
