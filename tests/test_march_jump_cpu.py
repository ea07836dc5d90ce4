"""The raycast's jump through a clear superblock span (csrc/tsdf.hip march_jump): the
march advances t <- fl(t + step) one step at a time (K/raycast.cl:193-276), and the jump
claims to land, in one operation, exactly where the last of those steps inside the span
lands.  Checked in float32 against the iterated additions for random times, steps and
spans, including spans that reach a power of two, steps that are half an ulp off a
multiple of the ulp (the tie case the jump refuses) and times near zero."""
import numpy as np

f32 = np.float32


def march_jump(t, step, s_out, max_t):
    """The device function, step for step, in float32 (the reciprocal is exact here,
    which only changes the estimate that the result check makes safe)."""
    t, step, s_out, max_t = f32(t), f32(step), f32(s_out), f32(max_t)
    ex = np.array(t, f32).view(np.uint32) & np.uint32(0x7F800000)
    if not (t > 0) or ex < (24 << 23) or ex >= (254 << 23):
        return t
    d = f32(f32(t + step) - t)
    u = np.array(ex - (23 << 23), np.uint32).view(f32)
    top = f32(np.array(ex + (1 << 23), np.uint32).view(f32) - f32(2) * u)
    if abs(f32(step - d)) == f32(0.5) * u:
        return t
    lim = min(s_out, top, max_t)
    m = f32(np.floor(f32(f32(lim - t) * f32(1 / d)))) - f32(1)
    if not (m >= 1):
        return t
    tm = f32(t + f32(m * d))
    return tm if (tm <= s_out and tm <= top and tm < max_t) else t


def iterate(t, step, s_out, max_t):
    """The march's own loop: steps while the next time stays in the span and the loop
    guard holds."""
    t, step = f32(t), f32(step)
    while t < max_t:
        u = f32(t + step)
        if not (u <= s_out):
            break
        t = u
    return t


def test_jump_lands_on_the_iterated_time():
    rng = np.random.default_rng(7)
    jumps = 0
    for _ in range(4000):
        t = f32(rng.uniform(0.05, 6.0))
        step = f32(rng.uniform(0.002, 0.05))
        s_out = f32(t + rng.uniform(0.0, 0.5))
        max_t = f32(rng.choice([27.0, t + rng.uniform(0.0, 0.6)]))
        j = march_jump(t, step, s_out, max_t)
        ref = iterate(t, step, s_out, max_t)
        # the jump may stop short (then the loop goes on from there) but never past, and
        # where it lands is a time of the iterated sequence
        seq = [t]
        while seq[-1] < ref:
            seq.append(f32(seq[-1] + step))
        assert j in seq, (t, step, s_out, max_t, j)
        assert j <= ref
        jumps += j != t
        # from the landing point the loop's remaining steps reach the same end
        assert iterate(j, step, s_out, max_t) == ref
    assert jumps > 2000


def test_jump_stops_before_a_power_of_two():
    # the increment changes in the next binade: the jump ends below 2 - 2 ulp
    t, step = f32(1.9), f32(0.0123)
    j = march_jump(t, step, f32(2.5), f32(27.0))
    assert j <= f32(2.0) - 2 * np.spacing(f32(1.9)) and j == iterate(t, step, j, f32(27.0))


def test_tie_steps_do_not_jump():
    # step = k ulp + ulp/2 at t in [1, 2): fl(t + step) rounds to even, alternating
    u = np.spacing(f32(1.5))
    step = f32(1000 * u + u / 2)
    assert march_jump(f32(1.5), step, f32(1.9), f32(27.0)) == f32(1.5)


def test_no_jump_from_zero_or_tiny_times():
    assert march_jump(f32(0.0), f32(0.02), f32(1.0), f32(27.0)) == f32(0.0)
    assert march_jump(f32(1e-33), f32(0.02), f32(1.0), f32(27.0)) == f32(1e-33)
