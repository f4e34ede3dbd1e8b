"""The panel test's in-rectangle decision fabsf(x / y) <= 1 (light.rs:60-62)
is taken in the kernels without a division unless |x| lies within 2^-20 of y.
This checks the two comparisons behind that shortcut against the rounded
division itself, in IEEE single precision, on values packed around the
boundary where a wrong margin or comparison would show."""
import numpy as np

LO = np.float32(1.0 - 2.0 ** -20)
HI = np.float32(1.0 + 2.0 ** -20)


def _decide(x, y):
    """(decided, inside): the shortcut's answer where it gives one."""
    ax = np.abs(x)
    ok = (y >= np.float32(2.0 ** -100)) & (y <= np.float32(2.0 ** 100))
    inside = ok & (ax <= y * LO)
    outside = ok & (ax > y * HI)
    return inside | outside, inside


def test_shortcut_agrees_with_rounded_division():
    rng = np.random.default_rng(5)
    n = 1 << 21
    y = (10.0 ** rng.uniform(-30, 30, n)).astype(np.float32)
    # ratios packed within a few 2^-20 of 1, both signs, plus exact neighbours
    r = 1.0 + rng.uniform(-8, 8, n) * 2.0 ** -20
    x = (y.astype(np.float64) * r * rng.choice([-1.0, 1.0], n)).astype(np.float32)
    step = rng.integers(-40, 41, n).astype(np.int32)
    x = (x.view(np.int32) + step).view(np.float32)
    decided, inside = _decide(x, y)
    truth = np.abs(x / y) <= np.float32(1.0)
    assert decided.mean() > 0.5
    assert np.array_equal(inside[decided], truth[decided])


def test_shortcut_leaves_nan_and_extremes_to_the_division():
    x = np.array([np.nan, 1.0, np.inf, 1e-40, 3e38, 1.0], np.float32)
    y = np.array([1.0, np.inf, np.inf, 1e-41, 3e38, 3e20], np.float32)
    decided, inside = _decide(x, y)
    assert not decided[:5].any()
    assert decided[5] and inside[5]
