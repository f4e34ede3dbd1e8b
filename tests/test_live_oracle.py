"""The oracle's relative_change (live.rs:208-218), the live view's settle
measure, pinned by closed forms and a vectorized restatement."""
import numpy as np

import oracle as O


def _img(seed, h=37, w=29):
    return np.random.default_rng(seed).uniform(0.0, 3.0, (h, w, 3)).astype(np.float32)


def test_closed_forms():
    a = _img(1)
    assert O.relative_change(a, a) == 0.0
    # a = 2b: |la - lb| = |lb| on every sampled pixel
    assert abs(O.relative_change(2.0 * a, a) - 1.0) < 1e-12
    assert O.relative_change(a, np.zeros_like(a)) == 0.0  # base below 1e-9
    # only pixels off the every-4th sample differ
    b = a.copy().reshape(-1, 3)
    idx = np.arange(len(b))
    b[idx % 4 != 0] += 1.0
    assert O.relative_change(b.reshape(a.shape), a) == 0.0
    # one sampled pixel moved by d in luma: d / sum of |luma|
    c = a.copy().reshape(-1, 3)
    c[8] += np.float32(0.5)
    lum = a.reshape(-1, 3)[::4].astype(np.float64) @ np.array([0.2126, 0.7152, 0.0722])
    assert abs(O.relative_change(c.reshape(a.shape), a) - 0.5 / np.abs(lum).sum()) < 1e-12


def test_matches_vectorized():
    w = np.array([0.2126, 0.7152, 0.0722])
    for seed in range(4):
        a, b = _img(seed), _img(seed + 10) - 1.0  # negative luma in b: |lb|
        la = a.reshape(-1, 3)[::4].astype(np.float64) @ w
        lb = b.reshape(-1, 3)[::4].astype(np.float64) @ w
        want = np.abs(la - lb).sum() / np.abs(lb).sum()
        assert abs(O.relative_change(a, b) - want) <= 1e-12 * want
