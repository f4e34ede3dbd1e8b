"""The oracle's preview development (argent-post tonemap.rs, denoise.rs),
pinned by the reference's own tests restated as properties: AgX black and mid
grey, highlight roll-off without hue flips, the other operators' sanity and
known values (ACES 0.18 -> ~0.106, Khronos Neutral (0.5, 0.3, 0.2) ->
(0.46, ., 0.16)), saturation and exposure, the table fast path within 2e-4 of
the exact operators (tonemap.rs:509-531), the dither keeping black and white
clean with the mean preserved (tonemap.rs:533-548), à-trous cutting the MSE
of a noisy gradient below 0.15x and not bleeding across a material edge
(denoise.rs:770-813)."""
import numpy as np
import pytest

import oracle as O

OPS = ["agx", "agx_punchy", "agx_golden", "aces", "khronos", "reinhard", "clamp"]


def test_agx_black_and_mid_grey():
    black = O.tonemap_pixel([[0, 0, 0]], "agx")[0]
    assert np.all(black < 1e-4)
    grey = O.tonemap_pixel([[0.18] * 3], "agx")[0]
    assert 0.45 < O.srgb_encode(grey[1]) < 0.55
    assert abs(grey[0] - grey[2]) < 2e-3
    punchy = O.tonemap_pixel([[0.18] * 3], "agx_punchy")[0]
    golden = O.tonemap_pixel([[0.18] * 3], "agx_golden")[0]
    assert punchy[1] < grey[1]
    assert golden[0] > golden[1] > golden[2]


def test_agx_rolls_off_without_hue_clipping():
    prev, sats = -1.0, []
    for e in range(-4, 15):
        k = 2.0 ** e
        o = O.tonemap_pixel([[1.0 * k, 0.5 * k, 0.1 * k]], "agx")[0]
        assert np.all(np.isfinite(o)) and np.all((o >= 0) & (o <= 1))
        assert o[0] >= o[1] - 2e-3 and o[1] >= o[2] - 2e-3
        luma = 0.2126 * o[0] + 0.7152 * o[1] + 0.0722 * o[2]
        assert luma >= prev - 1e-5
        prev = luma
        sats.append((o[0] - o[2]) / max(o[0], 1e-6))
    assert all(b <= a + 1e-3 for a, b in zip(sats[8:], sats[9:]))
    assert sats[-1] < 0.1
    assert np.all(O.tonemap_pixel([[16.0, 8.0, 1.6]], "clamp")[0] == 1.0)
    agx = O.tonemap_pixel([[16.0, 8.0, 1.6]], "agx")[0]
    assert agx[0] - agx[2] > 0.15
    huge = O.tonemap_pixel([[1e6, 2e5, 1e4]], "agx")[0]
    assert np.all((huge > 0.9) & (huge <= 1.0))


def test_other_operators_are_sane():
    for op in ["aces", "khronos", "reinhard", "clamp"]:
        assert np.all(np.abs(O.tonemap_pixel([[0, 0, 0]], op)) < 1e-3)
        hi = O.tonemap_pixel([[50.0, 40.0, 30.0]], op)[0]
        assert np.all((hi >= 0) & (hi <= 1))
        assert np.all(np.isfinite(O.tonemap_pixel([[np.nan, np.inf, -1.0]], op)))
    assert abs(O.tonemap_pixel([[0.18] * 3], "aces")[0, 1] - 0.106) < 0.01
    kn = O.tonemap_pixel([[0.5, 0.3, 0.2]], "khronos")[0]
    assert abs(kn[0] - 0.46) < 1e-3 and abs(kn[2] - 0.16) < 1e-3


def test_saturation_and_exposure():
    assert abs(O.tonemap_pixel([[1.0] * 3], "reinhard", ev=1.0)[0, 0] - 2.0 / 3.0) < 1e-5
    o = O.tonemap_pixel([[1.0, 0.2, 0.05]], "agx", sat=0.0)[0]
    assert abs(o[0] - o[1]) < 2e-3 and abs(o[1] - o[2]) < 2e-3


def _test_image(w, h):
    x = np.arange(w, dtype=np.float32) / np.float32(w - 1)
    c = np.float32(0.18) * np.power(np.float32(2.0), np.float32(-10.0) + np.float32(22.0) * x)
    hue = (np.arange(h, dtype=np.float32) / np.float32(h))[:, None]
    return np.stack([c[None, :] * (1 - hue), np.broadcast_to(c * np.float32(0.6), (h, w)), c[None, :] * hue], -1)


@pytest.mark.parametrize("op", OPS)
def test_fast_path_matches_exact(op):
    img = _test_image(512, 64).reshape(-1, 3)
    fast = O.develop_encoded(img, op, ev=0.5, sat=1.2)
    exact = O.srgb_encode(O.tonemap_pixel(img, op, ev=0.5, sat=1.2))
    assert np.abs(fast - exact).max() < 2e-4


def test_develop_dithers_and_keeps_extremes():
    img = np.zeros((4, 256, 3), np.float32)
    img[1] = 1e9
    img[2] = 0.18
    out = O.develop(img, "clamp")
    assert np.all(out[0] == 0) and np.all(out[1] == 255)
    row = out[2].astype(np.float64)
    want = float(O.srgb_encode(0.18)) * 255.0
    assert abs(row.mean() - want) < 0.15
    assert np.any(row != row.flat[0])


def _gauss(i):
    """denoise.rs:731-736: a hashed standard normal (computed in float64 here)."""
    def h(x):
        x = np.asarray(x, np.uint64) & np.uint64(0xFFFFFFFF)
        x ^= x >> np.uint64(16)
        x = (x * np.uint64(0x7FEB352D)) & np.uint64(0xFFFFFFFF)
        x ^= x >> np.uint64(15)
        x = (x * np.uint64(0x846CA68B)) & np.uint64(0xFFFFFFFF)
        return x ^ (x >> np.uint64(16))
    i = np.asarray(i, np.uint64)
    u1 = (h(i * 2).astype(np.float64) + 1.0) / 4294967296.0
    u2 = h(i * 2 + 1).astype(np.float64) / 4294967296.0
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * np.pi * u2)


def gradient(w, h, sigma):
    """denoise.rs:746-768: a smooth colour gradient plus noise, flat albedo 0.8
    and normal +z."""
    x = np.arange(w) / (w - 1)
    y = np.arange(h) / (h - 1)
    X, Y = np.meshgrid(x, y)
    clean = np.stack([0.1 + 0.8 * X, 0.2 + 0.5 * Y, 0.6 - 0.4 * X], -1).astype(np.float32)
    idx = (np.arange(h)[:, None] * w + np.arange(w)[None, :]) * 3
    noise = np.stack([_gauss(idx), _gauss(idx + 1), _gauss(idx + 2)], -1)
    noisy = np.maximum(clean + sigma * noise, 0).astype(np.float32)
    albedo = np.full_like(clean, 0.8)
    normal = np.zeros_like(clean)
    normal[..., 2] = 1.0
    return clean, noisy, albedo, normal


def _mse(a, b):
    return float(np.mean((a.astype(np.float64) - b) ** 2))


def test_atrous_reduces_mse():
    clean, noisy, albedo, normal = gradient(256, 192, 0.08)
    for a, n in [(None, None), (albedo, normal)]:
        out = O.atrous(noisy, a, n)
        assert _mse(out, clean) < 0.15 * _mse(noisy, clean)


def test_atrous_keeps_material_edges():
    w, h = 128, 96
    v = np.where(np.arange(w) < 64, 0.2, 0.8)[None, :].repeat(h, 0)
    clean = np.repeat(v[..., None], 3, -1).astype(np.float32)
    idx = (np.arange(h)[:, None] * w + np.arange(w)[None, :]) * 3
    noisy = (clean + 0.05 * np.stack([_gauss(idx), _gauss(idx + 1), _gauss(idx + 2)], -1)).astype(np.float32)
    out = O.atrous(noisy, clean, None)
    band = slice(61, 67)
    before = _mse(noisy[:, band], clean[:, band])
    after = _mse(out[:, band], clean[:, band])
    assert after < 0.5 * before


def test_atrous_of_a_constant_is_the_constant():
    img = np.full((40, 50, 3), 0.37, np.float32)
    out = O.atrous(img)
    assert np.allclose(out, 0.37, rtol=0, atol=1e-6)
