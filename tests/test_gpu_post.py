"""Preview development on the GPU against the oracle (argent-post tonemap.rs
and denoise.rs restated): develop is exact -- the same per-grade tables built
with the same libm, the same f32 expressions and the same dither -- so its
bytes must be equal; the à-trous denoiser differs only through expf (device
vs host libm), so its result is held to 1e-4 relative, and its developed bytes
to at most one code on at most 0.1 % of values."""
import numpy as np
import pytest

import argent_amd as ag
import oracle as O
from argent_amd import synth, workloads

from test_post_oracle import OPS, _test_image, gradient

pytestmark = pytest.mark.gpu


def _hard_image():
    img = _test_image(512, 64)
    img[0, :8] = [[0, 0, 0], [1e9, 1e9, 1e9], [np.nan, 0.5, 0.2], [np.inf, -1.0, 0.3], [-np.inf, 2.0, np.nan],
                  [65504.0, 1e-30, 0.0], [0.18, 0.18, 0.18], [3.0, 1e-8, 7.0]]
    return img


@pytest.mark.parametrize("op", OPS)
def test_develop_is_bit_exact(ctx, op):
    img = _hard_image()
    for ev, sat in [(0.0, 1.0), (0.5, 1.2), (-1.5, 0.0)]:
        got = ag.develop_image(ctx, img, op, ev, sat)
        want = O.develop(img, op, ev, sat)
        assert np.array_equal(got, want), (op, ev, sat, np.argwhere(got != want)[:5])


def _close(got, want):
    err = np.abs(got.astype(np.float64) - want)
    assert np.all(err <= 1e-4 * np.maximum(np.abs(want), 1e-3)), float(err.max())


def test_denoise_matches_oracle(ctx):
    clean, noisy, albedo, normal = gradient(256, 192, 0.08)
    for a, n in [(None, None), (albedo, None), (albedo, normal)]:
        got = ag.denoise_image(ctx, noisy, a, n)
        _close(got, O.atrous(noisy, a, n))
        assert np.mean((got - clean) ** 2) < 0.15 * np.mean((noisy - clean) ** 2)


def _film(ctx):
    pa = workloads.Piece(*synth.torus_ring(seed=1, nu=128, nv=32))
    sc = ag.Scene(ctx, pa.scene("gold14k", 0.07, 0.8, "noir"))
    cam, turn = pa.shot("hero", 100.0, 0.0, 96 / 64)
    film = ag.Film(ctx, 96, 64)
    ag.render_views(ctx, sc, [dict(camera=cam, film=film, count=8, seed=3, turntable_deg=turn)],
                    workloads.path_settings(8))
    sc.close()
    return film


def test_film_development_matches_oracle(ctx):
    film = _film(ctx)
    color, albedo, normal = film.color(), film.albedo(), film.normal()
    assert np.array_equal(film.develop("agx_punchy"), O.develop(color, "agx_punchy"))
    assert np.array_equal(film.develop("khronos", 0.7, 1.1), O.develop(color, "khronos", 0.7, 1.1))
    den = film.denoise()
    ref = O.atrous(color, albedo, normal)
    _close(den, ref)
    got8, want8 = film.develop("agx_punchy", denoise=True), O.develop(ref, "agx_punchy")
    d = np.abs(got8.astype(int) - want8.astype(int))
    assert d.max() <= 1 and (d > 0).mean() <= 1e-3
    film.close()
