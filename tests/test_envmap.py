"""The importance-sampled HDRI (envmap.rs) pinned on the CPU: the LinearRgb
bases against an independent Planck curve and the partition of unity they
are built on, the RGB round trip through the CIE observer that defines them
(envmap.rs:10-13), the sampling tables of the host library against the
oracle's restatement of EnvData::load, the pdf's normalisation over the
sphere, and the histogram of EnvMap::sample against EnvMap::pdf texel by
texel.  The reference holds no test or fixture for this file."""
import ctypes as C

import numpy as np
import pytest

import argent_amd as ag
from argent_amd import _abi as A
from argent_amd import _ptr, synth

O = pytest.importorskip("oracle")

SRGB_FROM_XYZ = np.array([[3.2404542, -1.5371385, -0.4985314],
                          [-0.969266, 1.8760108, 0.041556],
                          [0.0556434, -0.2040259, 1.0572252]])


def _host_env(rgb, exposure=0.0):
    h, w = rgb.shape[:2]
    env = A.Env()
    ag.check(A.lib().ag_env_from_rgb(_ptr(np.ascontiguousarray(rgb, np.float32).reshape(-1)), w, h, exposure, 0.0,
                                     C.byref(env)))
    weights = np.ctypeslib.as_array(env.weights, (h * w * 3,)).reshape(h, w, 3).copy()
    marg = np.ctypeslib.as_array(env.marginal, (h + 1,)).copy()
    cond = np.ctypeslib.as_array(env.conditional, (h * (w + 1),)).reshape(h, w + 1).copy()
    basis = np.array([env.basis[c][:] for c in range(3)], np.float32)
    total = env.total
    A.lib().ag_env_free(C.byref(env))
    return weights, marg, cond, total, basis


def _planck(lam_nm, t):
    c, h, kb = 299792458.0, 6.62606957e-34, 1.3806488e-23
    l = np.asarray(lam_nm, np.float64) * 1e-9
    return 2 * h * c * c / (l ** 5 * np.expm1(h * c / (l * kb * t)))


def test_linear_rgb_bases():
    """Bases = (red, green, blue) smoothstep windows that sum to one, times a
    6504 K Planck curve scaled to 1 at Wien's peak (envmap.rs:31-49)."""
    b, _ = O.linear_rgb()
    lam = 360.0 + np.arange(471)
    bb = _planck(lam, 6504.0) / _planck(2.897772e-3 / 6504.0 * 1e9, 6504.0)
    assert np.allclose(b.sum(axis=0), bb, rtol=2e-6, atol=0)
    assert np.allclose(b[2][lam <= 470], bb[lam <= 470], rtol=2e-6)   # blue alone below 470 nm
    assert np.allclose(b[0][lam >= 610], bb[lam >= 610], rtol=2e-6)   # red alone above 610 nm
    g = (lam >= 515) & (lam <= 565)
    assert np.allclose(b[1][g], bb[g], rtol=2e-6)                      # green alone in 515..565 nm
    assert np.all(b >= 0)


def test_linear_rgb_round_trip():
    """Any linear sRGB colour comes back from weights -> spectrum -> XYZ ->
    sRGB (the reason for the inverse response matrix)."""
    b, inv = O.linear_rgb()
    cmf = np.empty((471, 3), np.float32)
    for i in range(471):
        A.lib().ag_cie_xyz_bar(360.0 + i, cmf[i].ctypes.data_as(A.f32p))
    rng = np.random.default_rng(3)
    for rgb in rng.random((40, 3)) * 4.0:
        w = inv.astype(np.float64) @ rgb
        spec = w @ b.astype(np.float64)
        xyz = spec @ cmf.astype(np.float64) / 106.85692
        back = SRGB_FROM_XYZ @ xyz
        assert np.allclose(back, rgb, rtol=1e-4, atol=1e-5), (rgb, back)


def test_env_tables_match_host():
    rgb = synth.hdri(64, 32, seed=5)
    hw, hm, hc, ht, hb = _host_env(rgb, 1.0)
    ow, om, oc, ot = O.env_build(rgb, 1.0)
    b, _ = O.linear_rgb()
    assert np.allclose(hb, b, rtol=1e-6, atol=0)
    assert np.allclose(hw, ow, rtol=1e-5, atol=1e-7)
    assert np.array_equal(hm, om) and np.array_equal(hc, oc)
    assert abs(ht - ot) <= 1e-6 * abs(ot)
    assert hm[0] == 0.0 and hm[-1] == 1.0 and np.all(np.diff(hm) >= 0)
    assert np.all(hc[:, 0] == 0.0) and np.allclose(hc[:, -1], 1.0) and np.all(np.diff(hc, axis=1) >= 0)


def _uv(d, rot=0.0):
    r = -np.deg2rad(rot)
    s, c = np.sin(r), np.cos(r)
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    x, z = c * x + s * z, -s * x + c * z
    u = np.mod(0.5 + np.arctan2(x, -z) / (2 * np.pi), 1.0)
    v = np.arccos(np.clip(y, -1, 1)) / np.pi
    return u, v


def _dir(u, v, rot=0.0):
    phi, th = (u - 0.5) * 2 * np.pi, v * np.pi
    d = np.stack([np.sin(th) * np.sin(phi), np.cos(th), -np.sin(th) * np.cos(phi)], axis=1)
    r = np.deg2rad(rot)
    s, c = np.sin(r), np.cos(r)
    return np.stack([c * d[:, 0] + s * d[:, 2], d[:, 1], -s * d[:, 0] + c * d[:, 2]], axis=1)


@pytest.mark.parametrize("rot", [0.0, 35.0])
def test_env_pdf_integrates_to_one(rot):
    rgb = synth.hdri(48, 24, seed=6)
    _, m, c, _ = O.env_build(rgb)
    k = 4  # sub-cells per texel and axis (midpoint rule in u, v)
    u = (np.arange(48 * k) + 0.5) / (48 * k)
    v = (np.arange(24 * k) + 0.5) / (24 * k)
    uu, vv = np.meshgrid(u, v)
    d = _dir(uu.ravel(), vv.ravel(), rot).astype(np.float32)
    p = O.env_pdf(m, c, d, rot).astype(np.float64)
    domega = (2 * np.pi / (48 * k)) * (np.pi / (24 * k)) * np.sin(np.pi * vv.ravel())
    assert abs((p * domega).sum() - 1.0) < 2e-3


@pytest.mark.parametrize("rot", [0.0, 35.0])
def test_env_sampling_follows_pdf(rot):
    w, h = 32, 16
    rgb = synth.hdri(w, h, seed=7)
    _, m, c, _ = O.env_build(rgb)
    n = 1 << 19
    rng = np.random.default_rng(11)
    u = rng.random((n, 2)).astype(np.float32)
    d, pdf = O.env_sample(m, c, u, rot)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    assert np.array_equal(pdf, O.env_pdf(m, c, d, rot))  # sample's pdf is pdf() of its direction
    uu, vv = _uv(d.astype(np.float64), rot)
    x = np.minimum((uu * w).astype(int), w - 1)
    y = np.minimum((vv * h).astype(int), h - 1)
    counts = np.bincount(y * w + x, minlength=w * h).astype(np.float64)
    expect = n * (np.diff(c, axis=1) * np.diff(m)[:, None]).ravel()
    big = expect > 200
    chi2 = ((counts[big] - expect[big]) ** 2 / expect[big]).sum() / big.sum()
    assert chi2 < 1.5, chi2
    assert counts[~big].sum() <= expect[~big].sum() * 1.2 + 50
