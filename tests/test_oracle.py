"""The oracle pinned against the reference's own tests and closed forms
(CPU only): the BVH against brute force (bvh.rs:572-598), the white furnace
(furnace.rs:46-95), wavelength sampling (spec.rs:226-245, cie.rs:262-272),
the sampler's stratification, and agreement of the oracle's independently
written tables with the product's host library."""
import numpy as np
import pytest

import argent_amd as ag
from argent_amd import _abi as A
from argent_amd import _ptr

O = pytest.importorskip("oracle")

from test_gpu_parity import _furnace_desc, grid, xorshift_rays  # noqa: E402


def test_bvh_matches_brute_force():
    p, t = grid(60)
    o, d = xorshift_rays()
    desc = ag.SceneDesc(p, t)
    desc.add_material(ag.Material.diffuse([0.5, 0.5, 0.5]))
    osc = O.OracleScene(desc)
    bt, _, _, bp = O.brute_intersect(p, t, o, d)
    hit = bp != 0xFFFFFFFF
    for mode in (O.MODE_REFERENCE, O.MODE_ROBUST):
        tt, _, _, pp = osc.intersect(o, d, mode=mode)
        assert np.array_equal(pp != 0xFFFFFFFF, hit)
        assert np.all(np.abs(tt[hit] - bt[hit]) < 1e-4)
        assert np.array_equal(osc.occluded(o, d, mode=mode), hit)


def test_bvh_random_soup_vs_brute_force():
    """Random triangle soup with shared edges and coplanar pieces: the robust
    traversal reproduces brute force exactly (ids, t, u, v)."""
    rng = np.random.default_rng(9)
    p, t = grid(12)
    q = rng.random((600, 3)).astype(np.float32) * 10 - 5
    tq = np.arange(600, dtype=np.uint32).reshape(-1, 3) + len(p)
    P = np.concatenate([p, q]).astype(np.float32)
    T = np.concatenate([t, tq]).astype(np.uint32)
    o = (rng.random((5000, 3)) * 14 - 7).astype(np.float32)
    d = rng.normal(size=(5000, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    desc = ag.SceneDesc(P, T)
    desc.add_material(ag.Material.diffuse([0.5, 0.5, 0.5]))
    osc = O.OracleScene(desc)
    bt, bu, bv, bp = O.brute_intersect(P, T, o, d)
    rt, ru, rv, rp = osc.intersect(o, d, mode=O.MODE_ROBUST)
    assert np.array_equal(rp, bp)
    h = bp != 0xFFFFFFFF
    assert np.array_equal(rt[h], bt[h]) and np.array_equal(ru[h], bu[h]) and np.array_equal(rv[h], bv[h])


def test_wavelength_sampling():
    for i in range(1000):
        lam, pdf = O.sample_wavelengths((i + 0.5) / 1000)
        assert np.all((lam >= 360) & (lam <= 830)) and np.all(pdf > 0)
    # cie.rs:262-272: MC estimate of a constant spectrum's luminance
    n = 4096
    cmf_y = np.array([A.lib().ag_cie_xyz_bar(0.0, _ptr(np.zeros(3, np.float32))) for _ in range(1)])
    y = 0.0
    ybar = np.zeros(3, np.float32)
    for i in range(n):
        lam, pdf = O.sample_wavelengths((i + 0.5) / n)
        s = 0.0
        for k in range(4):
            A.lib().ag_cie_xyz_bar(float(lam[k]), _ptr(ybar))
            s += ybar[1] / pdf[k]
        y += s / (4 * 106.85692)
    assert abs(y / n - 1.0) < 2e-3


def test_sampler_is_stratified():
    """Owen-scrambled Sobol: dims (4k, 4k+1) of 256 samples form a (0,8,2)-net."""
    for dim in (0, 8, 20):
        pts = np.array([[O.rng_dim(777, i, dim), O.rng_dim(777, i, dim + 1)] for i in range(256)])
        assert np.all((pts >= 0) & (pts < 1))
        for a, b in [(16, 16), (256, 1), (1, 256), (4, 64)]:
            cells = set((int(x * a), int(y * b)) for x, y in pts)
            assert len(cells) == 256, (dim, a, b)


@pytest.mark.parametrize("name,tol", [("grey", 0.04), ("glass", 0.05), ("diamond", 0.06)])
def test_furnace_oracle(name, tol):
    mat = ag.Material.diffuse([0.5, 0.5, 0.5]) if name == "grey" else ag.Material.gem(name)
    desc = _furnace_desc(mat)
    cam = ag.look_at([0, 0, 200], [0, 0, 0], [0, 1, 0], 100.0, 0.0, 1.0)
    osc = O.OracleScene(desc)
    sums = np.zeros((3, 64, 64, 3), np.float32)
    osc.render_pass(cam, 64, 64, sums, 0, ag.PathSettings(), 256, 1)
    img = sums[0] / 256
    bg = img[1, 1]
    centre = img[30:34, 30:34].reshape(-1, 3).mean(0)
    assert np.all(np.abs(bg - 1.0) < 0.03), bg
    if name == "grey":
        assert np.all(np.abs(centre - 0.5) < tol), centre
    else:
        assert np.all(np.abs(centre - bg) < tol), (centre, bg)


def test_oracle_tables_match_host_library():
    e = np.zeros(1024, np.float32)
    A.lib().ag_energy_table(_ptr(e))
    assert np.array_equal(e, O.energy_table())
    for white in [None, np.array([0.98, 1.0, 0.9], np.float32)]:
        m = np.zeros(9, np.float32)
        w = np.zeros(3, np.float32) if white is None else white
        A.lib().ag_camera_matrix(0 if white is None else 1, _ptr(w), _ptr(m))
        assert np.array_equal(m.reshape(3, 3), O.camera_matrix(white))
    for c in np.linspace(-1, 1, 41):
        assert O.fresnel_dielectric(float(c), 1.7) == A.lib().ag_fresnel_dielectric(float(c), 1.7)
        assert O.fresnel_conductor1(float(c), 0.2, 3.1) == A.lib().ag_fresnel_conductor1(float(c), 0.2, 3.1)
