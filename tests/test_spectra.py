"""Known-answer tests of the host spectral library (the scene-build half of
argent-spectra and argent-core's tables), against the values the reference's
own tests and module docs pin.  CPU only."""
import ctypes as C

import numpy as np
import pytest

import argent_amd as ag
from argent_amd import _abi as A
from argent_amd import _ptr

L = A.lib()


def tab(fn, *args):
    out = np.zeros(A.SPEC_N, np.float32)
    A.check(fn(*args, _ptr(out)))
    return out


def tab_eval(v, l):
    """tab.rs:44-52"""
    x = min(max(np.float32(l) - np.float32(360.0), np.float32(0.0)), np.float32(470.0))
    i = min(int(x), 469)
    f = np.float32(x - np.float32(i))
    return v[i] * (np.float32(1.0) - f) + v[i + 1] * f


# alloys.rs:62-78: the module-doc F0 table (linear sRGB, 3 decimals)
F0_TABLE = {
    "fine-silver": (0.989, 0.984, 0.978), "sterling": (0.990, 0.964, 0.927), "gold24k": (1.040, 0.730, 0.363),
    "gold18k": (0.948, 0.734, 0.476), "gold14k": (0.948, 0.748, 0.531), "gold10k": (0.958, 0.765, 0.580),
    "rose14k": (0.937, 0.650, 0.486), "white14k": (0.797, 0.765, 0.742), "rhodium": (0.797, 0.765, 0.742),
    "platinum": (0.679, 0.640, 0.584), "palladium": (0.733, 0.697, 0.652), "copper": (0.935, 0.620, 0.523),
    "bronze": (0.834, 0.672, 0.492), "brass": (0.910, 0.778, 0.423),
}


@pytest.mark.parametrize("alloy", sorted(F0_TABLE))
def test_alloy_f0_table(alloy):
    got = ag.alloy_f0_srgb(alloy)
    assert np.all(np.abs(got - np.array(F0_TABLE[alloy])) <= 6e-4), (alloy, got)


def test_metal_kats():
    """metals.rs:179-266 spot values through the pure-metal alloys."""
    n, k = ag.alloy_ior("gold24k")
    lam = 360.0 + 5.0 * np.arange(95)
    assert abs(np.interp(659.5, lam, n) - 0.14) < 0.01 and abs(np.interp(659.5, lam, k) - 3.697) < 0.01
    n, k = ag.alloy_ior("copper")
    assert abs(np.interp(548.6, lam, n) - 1.02) < 0.02 and abs(np.interp(548.6, lam, k) - 2.577) < 0.01
    for name, want, tol in [("palladium", (0.733, 0.697, 0.652), 0.005), ("platinum", (0.673, 0.637, 0.585), 0.01),
                            ("copper", (0.955, 0.638, 0.538), 0.025)]:
        assert np.all(np.abs(ag.alloy_f0_srgb(name) - np.array(want)) < tol), name
    for a in F0_TABLE:  # physical: a metal has eps1 = n^2 - k^2 < 0
        n, k = ag.alloy_ior(a)
        assert np.all(n > 0) and np.all(k > 0) and np.all(n * n - k * k < 0), a


def test_brass_and_end_members():
    """alloys.rs:556-590: brass is Querry's measured alloy; pure end members of
    the karat model are the measured metals."""
    assert np.all(np.abs(ag.alloy_f0_srgb("brass") - np.array([0.910, 0.778, 0.423])) < 5e-3)
    for (au, agv, cu), metal in [((1, 0, 0), "gold24k"), ((0, 1, 0), "fine-silver"), ((0, 0, 1), "copper")]:
        n = np.zeros(95, np.float32)
        k = np.zeros(95, np.float32)
        A.check(L.ag_karat_gold(au, agv, cu, _ptr(n), _ptr(k)))
        n2, k2 = ag.alloy_ior(metal)
        assert np.max(np.abs(n - n2)) < 1e-4 and np.max(np.abs(k - k2)) < 1e-4, metal


def gem(name):
    kind = C.c_int32()
    ior = np.zeros(A.SPEC_N, np.float32)
    ab = np.zeros(A.SPEC_N, np.float32)
    body = np.zeros(3, np.float32)
    hb = C.c_int32()
    A.check(L.ag_gem_optics(name.encode(), C.byref(kind), _ptr(ior), _ptr(ab), _ptr(body), C.byref(hb)))
    return kind.value, ior, ab, body, hb.value


def test_gem_dispersion_kats():
    """gems.rs:515-542 (n evaluated from the 1 nm bake)."""
    def n(name, l):
        return tab_eval(gem(name)[1], l)

    def bg(name):
        return n(name, 430.8) - n(name, 686.7)

    assert abs(n("diamond", 589.3) - 2.417) < 0.002
    assert abs(bg("diamond") - 0.044) < 0.002
    assert abs((n("diamond", 486.13) - n("diamond", 656.27)) - 0.0256) < 0.001
    assert abs(bg("cubic_zirconia") - 0.060) < 0.003
    assert abs(n("moissanite", 589.3) - 2.648) < 0.005
    assert abs(n("sapphire", 589.3) - 1.768) < 0.001 and abs(bg("sapphire") - 0.018) < 0.001
    nd = n("glass", 587.56)
    assert abs(nd - 1.5168) < 2e-4
    assert abs((nd - 1.0) / (n("glass", 486.13) - n("glass", 656.27)) - 64.17) < 0.1
    for name, nD, b in [("emerald", 1.580, 0.014), ("garnet", 1.790, 0.027), ("tanzanite", 1.694, 0.030)]:
        assert abs(n(name, 589.3) - nD) < 2e-4 and abs(bg(name) - b) < 2e-4
    assert bg("moissanite") > bg("cubic_zirconia") > bg("diamond") > bg("sapphire") > bg("amethyst")


def test_gem_kinds_and_colourless():
    assert gem("onyx")[0] == 2 and gem("black_opal")[0] == 1 and gem("diamond")[0] == 0
    for name in ["diamond", "cubic_zirconia", "white_topaz", "glass"]:
        ab = gem(name)[2]
        t = np.exp(-ab * 4.0).astype(np.float32)
        rgb = np.zeros(3, np.float32)
        A.check(L.ag_reflectance_srgb(_ptr(t), _ptr(rgb)))
        assert np.all(np.abs(rgb - 1.0) < 0.01), (name, rgb)


def test_cie_kats():
    """cie.rs:191-211"""
    out = np.zeros(3, np.float32)
    L.ag_cie_xyz_bar(555.0, _ptr(out))
    assert out[1] == np.float32(1.0)
    L.ag_cie_xyz_bar(360.0, _ptr(out))
    assert list(out) == [np.float32(0.0001299), np.float32(3.917e-6), np.float32(0.0006061)]
    for l in (359.9, 830.1, float("nan")):
        L.ag_cie_xyz_bar(l, _ptr(out))
        assert list(out) == [0.0, 0.0, 0.0]
    white = np.ones(A.SPEC_N, np.float32)
    assert abs(L.ag_luminance_y(_ptr(white)) - 1.0) < 1e-5


def test_illuminants_and_emitters():
    """spectrum.rs:502-546 and studio.rs:15-25: emitter luminance = intensity."""
    w = tab(L.ag_spectrum_rgb_illuminant, _ptr(np.ones(3, np.float32)))
    assert abs(L.ag_luminance_y(_ptr(w)) - 1.0) < 1e-5
    for kelvin, inten in [(1900.0, 9.0), (5600.0, 3.2), (5800.0, 1.7), (9000.0, 1.4)]:
        e = tab(L.ag_spectrum_emitter, kelvin, inten)
        assert abs(L.ag_luminance_y(_ptr(e)) - inten) < 1e-3 * inten


def test_albedo_round_trip():
    """spectrum.rs:461-482: sigmoid upsampling round trips the sRGB cube."""
    vals = [0.0, 0.02, 0.1, 0.3, 0.6, 0.9, 0.98, 1.0]
    worst = 0.0
    rgb = np.zeros(3, np.float32)
    for r in vals:
        for g in vals:
            for b in vals:
                s = tab(L.ag_spectrum_rgb_albedo, _ptr(np.array([r, g, b], np.float32)))
                A.check(L.ag_reflectance_srgb(_ptr(s), _ptr(rgb)))
                worst = max(worst, float(np.max(np.abs(rgb - np.array([r, g, b])))))
    assert worst < 1e-4, worst
    for gv in [0.0, 0.04, 0.18, 0.5, 1.0]:  # greys are exact constants
        s = tab(L.ag_spectrum_rgb_albedo, _ptr(np.array([gv, gv, gv], np.float32)))
        assert np.all(s == np.float32(gv))


def test_fresnel_kats():
    """fresnel.rs:122-148"""
    assert abs(L.ag_fresnel_dielectric(1.0, 1.5) - 0.04) < 1e-4
    assert L.ag_fresnel_dielectric(-0.1, 1.5) == 1.0
    n, k = 0.2, 3.0
    want = ((n - 1) ** 2 + k * k) / ((n + 1) ** 2 + k * k)
    assert abs(L.ag_fresnel_conductor1(1.0, n, k) - want) < 1e-5
    assert L.ag_fresnel_conductor1(0.0, n, k) > 0.999


def test_white_balance():
    """color.rs:66-79"""
    w = np.zeros(3, np.float32)
    L.ag_blackbody_white(5600.0, _ptr(w))
    m = np.zeros(9, np.float32)
    L.ag_camera_matrix(1, _ptr(w), _ptr(m))
    rgb = m.reshape(3, 3) @ w
    assert abs(rgb[0] - rgb[1]) < 0.01 and abs(rgb[2] - rgb[1]) < 0.01
    d65 = np.array([0.95047, 1.0, 1.08883], np.float32)
    L.ag_camera_matrix(1, _ptr(d65), _ptr(m))
    m0 = np.zeros(9, np.float32)
    L.ag_camera_matrix(0, _ptr(d65), _ptr(m0))
    assert np.allclose(m.reshape(3, 3) @ [0.3, 0.4, 0.5], m0.reshape(3, 3) @ [0.3, 0.4, 0.5], atol=1e-3)


def energy_lookup(e, mu, alpha):
    """microfacet.rs:100-110"""
    x = min(min(max(mu, 0.0), 1.0) * 31.0, 31.0 - 1e-4)
    y = min(min(max(np.sqrt(alpha), 0.0), 1.0) * 31.0, 31.0 - 1e-4)
    xi, yi = int(x), int(y)
    fx, fy = x - xi, y - yi
    g = lambda i, j: e[j * 32 + i]
    a = g(xi, yi) * (1 - fx) + g(xi + 1, yi) * fx
    b = g(xi, yi + 1) * (1 - fx) + g(xi + 1, yi + 1) * fx
    return min(max(a * (1 - fy) + b * fy, 0.05), 1.0)


def test_energy_table_kats():
    """microfacet.rs:146-153"""
    e = np.zeros(1024, np.float32)
    L.ag_energy_table(_ptr(e))
    assert energy_lookup(e, 0.8, max(0.05 ** 2, 1e-4)) > 0.97
    v = energy_lookup(e, 0.5, 0.6 ** 2)
    assert 0.5 < v < 0.95


def test_studio_rigs_and_light_cdf():
    """studio.rs:87-197 + scene.rs:88-110"""
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    t = np.array([[0, 1, 2]], np.uint32)
    counts = {"noir": (5, 5), "white": (8, 4), "velvet": (3, 3), "ember": (4, 4), "window": (3, 2)}
    for studio, (n_panels, n_lights) in counts.items():
        d = ag.SceneDesc(p, t)
        d.add_material(ag.Material.metal("sterling"))
        d.apply_studio(studio, 15.0, -5.0)
        assert d.d.n_panels == n_panels and d.d.floor.present == 1 and d.d.n_materials == 2
        assert abs(d.d.floor.y - (-5.005)) < 1e-6
        nl = C.c_int32()
        lights = (C.c_int32 * 16)()
        cdf = (C.c_float * 16)()
        pick = C.c_float()
        A.check(L.ag_desc_light_cdf(C.byref(d.d), C.byref(nl), lights, cdf, C.byref(pick)))
        assert nl.value == n_lights and abs(cdf[nl.value - 1] - 1.0) < 1e-6 and pick.value == 0.0
        assert all(cdf[i] <= cdf[i + 1] for i in range(nl.value - 1))
    with pytest.raises(ag.ArgentError):
        ag.SceneDesc(p, t).apply_studio("moon", 15.0, 0.0)
    with pytest.raises(ag.ArgentError):
        ag.Material.metal("unobtainium")
