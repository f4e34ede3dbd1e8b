"""Scenes whose rendered values are fixed by closed forms, for pinning the
integrator's light transport (integrator.rs:50-221) independently of any
restatement of it.  Shared by the oracle tests (CPU) and the GPU tests.

Each case names what the mathematics fixes:
  * goldprobe (examples/goldprobe.rs:42-55): a polished metal sphere in a
    uniform white environment; the centre pixel is the alloy's normal-incidence
    reflectance F0, whose linear-sRGB values the reference tabulates
    (alloys.rs:62-78).
  * uniform furnaces with NEE: a convex object inside uniform radiance, given
    by an importance-sampled constant HDRI (env_pick 1), and by the same HDRI
    plus a softbox of the same radiance in front of the object (env_pick 0.5):
    every direction carries the same radiance, so a diffuse sphere shows its
    albedo and a rough conductor whose Fresnel factor is 1 shows 1 (the
    multiple-scattering compensation of bsdf.rs:55-65 restores the energy the
    single-scattering lobe loses).  Every MIS weight (integrator.rs:70-98,
    :130-178) enters; a wrong one breaks the uniformity.
  * a Lambertian plane under one rectangular emitter: the outgoing radiance is
    rho/pi times the irradiance, which Lambert's polygon formula gives in
    closed form.
  * an absorbing slab (Beer-Lambert, integrator.rs:64-67) of index eta in a
    uniform environment: R + (1-R)^2 T / (1 - R T), with R the Fresnel
    reflectance at the ray's incidence and T = exp(-sigma d / cos theta_t).
"""
import numpy as np

import argent_amd as ag
from argent_amd import _abi as A

from test_gpu_parity import sphere

GOLDPROBE_ALLOYS = ("sterling", "gold24k", "gold18k", "gold14k", "rose14k", "bronze")
# alloys.rs:62-78, linear sRGB (D65) normal-incidence reflectance, 3 decimals
F0_DOC = {"sterling": (0.990, 0.964, 0.927), "gold24k": (1.040, 0.730, 0.363), "gold18k": (0.948, 0.734, 0.476),
          "gold14k": (0.948, 0.748, 0.531), "rose14k": (0.937, 0.650, 0.486), "bronze": (0.834, 0.672, 0.492)}


def white_spectrum():
    """RgbIlluminant([1, 1, 1]) baked to the 1 nm table (spectrum.rs:380-420)."""
    out = np.zeros(A.SPEC_N, np.float32)
    A.check(A.lib().ag_spectrum_rgb_illuminant(ag._ptr(np.ones(3, np.float32)), ag._ptr(out)))
    return out


def front_camera(w=64, h=64):
    """furnace.rs:52: 200 mm in front of the origin, 100 mm lens, pinhole."""
    return ag.look_at([0, 0, 200], [0, 0, 0], [0, 1, 0], 100.0, 0.0, w / h)


def pixel_rays(cam, w, h):
    """Directions of pinhole camera rays through pixel centres (camera.rs:57-70
    with no jitter): (h, w, 3)."""
    sx = (np.arange(w) + 0.5) * (2.0 / w) - 1.0
    sy = 1.0 - (np.arange(h) + 0.5) * (2.0 / h)
    X, Y = np.meshgrid(sx, sy)
    f, r, u = (np.array(getattr(cam, k), np.float64) for k in ("forward", "right", "up"))
    d = f + r * ((X + cam.shift[0]) * cam.half_w)[..., None] + u * ((Y + cam.shift[1]) * cam.half_h)[..., None]
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def sphere_desc(mat, n=64):
    p, t, nrm = sphere(10.0, n)
    desc = ag.SceneDesc(p, t, normals=nrm)
    desc.add_material(mat)
    return desc


def goldprobe_desc(alloy):
    """goldprobe.rs:45-49: Material::metal(alloy, 0.02) (no patina) in a white gradient furnace."""
    desc = sphere_desc(ag.Material.metal(alloy, 0.02, None))
    desc.set_gradient_env([1, 1, 1], [1, 1, 1], [1, 1, 1], 1.0)
    return desc


def unit_fresnel_conductor(roughness):
    """A GGX conductor whose Fresnel factor is 1 to 4e-6 at every angle and
    wavelength (n = 1, k = 1000)."""
    m = ag.Material.metal("sterling", roughness, None)
    for i in range(A.GRID_N):
        m.metal_n[i] = 1.0
        m.metal_k[i] = 1000.0
    return m


def constant_hdri(w=64, h=32, rgb=(1.0, 1.0, 1.0)):
    return np.broadcast_to(np.asarray(rgb, np.float32), (h, w, 3)).copy()


def hdri_texel_spectrum(desc):
    """Radiance spectrum of texel 0 of the scene's map: sum of the three
    LinearRgb bases (envmap.rs:31-47) by the texel's weights."""
    e = desc.d.env
    wts = np.ctypeslib.as_array(e.weights, shape=(3,)).astype(np.float64)
    basis = np.array([list(e.basis[k]) for k in range(3)], np.float64)
    return (wts[:, None] * basis).sum(0).astype(np.float32)


def furnace_desc(mat, env="hdri"):
    """Uniform radiance around a 10 mm sphere.  env: "gradient" (no NEE),
    "hdri" (a constant importance-sampled map, env_pick 1), "hdri+panel" (the
    map plus a hard-edged softbox of the map's radiance 40 mm in front of the
    sphere, facing it and invisible to the camera: env_pick 0.5)."""
    desc = sphere_desc(mat)
    if env == "gradient":
        desc.set_gradient_env([1, 1, 1], [1, 1, 1], [1, 1, 1], 1.0)
        return desc
    desc.set_gradient_env([0, 0, 0], [0, 0, 0], [0, 0, 0], 1.0)
    desc.set_env_map(constant_hdri())
    if env == "hdri+panel":
        # hu x hv = (12,0,0) x (0,12,0) points along +z; the panel must face -z
        desc.add_panel([0, 0, 40], [0, 12, 0], [12, 0, 0], hdri_texel_spectrum(desc), camera_visible=False,
                       feather=0.0)
    return desc


# -------------------------------------------------------- plane under a panel
PLANE_HALF = 20.0
PANEL_CENTER = np.array([2.0, 12.0, -1.0])
PANEL_HU = np.array([6.0, 0.0, 0.0])
PANEL_HV = np.array([0.0, 0.0, 4.0])  # hu x hv = (0, -24, 0): the panel faces down


def plane_under_panel_desc(rho=0.5):
    """A 40 x 40 mm Lambertian square in the plane y = 0 (normal +y) under one
    hard-edged rectangular emitter of white-illuminant radiance, in a black
    gradient environment.  Every path ends at its first bounce: the panel is
    the only light and nothing else reflects."""
    h = PLANE_HALF
    p = np.array([[-h, 0, -h], [h, 0, -h], [h, 0, h], [-h, 0, h]], np.float32)
    t = np.array([[0, 2, 1], [0, 3, 2]], np.uint32)  # (v2-v0) x (v1-v0) winding: +y
    desc = ag.SceneDesc(p, t, normals=np.tile(np.float32([0, 1, 0]), (4, 1)))
    desc.add_material(ag.Material.diffuse([rho, rho, rho]))
    desc.set_gradient_env([0, 0, 0], [0, 0, 0], [0, 0, 0], 1.0)
    desc.add_panel(PANEL_CENTER, PANEL_HU, PANEL_HV, white_spectrum(), camera_visible=False, feather=0.0)
    return desc


def plane_camera(w=48, h=48):
    """Looking down at the plane from 150 mm, slightly off axis."""
    return ag.look_at([3.0, 150.0, 4.0], [0.0, 0.0, 0.0], [0, 0, -1], 100.0, 0.0, w / h)


def lambert_polygon_irradiance(x, n, verts):
    """Irradiance at points x (..., 3) with normal n from a Lambertian polygon
    of unit radiance (Lambert 1760): E = 1/2 sum_i gamma_i n . unit(r_i x r_i+1),
    r_i = v_i - x, gamma_i the angle between r_i and r_i+1.  The sign of the
    sum follows the winding; the emitter faces x, so its magnitude is taken."""
    x = np.asarray(x, np.float64)
    e = np.zeros(x.shape[:-1])
    m = len(verts)
    for i in range(m):
        a = np.asarray(verts[i], np.float64) - x
        b = np.asarray(verts[(i + 1) % m], np.float64) - x
        a /= np.linalg.norm(a, axis=-1, keepdims=True)
        b /= np.linalg.norm(b, axis=-1, keepdims=True)
        c = np.cross(a, b)
        cn = np.linalg.norm(c, axis=-1)
        gamma = np.arctan2(cn, (a * b).sum(-1))
        e += 0.5 * gamma * (c @ np.asarray(n, np.float64)) / cn
    return np.abs(e)


def plane_expected(cam, w, h, rho=0.5):
    """rho/pi * E(x) per pixel centre (h, w), nan where the pixel misses the square."""
    d = pixel_rays(cam, w, h)
    o = np.array(cam.pos, np.float64)
    t = -o[1] / d[..., 1]
    x = o + d * t[..., None]
    inside = (np.abs(x[..., 0]) < PLANE_HALF) & (np.abs(x[..., 2]) < PLANE_HALF)
    c, hu, hv = PANEL_CENTER, PANEL_HU, PANEL_HV
    verts = [c - hu - hv, c + hu - hv, c + hu + hv, c - hu + hv]
    e = lambert_polygon_irradiance(x, [0.0, 1.0, 0.0], verts)
    out = rho / np.pi * e
    out[~inside] = np.nan
    return out


# -------------------------------------------------------------- absorbing slab
SLAB = (40.0, 40.0, 4.0)


def slab_material(eta, sigma):
    """A transparent gem with constant index eta and constant absorption
    sigma (1/mm) at every wavelength."""
    m = ag.Material.gem("glass")
    for i in range(A.SPEC_N):
        m.spec_a[i] = eta
        m.spec_b[i] = sigma
    return m


def slab_desc(eta, sigma):
    from argent_amd import synth
    p, t = synth.cuboid(SLAB)
    desc = ag.SceneDesc(p, t)
    desc.add_material(slab_material(eta, sigma))
    desc.set_gradient_env([1, 1, 1], [1, 1, 1], [1, 1, 1], 1.0)
    return desc


def fresnel_unpolarised(cos_i, eta):
    """Exact unpolarised Fresnel reflectance of a dielectric interface, light
    arriving from index 1 (Born & Wolf), f64."""
    cos_i = np.asarray(cos_i, np.float64)
    sin_t2 = (1.0 - cos_i ** 2) / eta ** 2
    cos_t = np.sqrt(np.maximum(0.0, 1.0 - sin_t2))
    rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t)
    rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t)
    return 0.5 * (rs * rs + rp * rp), cos_t


def slab_expected(cam, w, h, eta, sigma):
    """Per pixel centre (h, w): R + (1-R)^2 T / (1 - R T) in units of the
    environment's radiance, for rays crossing both large faces (nan elsewhere)."""
    d = pixel_rays(cam, w, h)
    o = np.array(cam.pos, np.float64)
    hz = SLAB[2] / 2
    # entry at z = +hz, exit (straight through for eta = 1; refracted rays
    # stay within 0.3 mm of it) at z = -hz
    t0 = (hz - o[2]) / d[..., 2]
    x0 = o + d * t0[..., None]
    inside = (np.abs(x0[..., 0]) < SLAB[0] / 2 - 2) & (np.abs(x0[..., 1]) < SLAB[1] / 2 - 2)
    cos_i = np.abs(d[..., 2])
    r, cos_t = fresnel_unpolarised(cos_i, eta)
    tr = np.exp(-sigma * SLAB[2] / cos_t)
    out = r + (1 - r) ** 2 * tr / (1 - r * tr)
    out[~inside] = np.nan
    return out


def slab_spectral_expected(sigma_table, eta=1.0):
    """XYZ of the slab's transmitted white (eta = 1: straight through, no
    reflection) relative to the unobstructed environment, by 1 nm quadrature
    of the CIE 1931 functions against the environment's spectrum:
    sum S(l) exp(-sigma(l) d) cmf(l) / sum S(l) cmf(l), per X, Y, Z.  Normal
    incidence (d = slab thickness)."""
    assert eta == 1.0
    lam = np.arange(360, 831, dtype=np.float32)
    cmf = np.zeros((len(lam), 3), np.float32)
    for i, l in enumerate(lam):
        A.lib().ag_cie_xyz_bar(float(l), ag._ptr(cmf[i]))
    s = white_spectrum().astype(np.float64)
    tr = np.exp(-np.asarray(sigma_table, np.float64) * SLAB[2])
    return (s[:, None] * tr[:, None] * cmf).sum(0) / (s[:, None] * cmf).sum(0)


def flat_furnace_desc(mat, radiance):
    """A furnace whose environment has the same radiance at every wavelength
    (a flat spectrum, not an illuminant): every spectral sample of a path
    that escapes after a bounce of factor f carries exactly f * radiance."""
    desc = sphere_desc(mat)
    for name in ("top", "horizon", "bottom"):
        getattr(desc.d.env, name)[:] = [float(radiance)] * A.SPEC_N
    desc.d.env.power = 1.0
    return desc


def render_oracle(desc, cam, w, h, spp, seed=1, ps=None, strategy=0, rows=None):
    """Colour (h, w, 3) of an oracle render; rows outside `rows` stay zero."""
    import oracle as O
    osc = O.OracleScene(desc)
    osc.set_strategy(strategy)
    sums = np.zeros((3, h, w, 3), np.float32)
    osc.render_pass(cam, w, h, sums, 0, ps or ag.PathSettings(), spp, seed, rows=rows)
    osc.close()
    return sums[0] / spp


def block_ratio(img, bg_rows=4, c0=28, c1=36):
    """Mean of the centre block over the mean of the top-left background block."""
    bg = img[:bg_rows, :bg_rows].reshape(-1, 3).mean(0)
    return img[c0:c1, c0:c1].reshape(-1, 3).mean(0) / bg


def interior(expected, px=2):
    from scipy import ndimage
    return ndimage.binary_erosion(~np.isnan(expected), iterations=px)


HDRI_SPOTS = dict(width=128, height=64, seed=4, sigma=1.0)
