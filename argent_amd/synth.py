"""Seeded procedural meshes, rays and environment maps for the workloads.

Surfaces of revolution follow argent-mesh synth.rs:147-172 (torus_profile:
vertex (rad cos u, rad sin u, z), two triangles per quad with the same
winding), and binary STL follows synth.rs:283-302 / io/stl.rs:34-62.  The
workload sizes:

  torus ring   R 9.0, r 1.2 mm, 512 x 128 quads (131,072 triangles), radial
               value-noise displacement +/-0.05 mm
  signet       torus 1024 x 256 (524,288 triangles) joined to an ellipsoidal
               head plate with 6-octave fBm relief (0.3 mm), ~1.2 M triangles
  ray batch    origins uniform on a 30 mm sphere, directions toward the
               centre jittered up to 20 degrees
"""
import struct

import numpy as np

TAU = np.float32(2.0 * np.pi)


def torus_profile(nu, nv, profile):
    """synth.rs:153-172.  profile(v) -> (radius, z) on float32 arrays."""
    i = np.arange(nu, dtype=np.float32)
    j = np.arange(nv, dtype=np.float32)
    u = TAU * i / np.float32(nu)
    v = TAU * j / np.float32(nv)
    rad, z = profile(u[:, None], v[None, :])
    rad = np.broadcast_to(rad, (nu, nv)).astype(np.float32)
    z = np.broadcast_to(z, (nu, nv)).astype(np.float32)
    cu = np.cos(u)[:, None].astype(np.float32)
    su = np.sin(u)[:, None].astype(np.float32)
    pos = np.stack([rad * cu, rad * su, z], axis=-1).reshape(-1, 3).astype(np.float32)
    ii, jj = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    ii, jj = ii.reshape(-1), jj.reshape(-1)

    def idx(a, b):
        return ((a % nu) * nv + (b % nv)).astype(np.uint32)

    t0 = np.stack([idx(ii, jj), idx(ii + 1, jj), idx(ii + 1, jj + 1)], axis=-1)
    t1 = np.stack([idx(ii, jj), idx(ii + 1, jj + 1), idx(ii, jj + 1)], axis=-1)
    tris = np.stack([t0, t1], axis=1).reshape(-1, 3)
    return pos, tris


def torus(big_r, r, nu, nv):
    """synth.rs:147-149: a plain torus of revolution."""
    def prof(u, v):
        vv = np.broadcast_to(v, (u.shape[0], v.shape[1]))
        return np.float32(big_r) + np.float32(r) * np.cos(vv), np.float32(r) * np.sin(vv)
    return torus_profile(nu, nv, prof)


def cuboid(size):
    """synth.rs:9-35: an axis-aligned box centred on the origin, 8 vertices and
    12 outward-wound triangles."""
    hx, hy, hz = (np.float32(s) * np.float32(0.5) for s in size)
    pos = np.array([[-hx, -hy, -hz], [hx, -hy, -hz], [hx, hy, -hz], [-hx, hy, -hz],
                    [-hx, -hy, hz], [hx, -hy, hz], [hx, hy, hz], [-hx, hy, hz]], np.float32)
    tris = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                     [3, 7, 6], [3, 6, 2], [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], np.uint32)
    return pos, tris


def icosphere(r, subdiv):
    """synth.rs:81-144: icosahedron, `subdiv` midpoint subdivisions (shared
    midpoints), vertices pushed to radius r (20 * 4^subdiv triangles)."""
    t = np.float32((1.0 + np.sqrt(np.float32(5.0))) / 2.0)
    p = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    p = [np.array(q, np.float32) for q in p]
    tris = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
            [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
            [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    for _ in range(subdiv):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p.append(((p[a] + p[b]) * np.float32(0.5)).astype(np.float32))
                mid[key] = len(p) - 1
            return mid[key]

        nxt = []
        for a, b, c in tris:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nxt += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        tris = nxt
    P = np.array(p, np.float32)
    ln = np.sqrt(P[:, 0] * P[:, 0] + P[:, 1] * P[:, 1] + P[:, 2] * P[:, 2]).astype(np.float32)
    P = (P / ln[:, None] * np.float32(r)).astype(np.float32)
    return P, np.array(tris, np.uint32)


def subdivided_cuboid(size, n):
    """synth.rs:39-78: each face an n x n grid of quads, outward wound, shared
    corner and edge vertices merged (they coincide bit for bit)."""
    n = max(int(n), 1)
    h = [np.float32(s) * np.float32(0.5) for s in size]
    size = [np.float32(s) for s in size]
    pos, tris = [], []
    for axis in range(3):
        for sign in (-1.0, 1.0):
            a, b = (axis + 1) % 3, (axis + 2) % 3
            base = len(pos)
            for i in range(n + 1):
                for j in range(n + 1):
                    q = [np.float32(0)] * 3
                    q[axis] = np.float32(sign) * h[axis]
                    q[a] = np.float32(-h[a] + np.float32(size[a] * np.float32(i) / np.float32(n)))
                    q[b] = np.float32(-h[b] + np.float32(size[b] * np.float32(j) / np.float32(n)))
                    pos.append(q)
            idx = lambda i, j: base + i * (n + 1) + j
            for i in range(n):
                for j in range(n):
                    if sign > 0:
                        tris += [[idx(i, j), idx(i + 1, j), idx(i + 1, j + 1)], [idx(i, j), idx(i + 1, j + 1), idx(i, j + 1)]]
                    else:
                        tris += [[idx(i, j), idx(i + 1, j + 1), idx(i + 1, j)], [idx(i, j), idx(i, j + 1), idx(i + 1, j + 1)]]
    P = np.array(pos, np.float32)
    key = (P + np.float32(0.0)).view(np.uint32)
    _, first, inv = np.unique(key.view([("", np.uint32)] * 3).ravel(), return_index=True, return_inverse=True)
    order = np.argsort(first)
    new_id = np.empty(len(first), np.int64)
    new_id[order] = np.arange(len(first))
    return P[np.sort(first)], new_id[inv.ravel()][np.array(tris)].astype(np.uint32)


def unweld(pos, tris):
    """synth.rs:257-268: every triangle gets its own three vertices."""
    p = np.asarray(pos, np.float32)[np.asarray(tris).reshape(-1)]
    return p, np.arange(len(p), dtype=np.uint32).reshape(-1, 3)


def signed_volume(pos, tris):
    """Enclosed volume by the divergence theorem (f64)."""
    p = np.asarray(pos, np.float64)
    a, b, c = p[tris[:, 0]], p[tris[:, 1]], p[tris[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)


def vertex_normals(pos, tris):
    """Area-weighted vertex normals (unit, float32)."""
    p = pos.astype(np.float64)
    a, b, c = p[tris[:, 0]], p[tris[:, 1]], p[tris[:, 2]]
    fn = np.cross(b - a, c - a)
    n = np.zeros_like(p)
    for k in range(3):
        np.add.at(n, tris[:, k], fn)
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    ln[ln == 0] = 1.0
    return (n / ln).astype(np.float32)


def _periodic_value_noise(u, v, nu_cells, nv_cells, rng):
    """Smooth periodic value noise in [0,1) over the (u, v) torus."""
    lat = rng.random((nu_cells, nv_cells))
    x = (u / (2 * np.pi)) * nu_cells
    y = (v / (2 * np.pi)) * nv_cells
    x0 = np.floor(x).astype(np.int64)
    y0 = np.floor(y).astype(np.int64)
    fx = x - x0
    fy = y - y0
    sx = fx * fx * (3 - 2 * fx)
    sy = fy * fy * (3 - 2 * fy)

    def L(a, b):
        return lat[a % nu_cells, b % nv_cells]

    top = L(x0, y0) * (1 - sx) + L(x0 + 1, y0) * sx
    bot = L(x0, y0 + 1) * (1 - sx) + L(x0 + 1, y0 + 1) * sx
    return top * (1 - sy) + bot * sy


def torus_ring(seed=1, big_r=9.0, r=1.2, nu=512, nv=128, amp=0.05):
    """The 131,072-triangle ring: tube radius r + amp * (2 noise - 1)."""
    rng = np.random.default_rng(seed)

    def prof(u, v):
        uu = np.broadcast_to(u, (u.shape[0], v.shape[1])).astype(np.float64)
        vv = np.broadcast_to(v, (u.shape[0], v.shape[1])).astype(np.float64)
        n = _periodic_value_noise(uu, vv, 64, 16, rng)
        rr = (r + amp * (2.0 * n - 1.0)).astype(np.float32)
        return (np.float32(big_r) + rr * np.cos(v)).astype(np.float32), (rr * np.sin(v)).astype(np.float32)

    pos, tris = torus_profile(nu, nv, prof)
    return pos, tris, vertex_normals(pos, tris)


def _fbm(x, y, z, octaves, rng):
    """Seeded 3D value-noise fBm in [-1, 1]."""
    size = 64
    lat = rng.random((size, size, size)) * 2.0 - 1.0
    total = np.zeros_like(x)
    amp, freq, norm = 1.0, 1.0, 0.0
    for _ in range(octaves):
        px, py, pz = x * freq, y * freq, z * freq
        x0, y0, z0 = np.floor(px).astype(np.int64), np.floor(py).astype(np.int64), np.floor(pz).astype(np.int64)
        fx, fy, fz = px - x0, py - y0, pz - z0
        sx, sy, sz = fx * fx * (3 - 2 * fx), fy * fy * (3 - 2 * fy), fz * fz * (3 - 2 * fz)
        acc = 0.0
        for dx in (0, 1):
            wx = sx if dx else 1 - sx
            for dy in (0, 1):
                wy = sy if dy else 1 - sy
                for dz in (0, 1):
                    wz = sz if dz else 1 - sz
                    acc = acc + wx * wy * wz * lat[(x0 + dx) % size, (y0 + dy) % size, (z0 + dz) % size]
        total += amp * acc
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def signet(seed=2, band_nu=1024, band_nv=256, plate_nt=412, plate_np=820, relief=0.3):
    """Signet ring: a 1024 x 256 torus band (R 9.0, r 1.3, around Z) joined to
    an ellipsoidal head plate (semi-axes 7.0 x 2.2 x 5.5 mm, centred on +Y at
    the band's outer edge) with 6-octave fBm relief of `relief` mm along the
    normal.  ~1.2 M triangles, indexed, with vertex normals."""
    rng = np.random.default_rng(seed)
    pos_b, tri_b = torus_profile(band_nu, band_nv,
                                 lambda u, v: (np.float32(9.0) + np.float32(1.3) * np.cos(v),
                                               np.float32(1.3) * np.sin(v)))
    # ellipsoid, poles along Y
    th = (np.arange(plate_nt + 1, dtype=np.float64) / plate_nt) * np.pi
    ph = (np.arange(plate_np, dtype=np.float64) / plate_np) * 2.0 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    ax, ay, az = 7.0, 2.2, 5.5
    nx, ny, nz = np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)
    x, y, z = ax * nx, ay * ny, az * nz
    g = np.stack([x / ax ** 2, y / ay ** 2, z / az ** 2], axis=-1)
    g /= np.linalg.norm(g, axis=-1, keepdims=True)
    disp = relief * _fbm(x * 0.35 + 10.0, y * 0.35 + 10.0, z * 0.35 + 10.0, 6, rng)
    cy = 9.0 + 1.3 + 1.2
    px, py, pz = x + g[..., 0] * disp, y + g[..., 1] * disp + cy, z + g[..., 2] * disp
    pos_p = np.stack([px, py, pz], axis=-1).reshape(-1, 3).astype(np.float32)
    ii, jj = np.meshgrid(np.arange(plate_nt), np.arange(plate_np), indexing="ij")
    ii, jj = ii.reshape(-1), jj.reshape(-1)

    def idx(a, b):
        return (a * plate_np + (b % plate_np)).astype(np.uint32)

    t0 = np.stack([idx(ii, jj), idx(ii + 1, jj), idx(ii + 1, jj + 1)], axis=-1)
    t1 = np.stack([idx(ii, jj), idx(ii + 1, jj + 1), idx(ii, jj + 1)], axis=-1)
    tri_p = np.stack([t0, t1], axis=1).reshape(-1, 3)
    # drop the degenerate triangles at the poles
    a, b, c = pos_p[tri_p[:, 0]], pos_p[tri_p[:, 1]], pos_p[tri_p[:, 2]]
    keep = np.linalg.norm(np.cross(b - a, c - a), axis=1) > 0
    tri_p = tri_p[keep] + np.uint32(len(pos_b))
    pos = np.concatenate([pos_b, pos_p]).astype(np.float32)
    tris = np.concatenate([tri_b, tri_p]).astype(np.uint32)
    return pos, tris, vertex_normals(pos, tris)


def _clip(poly, n, c):
    """Sutherland-Hodgman: keep the part of a convex polygon with n.x <= c."""
    out = []
    m = len(poly)
    for i in range(m):
        a, b = poly[i], poly[(i + 1) % m]
        da, db = np.dot(n, a) - c, np.dot(n, b) - c
        if da <= 0.0:
            out.append(a)
        if (da < 0.0) != (db < 0.0) and da != db:
            out.append(a + (b - a) * (da / (da - db)))
    return out


def convex_hull_of_planes(normals, offsets):
    """The polytope {x : n_i . x <= c_i} as a triangle mesh (one fan per face,
    faces outward).  Every plane must bound the solid."""
    normals = np.asarray(normals, np.float64)
    offsets = np.asarray(offsets, np.float64)
    big = 4.0 * float(np.max(np.abs(offsets))) + 1.0
    verts, tris = [], []
    for i, (n, c) in enumerate(zip(normals, offsets)):
        n = n / np.linalg.norm(n)
        a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
        t1 = np.cross(n, a)
        t1 /= np.linalg.norm(t1)
        t2 = np.cross(n, t1)
        o = n * c
        poly = [o + big * (sx * t1 + sy * t2) for sx, sy in ((-1, -1), (1, -1), (1, 1), (-1, 1))]
        for j, (m, d) in enumerate(zip(normals, offsets)):
            if j != i and poly:
                poly = _clip(poly, m / np.linalg.norm(m), d / np.linalg.norm(m))
        if len(poly) < 3:
            continue
        base = len(verts)
        verts.extend(poly)
        for k in range(1, len(poly) - 1):
            tri = [base, base + k, base + k + 1]
            v0, v1, v2 = (poly[0], poly[k], poly[k + 1])
            if np.dot(np.cross(v1 - v0, v2 - v0), n) < 0.0:
                tri = [base, base + k + 1, base + k]
            tris.append(tri)
    return np.asarray(verts, np.float32), np.asarray(tris, np.uint32)


def brilliant(diameter=1.0, seed=0, jitter_deg=0.4):
    """Round brilliant of girdle diameter `diameter` (mm), table up (+Y),
    girdle centred on the origin: the convex hull of 57 planes -- table,
    8 stars, 8 bezels, 16 upper and 16 lower girdle facets, 8 pavilion
    mains -- with table 57 %, crown 34.5 deg, pavilion 40.75 deg, girdle
    2 %, each facet angle jittered by up to `jitter_deg` (seeded)."""
    rng = np.random.default_rng(seed)
    r, g = 0.5, 0.01
    crown, pav = np.deg2rad(34.5), np.deg2rad(40.75)
    table = 0.285
    h_table = g + (r - table) * np.tan(crown)
    normals, offsets = [], []

    def add(az_deg, tilt, through, up=True):
        j = np.deg2rad(jitter_deg) * (2.0 * rng.random() - 1.0)
        tilt = tilt + j
        az = np.deg2rad(az_deg)
        n = np.array([np.sin(tilt) * np.cos(az), np.cos(tilt) if up else -np.cos(tilt), np.sin(tilt) * np.sin(az)])
        normals.append(n)
        offsets.append(float(np.dot(n, through)))

    def rim(az_deg, radius, y):
        az = np.deg2rad(az_deg)
        return np.array([radius * np.cos(az), y, radius * np.sin(az)])

    normals.append(np.array([0.0, 1.0, 0.0]))
    offsets.append(h_table)
    for k in range(8):
        a = 45.0 * k
        add(a, crown, rim(a, r, g))                                        # bezel
        add(a + 22.5, np.deg2rad(21.5), rim(a + 22.5, table, h_table))     # star
        for s in (-11.25, 11.25):
            add(a + 22.5 + s, np.deg2rad(41.0), rim(a + 22.5 + s, r, g))   # upper girdle
            add(a + 22.5 + s, np.deg2rad(42.0), rim(a + 22.5 + s, r, -g), up=False)  # lower girdle
        add(a, pav, rim(a, r, -g), up=False)                               # pavilion main
    pos, tris = convex_hull_of_planes(np.asarray(normals) , np.asarray(offsets))
    return (pos * np.float32(diameter)).astype(np.float32), tris


def signet_stones(seed=2, n_stones=12, d_range=(2.0, 6.0), stone_seed=12, signet_kw=None):
    """The signet with `n_stones` brilliants set in its head plate: girdle
    diameters drawn from U[d_range] (largest placed first; a stone that finds
    no free spot after 400 tries shrinks by 10 %), tables level with the
    plate's top surface, no two stones overlapping.  Returns positions,
    triangles, vertex normals (stone normals flat per facet) and per-triangle
    material (0 metal, 1 + stone index % 2: two gem slots)."""
    pos, tris, nrm = signet(seed, **(signet_kw or {}))
    rng = np.random.default_rng(stone_seed)
    ds = np.sort(rng.uniform(d_range[0], d_range[1], n_stones))[::-1]
    ax, az, cy, ay = 7.0, 5.5, 9.0 + 1.3 + 1.2, 2.2
    placed = []
    P, T, N, M = [pos], [tris], [nrm], [np.zeros(len(tris), np.uint16)]
    base = len(pos)
    for i, d in enumerate(ds):
        for attempt in range(100000):
            if attempt and attempt % 400 == 0:
                d *= 0.9
            x, z = (2.0 * rng.random() - 1.0) * ax, (2.0 * rng.random() - 1.0) * az
            q = (x / ax) ** 2 + (z / az) ** 2
            rad = d / 2.0
            # keep the girdle inside the flatter part of the plate
            if ((abs(x) + rad) / ax) ** 2 + ((abs(z) + rad) / az) ** 2 > 0.8:
                continue
            if any((x - px) ** 2 + (z - pz) ** 2 < ((rad + pr) * 1.05) ** 2 for px, pz, pr in placed):
                continue
            break
        placed.append((x, z, rad))
        y = cy + ay * np.sqrt(max(1.0 - q, 0.0)) + 0.05
        sp, st = brilliant(d, seed=stone_seed * 1000 + i)
        sp = sp + np.array([x, y - 0.16 * d, z], np.float32)
        P.append(sp.astype(np.float32))
        T.append(st + np.uint32(base))
        a, b, c = sp[st[:, 0]], sp[st[:, 1]], sp[st[:, 2]]
        fn = np.cross(b - a, c - a)
        fn /= np.linalg.norm(fn, axis=1, keepdims=True)
        vn = np.zeros_like(sp)
        for k in range(3):
            vn[st[:, k]] = fn
        N.append(vn.astype(np.float32))
        M.append(np.full(len(st), 1 + i % 2, np.uint16))
        base += len(sp)
    return (np.concatenate(P).astype(np.float32), np.concatenate(T).astype(np.uint32),
            np.concatenate(N).astype(np.float32), np.concatenate(M).astype(np.uint16), [2 * p[2] for p in placed])


def sphere_rays(n, center, radius=30.0, jitter_deg=20.0, seed=3):
    """Origins uniform on a sphere; directions toward `center`, rotated by a
    uniform angle in [0, jitter] about a random perpendicular axis."""
    rng = np.random.default_rng(seed)
    z = rng.random(n) * 2.0 - 1.0
    phi = rng.random(n) * 2.0 * np.pi
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    dirs = np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)
    o = np.asarray(center, np.float64)[None, :] + radius * dirs
    d = -dirs
    helper = np.where(np.abs(d[:, 0:1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
    a1 = np.cross(d, helper)
    a1 /= np.linalg.norm(a1, axis=1, keepdims=True)
    a2 = np.cross(d, a1)
    ang = rng.random(n) * np.radians(jitter_deg)
    rot = rng.random(n) * 2.0 * np.pi
    axis = a1 * np.cos(rot)[:, None] + a2 * np.sin(rot)[:, None]
    # rotate d about axis by ang (axis is perpendicular to d)
    d = d * np.cos(ang)[:, None] + np.cross(axis, d) * np.sin(ang)[:, None]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(np.float32), d.astype(np.float32)


def hdri(width=2048, height=1024, seed=4, sigma=2.0):
    """Equirectangular linear-RGB map: log-normal luminance (sigma) smoothed
    over 16-texel cells plus three Gaussian hot spots."""
    rng = np.random.default_rng(seed)
    cw, ch = max(1, width // 16), max(1, height // 16)
    cells = np.exp(rng.normal(0.0, sigma, (ch, cw)))
    lum = np.kron(cells, np.ones((height // ch + 1, width // cw + 1)))[:height, :width]
    tint = 0.85 + 0.3 * rng.random((ch, cw, 3))
    tint = np.kron(tint, np.ones((height // ch + 1, width // cw + 1, 1)))[:height, :width]
    img = lum[..., None] * tint * 0.05
    yy, xx = np.meshgrid(np.arange(height) + 0.5, np.arange(width) + 0.5, indexing="ij")
    for _ in range(3):
        cx, cy = rng.random() * width, (0.15 + 0.35 * rng.random()) * height
        sx = 0.01 * width + 0.02 * width * rng.random()
        dx = np.minimum(np.abs(xx - cx), width - np.abs(xx - cx))
        g = np.exp(-0.5 * ((dx / sx) ** 2 + ((yy - cy) / sx) ** 2))
        img += g[..., None] * (50.0 + 150.0 * rng.random()) * np.array([1.0, 0.95, 0.85])
    return img.astype(np.float32)


def to_binary_stl(pos, tris):
    """synth.rs:283-302: 80-byte header, count, then per triangle the unit
    normal, three vertices and a zero attribute word."""
    p = pos.astype(np.float32)
    a, b, c = p[tris[:, 0]], p[tris[:, 1]], p[tris[:, 2]]
    n = np.cross(b - a, c - a)
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    ln[ln == 0] = 1.0
    n = (n / ln).astype(np.float32)
    rec = np.zeros(len(tris), dtype=[("n", "<f4", 3), ("v", "<f4", (3, 3)), ("attr", "<u2")])
    rec["n"] = n
    rec["v"] = np.stack([a, b, c], axis=1)
    header = b"solid argent-synth   ".ljust(80, b" ")
    return header + struct.pack("<I", len(tris)) + rec.tobytes()


def parse_binary_stl(data):
    """io/stl.rs:34-62: returns an unwelded triangle soup (positions, triangles)."""
    if len(data) < 84:
        raise ValueError("binary STL shorter than its header")
    (n,) = struct.unpack_from("<I", data, 80)
    if len(data) < 84 + 50 * n:
        raise ValueError("binary STL truncated")
    rec = np.frombuffer(data, dtype=[("n", "<f4", 3), ("v", "<f4", (3, 3)), ("attr", "<u2")], count=n, offset=84)
    pos = rec["v"].reshape(-1, 3).astype(np.float32)
    tris = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    return pos, tris
