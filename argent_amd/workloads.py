"""The benchmark workloads as scenes: Prepared::scene + Prepared::shot
(crates/argent-core/src/shot.rs:206-295) for the synthetic pieces.

    torus_still   131,072-triangle torus ring, 14k gold (roughness 0.07,
                  patina 0.8), noir studio, 512 x 512, 64 spp, depth 8, f/16
    signet_still  ~1.2 M-triangle signet, sterling, white studio,
                  1080 x 1920, 256 spp, depth 12, no depth of field

The pieces are generated already posed (band around Z, head toward +Y), as
`Prepared::build` with a given pose does (shot.rs:104-147); its patina step
(shot.rs:133-143) is Piece.prepare_patina, which computes the cavity mask on
the device and stores it on the piece, so the patina layer (strength 0.8)
shows in the recesses.  Lens: 100 mm, f/16 outside the white studio, no
depth of field in it (jobs.rs:403-406).
"""
import numpy as np

from . import Material, PathSettings, SceneDesc, VIEWS, frame, synth
from . import _abi as A


class Piece:
    """Prepared (shot.rs:73-147) for a posed mesh: radius, floor height, hull."""

    def __init__(self, positions, triangles, normals=None, tri_material=None, cavity=None):
        self.positions = np.asarray(positions, np.float32)
        self.triangles = np.asarray(triangles, np.uint32)
        self.normals = normals
        self.tri_material = tri_material
        self.cavity = cavity
        # per column: numpy reduces an (n, 3) array along axis 0 an order of magnitude slower
        lo = np.array([self.positions[:, k].min() for k in range(3)], np.float32)
        hi = np.array([self.positions[:, k].max() for k in range(3)], np.float32)
        self.radius = float(max(np.linalg.norm((hi - lo) * 0.5), 1e-3))
        self.floor_y = float(lo[1])
        step = max(len(self.positions) // 4000, 1)
        self.hull = self.positions[::step].copy()

    @classmethod
    def from_frame(cls, lo, hi, hull):
        """A Piece for a mesh that stays on the device (Mesh.frame_sample):
        the same radius, floor height and hull as Piece(positions, ...) on its
        positions; the scene descriptor it builds carries the hull only, for
        Scene.from_mesh, which takes the geometry from the device mesh."""
        p = cls.__new__(cls)
        lo, hi = np.asarray(lo, np.float32), np.asarray(hi, np.float32)
        p.positions = np.asarray(hull, np.float32)
        p.triangles = np.zeros((1, 3), np.uint32)
        p.normals = p.tri_material = p.cavity = None
        p.radius = float(max(np.linalg.norm((hi - lo) * 0.5), 1e-3))
        p.floor_y = float(lo[1])
        p.hull = p.positions.copy()
        return p

    def patina_settings(self):
        """PatinaSettings sized by the piece (shot.rs:133-141), in f32 as the
        reference computes them: a dict for the oracle's patina_mask."""
        lo = np.float32([self.positions[:, k].min() for k in range(3)])
        hi = np.float32([self.positions[:, k].max() for k in range(3)])
        h = (hi - lo) * np.float32(0.5)
        r = max(np.sqrt(np.float32(h[0] * h[0] + h[1] * h[1] + h[2] * h[2])), np.float32(1e-3))
        return dict(ao_samples=48, ao_distance=float(np.clip(r * np.float32(0.09), np.float32(0.4), np.float32(2.5))),
                    curvature_radius=float(np.clip(r * np.float32(0.03), np.float32(0.15), np.float32(1.0))),
                    concavity_weight=0.35, smooth_iterations=3)

    def prepare_patina(self, scene):
        """Prepared::build's patina step (shot.rs:133-143) on the device: the
        mask with settings sized by the piece, stored as the piece's cavity
        in `scene` (and in scenes created over it later); a host copy stays on
        the piece for scenes described from it later and for the oracle."""
        self.cavity = scene.patina_mask(scene.patina_settings(), store=True)
        return self.cavity

    def scene(self, metal="sterling", roughness=0.07, patina=0.8, studio="noir", gems=()):
        """Prepared::scene (shot.rs:206-245): slot 0 metal, then gems, then
        the studio's floor material."""
        d = SceneDesc(self.positions, self.triangles, self.normals, self.tri_material, self.cavity)
        d.add_material(Material.metal(metal, roughness, patina))
        for g in gems:
            d.add_material(Material.gem(g))
        d.apply_studio(studio, self.radius, self.floor_y)
        return d

    def shot(self, view="hero", focal_mm=100.0, f_stop=16.0, aspect=1.0, intersect=None):
        """Prepared::shot (shot.rs:263-273): the piece turns by -azimuth, the
        camera frames it from the rig's front; autofocus through `intersect`
        (a callable (o, d) -> t, u, v, prim) when the lens has an aperture."""
        az, el, fill = VIEWS[view] if isinstance(view, str) else view
        turn = -az
        pts = _rot_y(self.hull, turn)
        cam = frame(pts, (0.0, el, fill), focal_mm, f_stop, aspect)
        if cam.lens_radius > 0.0 and intersect is not None:
            autofocus(cam, turn, intersect)
        return cam, turn


def _rot_y(v, deg):
    r = np.float32(deg) * np.float32(np.pi / 180.0)
    s, c = np.float32(np.sin(r)), np.float32(np.cos(r))
    v = np.asarray(v, np.float32)
    return np.stack([c * v[..., 0] + s * v[..., 2], v[..., 1], -s * v[..., 0] + c * v[..., 2]], axis=-1)


def autofocus(cam, turn_deg, intersect):
    """shot.rs:278-295: eight probes near the frame centre, median distance."""
    probes = [(0.0, 0.0), (0.0, 0.15), (0.0, 0.3), (0.12, 0.22), (-0.12, 0.22), (0.0, 0.42), (0.2, 0.1),
              (-0.2, 0.1)]
    pos = np.array(cam.pos[:], np.float32)
    fwd, right, up = (np.array(getattr(cam, k)[:], np.float32) for k in ("forward", "right", "up"))
    os_, ds = [], []
    for sx, sy in probes:
        d = fwd + right * np.float32(sx * cam.half_w) + up * np.float32(sy * cam.half_h)
        d = d / np.linalg.norm(d)
        os_.append(pos)
        ds.append(d)
    o = _rot_y(np.array(os_), -turn_deg)
    d = _rot_y(np.array(ds), -turn_deg)
    t, _, _, prim = intersect(o, d)
    f = _rot_y(fwd, -turn_deg)
    hits = sorted(float(t[i] * (d[i] @ f)) for i in range(len(t)) if prim[i] != 0xFFFFFFFF)
    if hits:
        cam.focus_dist = hits[len(hits) // 2]


def torus_still():
    pos, tris, nrm = synth.torus_ring(seed=1)
    p = Piece(pos, tris, nrm)
    return dict(piece=p, metal="gold14k", studio="noir", width=512, height=512, spp=64, depth=8, f_stop=16.0,
                seed=7)


def signet_still():
    pos, tris, nrm = synth.signet(seed=2)
    p = Piece(pos, tris, nrm)
    return dict(piece=p, metal="sterling", studio="white", width=1080, height=1920, spp=256, depth=12, f_stop=0.0,
                seed=7)


def signet_gems_still(hdri_size=(2048, 1024)):
    """The signet with 12 brilliants (diamond and sapphire) in its plate,
    lit by a 2048 x 1024 HDRI (log-normal luminance sigma 2 plus three hot
    spots) together with the white rig's softboxes (light pick 0.5 between
    the map and the panels), 1080 x 1920, 512 spp, depth 24."""
    pos, tris, nrm, mat, diam = synth.signet_stones(seed=2)
    p = Piece(pos, tris, nrm, tri_material=mat)
    w, h = hdri_size
    return dict(piece=p, metal="sterling", studio="white", gems=("diamond", "sapphire"), width=1080, height=1920,
                spp=512, depth=24, f_stop=0.0, seed=7, hdri=synth.hdri(w, h, seed=4), stone_diameters=diam)


def path_settings(depth):
    return PathSettings(max_depth=depth, clamp_indirect=24.0, rr_depth=6)
