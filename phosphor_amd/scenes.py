"""Procedural scenes in the reference's SceneData form.

The reference imports glTF through assimp (src/parsers/scene.cpp); the
workloads here are procedural, so this module produces the same arrays the
parser would: triangles split to SoA (geometry.cpp:10-28), materials with the
parser defaults (material.cpp:9-30), one LIGHT_TEXTURED light per emissive
mesh with power = π·E·area (scene.cpp:115-130), the camera basis of
camera.cpp:15-56, and optionally an equirectangular RGB float envmap.

Meshes without UVs get zero UVs and the tangent (1, 0, 0), as the parser's
no-UV fallback produces (scene.cpp:80-91).
"""
from __future__ import annotations

import functools
import math
from dataclasses import dataclass, field

import numpy as np

from .abi import (CAMERA_DTYPE, LIGHT_DIRECTIONAL, LIGHT_DTYPE, LIGHT_TEXTURED, MATERIAL_DTYPE, NO_TEXTURE,
                  TEX_META_DTYPE, UV_BLOCK_DTYPE, SceneDesc, ptr)

f32 = np.float32
PI = f32(3.141592654)
EPS = 1e-5


def as_float(u: int) -> np.float32:
    return np.array([u], dtype=np.uint32).view(np.float32)[0]


def _norm(v):
    v = np.asarray(v, dtype=f32)
    return (v / f32(np.sqrt(np.dot(v, v)))).astype(f32)


def make_tbn(d: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Tangent frame of a unit direction (utils.h:44-53)."""
    d = np.asarray(d, f32)
    if abs(d[0]) > abs(d[1]):
        t = _norm(np.array([-d[2], 0.0, d[0]], f32))
    else:
        t = _norm(np.array([0.0, d[2], -d[1]], f32))
    return t, np.cross(d, t).astype(f32)


def make_camera(position, look_at, up, hfov_deg: float, aspect: float, defocus_angle: float = 0.0,
                focus_distance: float = 1.0) -> np.ndarray:
    """CameraParams of Camera::recalculate/focus/to_params (camera.cpp:15-56), in float32."""
    pos = np.asarray(position, f32)
    w = _norm(pos - np.asarray(look_at, f32))
    u = _norm(np.cross(np.asarray(up, f32), w).astype(f32))
    v = np.cross(w, u).astype(f32)
    half_w = f32(math.tan(math.radians(hfov_deg) * 0.5)) * f32(focus_distance)
    half_h = f32(half_w / f32(aspect))
    horizontal = (f32(2.0) * half_w * u).astype(f32)
    vertical = (f32(2.0) * half_h * v).astype(f32)
    lower_left = (pos - f32(0.5) * horizontal - f32(0.5) * vertical - w * f32(focus_distance)).astype(f32)
    r = f32(focus_distance) * f32(math.tan(math.radians(defocus_angle / 2.0)))
    cam = np.zeros(1, CAMERA_DTYPE)
    c = cam[0]
    c["position"][:3] = pos
    c["lower_left_corner"][:3] = lower_left
    c["horizontal"][:3] = horizontal
    c["vertical"][:3] = vertical
    c["defocus_disk_u"][:3] = u * r
    c["defocus_disk_v"][:3] = v * r
    c["defocus_angle"] = defocus_angle
    c["focus_distance"] = focus_distance
    return cam


def default_material(base_color=(1.0, 1.0, 1.0), emissive=(0.0, 0.0, 0.0), metallic=0.0, roughness=1.0,
                     transmission=0.0, ior=1.5) -> np.ndarray:
    """material.cpp:9-30 defaults with the given overrides."""
    m = np.zeros(1, MATERIAL_DTYPE)[0]
    m["base_color"] = (*base_color, 1.0)
    m["emissive"] = (*emissive, 0.0)
    m["att_color"] = (1.0, 1.0, 1.0, 0.0)
    m["metallic"] = metallic
    m["roughness"] = roughness
    m["transmission"] = transmission
    m["ior"] = ior
    m["att_dist"] = 1e9
    m["thickness"] = 0.0
    for k in ("diff_index", "emis_index", "norm_index", "occlusion_index", "metal_rough_index",
              "transmission_index"):
        m[k] = NO_TEXTURE
    return m


@dataclass
class Scene:
    """SceneData (include/host/scene_data.hpp:21-42) already split to SoA."""

    tri_verts: np.ndarray          # (n, 3, 4) f32
    tri_normals: np.ndarray        # (n, 3, 4) f32
    tri_tangents: np.ndarray       # (n, 3, 4) f32
    tri_uvs: np.ndarray            # (n, 8) f32: uv0, uv1, uv2, padding
    tri_mat_index: np.ndarray      # (n,) u32
    etri_verts: np.ndarray
    etri_normals: np.ndarray
    etri_uvs: np.ndarray
    etri_mat_index: np.ndarray
    materials: np.ndarray          # MATERIAL_DTYPE
    uv_transforms: np.ndarray      # UV_BLOCK_DTYPE
    lights: np.ndarray             # LIGHT_DTYPE
    camera: np.ndarray             # CAMERA_DTYPE (1,)
    envmap: np.ndarray | None = None   # (h, w, 3) f32
    tex_meta: np.ndarray | None = None
    tex_atlas: np.ndarray | None = None
    bvh_nodes: np.ndarray | None = None  # optional prebuilt (m, 8) f32
    name: str = "scene"
    meta: dict = field(default_factory=dict)

    @property
    def n_triangles(self) -> int:
        return int(self.tri_verts.shape[0])

    def desc(self) -> SceneDesc:
        """PhxSceneDesc pointing into this scene's arrays (keep the scene alive while it is used)."""
        d = SceneDesc()
        d.tri_verts = ptr(self.tri_verts)
        d.tri_normals = ptr(self.tri_normals)
        d.tri_tangents = ptr(self.tri_tangents)
        d.tri_uvs = ptr(self.tri_uvs)
        d.tri_mat_index = ptr(self.tri_mat_index)
        d.n_triangles = self.n_triangles
        d.etri_verts = ptr(self.etri_verts) if len(self.etri_verts) else None
        d.etri_normals = ptr(self.etri_normals) if len(self.etri_normals) else None
        d.etri_uvs = ptr(self.etri_uvs) if len(self.etri_uvs) else None
        d.etri_mat_index = ptr(self.etri_mat_index) if len(self.etri_mat_index) else None
        d.n_e_triangles = int(self.etri_verts.shape[0])
        d.materials = ptr(self.materials)
        d.uv_transforms = ptr(self.uv_transforms)
        d.n_materials = int(self.materials.shape[0])
        d.lights = ptr(self.lights) if len(self.lights) else None
        d.n_lights = int(self.lights.shape[0])
        if self.tex_meta is not None and len(self.tex_meta):
            d.tex_meta = ptr(self.tex_meta)
            d.n_textures = int(self.tex_meta.shape[0])
            d.tex_atlas = ptr(self.tex_atlas)
            d.tex_atlas_bytes = int(self.tex_atlas.nbytes)
        if self.envmap is not None:
            d.envmap = ptr(self.envmap)
            d.envmap_height, d.envmap_width = int(self.envmap.shape[0]), int(self.envmap.shape[1])
        d.camera = ptr(self.camera)
        if self.bvh_nodes is not None:
            d.bvh_nodes = ptr(self.bvh_nodes)
            d.n_bvh_nodes = int(self.bvh_nodes.shape[0])
        return d

    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        v = self.tri_verts[:, :, :3].reshape(-1, 3)
        return v.min(axis=0), v.max(axis=0)


class SceneBuilder:
    """Accumulates meshes the way parse_mesh does (scene.cpp:36-131)."""

    def __init__(self):
        self.verts: list[np.ndarray] = []
        self.normals: list[np.ndarray] = []
        self.mats: list[np.ndarray] = []
        self.mat_index: list[np.ndarray] = []
        self.ever: list[np.ndarray] = []
        self.enorm: list[np.ndarray] = []
        self.emat: list[np.ndarray] = []
        self.lights: list[np.ndarray] = []
        self.n_tris = 0
        self.n_etris = 0
        self.envmap = None

    def material(self, m: np.ndarray) -> int:
        self.mats.append(m)
        return len(self.mats) - 1

    def add_mesh(self, tris: np.ndarray, mat: int, normals: np.ndarray | None = None) -> None:
        """tris: (n, 3, 3) positions; normals: (n, 3, 3) vertex normals (face normals if None)."""
        tris = np.asarray(tris, f32).reshape(-1, 3, 3)
        n = tris.shape[0]
        if normals is None:
            fn = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]).astype(f32)
            fn = (fn / np.sqrt((fn * fn).sum(axis=1, keepdims=True))).astype(f32)
            normals = np.repeat(fn[:, None, :], 3, axis=1)
        normals = np.asarray(normals, f32).reshape(n, 3, 3)
        v4 = np.zeros((n, 3, 4), f32)
        v4[:, :, :3] = tris
        n4 = np.zeros((n, 3, 4), f32)
        n4[:, :, :3] = normals
        self.verts.append(v4)
        self.normals.append(n4)
        self.mat_index.append(np.full(n, mat, np.uint32))
        m = self.mats[mat]
        em = np.asarray(m["emissive"][:3], f32)
        if n > 0 and np.any(em > EPS):
            # scene.cpp:115-130: emissive mesh → LIGHT_TEXTURED with power = π·E·area
            total = f32(0.0)
            for t in tris:
                e1 = (t[1] - t[0]).astype(f32)
                e2 = (t[2] - t[0]).astype(f32)
                c = np.cross(e1, e2).astype(f32)
                total = f32(total + f32(0.5) * f32(np.sqrt(f32(c[0] * c[0] + c[1] * c[1]) + c[2] * c[2])))
            start = self.n_etris
            self.ever.append(v4.copy())
            self.enorm.append(n4.copy())
            self.emat.append(np.full(n, mat, np.uint32))
            self.n_etris += n
            light = np.zeros(1, LIGHT_DTYPE)[0]
            light["kind"] = LIGHT_TEXTURED
            light["power"][:3] = (PI * em).astype(f32) * total
            light["aux"] = (as_float(int(m["emis_index"])), as_float(start), as_float(n), total)
            self.lights.append(light)
        self.n_tris += n

    def add_quad(self, p0, p1, p2, p3, mat: int) -> None:
        q = np.asarray([p0, p1, p2, p3], f32)
        self.add_mesh(np.stack([q[[0, 1, 2]], q[[0, 2, 3]]]), mat)

    def add_box(self, lo, hi, mat: int) -> None:
        x0, y0, z0 = lo
        x1, y1, z1 = hi
        c = lambda x, y, z: (x, y, z)  # noqa: E731
        faces = [  # counter-clockwise seen from outside → outward face normals
            (c(x0, y0, z1), c(x1, y0, z1), c(x1, y1, z1), c(x0, y1, z1)),  # +z
            (c(x1, y0, z0), c(x0, y0, z0), c(x0, y1, z0), c(x1, y1, z0)),  # -z
            (c(x1, y0, z1), c(x1, y0, z0), c(x1, y1, z0), c(x1, y1, z1)),  # +x
            (c(x0, y0, z0), c(x0, y0, z1), c(x0, y1, z1), c(x0, y1, z0)),  # -x
            (c(x0, y1, z1), c(x1, y1, z1), c(x1, y1, z0), c(x0, y1, z0)),  # +y
            (c(x0, y0, z0), c(x1, y0, z0), c(x1, y0, z1), c(x0, y0, z1)),  # -y
        ]
        tris = []
        for q in faces:
            tris.append([q[0], q[1], q[2]])
            tris.append([q[0], q[2], q[3]])
        self.add_mesh(np.asarray(tris, f32), mat)

    def add_sphere(self, center, radius: float, mat: int, level: int = 5) -> None:
        v, f = icosphere(level)
        tris = (v[f] * f32(radius) + np.asarray(center, f32)).astype(f32)
        self.add_mesh(tris, mat, normals=v[f])

    def add_point_light(self, position, power) -> None:
        """light.cpp:4-11 make_point_light."""
        light = np.zeros(1, LIGHT_DTYPE)[0]
        light["kind"] = 0
        light["position"][:3] = position
        light["power"][:3] = power
        self.lights.append(light)

    def add_spot_light(self, position, power, direction, inner_rad: float, outer_rad: float) -> None:
        """light.cpp:13-22 make_spot_light."""
        light = np.zeros(1, LIGHT_DTYPE)[0]
        light["kind"] = 1
        light["position"][:3] = position
        light["power"][:3] = power
        light["direction"][:3] = direction
        light["aux"] = (inner_rad, outer_rad, 0.0, 0.0)
        self.lights.append(light)

    def add_directional_light(self, direction, power) -> None:
        """light.cpp:24-36 make_directional_light."""
        d = np.asarray(direction, f32)
        t, b = make_tbn(d)
        light = np.zeros(1, LIGHT_DTYPE)[0]
        light["kind"] = LIGHT_DIRECTIONAL
        light["direction"][:3] = d
        light["tangent"][:3] = t
        light["bitangent"][:3] = b
        light["power"][:3] = power
        self.lights.append(light)

    def build(self, camera: np.ndarray, name: str = "scene", envmap: np.ndarray | None = None, **meta) -> Scene:
        tv = np.ascontiguousarray(np.concatenate(self.verts))
        tn = np.ascontiguousarray(np.concatenate(self.normals))
        n = tv.shape[0]
        tt = np.zeros((n, 3, 4), f32)
        tt[:, :, 0] = 1.0
        tuv = np.zeros((n, 8), f32)
        tm = np.ascontiguousarray(np.concatenate(self.mat_index))
        if self.ever:
            ev = np.ascontiguousarray(np.concatenate(self.ever))
            en = np.ascontiguousarray(np.concatenate(self.enorm))
            em = np.ascontiguousarray(np.concatenate(self.emat))
        else:
            ev = np.zeros((0, 3, 4), f32)
            en = np.zeros((0, 3, 4), f32)
            em = np.zeros(0, np.uint32)
        euv = np.zeros((ev.shape[0], 8), f32)
        mats = np.ascontiguousarray(np.array(self.mats, dtype=MATERIAL_DTYPE))
        uvt = np.zeros(len(self.mats), UV_BLOCK_DTYPE)
        for k in UV_BLOCK_DTYPE.names:
            uvt[k]["uv_scale"] = 1.0
        lights = np.ascontiguousarray(np.array(self.lights, dtype=LIGHT_DTYPE)) if self.lights else \
            np.zeros(0, LIGHT_DTYPE)
        return Scene(tv, tn, tt, tuv, tm, ev, en, euv, em, mats, uvt, lights, camera, envmap=envmap,
                     tex_meta=np.zeros(0, TEX_META_DTYPE), tex_atlas=np.zeros(1, np.uint8), name=name, meta=meta)


@functools.lru_cache(maxsize=None)
def icosphere(level: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere: 20·4^level triangles (level 5 → 20480), vertices on the
    unit sphere.  Cached: callers must not modify the arrays."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache: dict[tuple[int, int], int] = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.asarray(verts, f32), np.asarray(faces, np.int64)


def grid_res_for_radius(extent: float, radius: float) -> int:
    """Grid resolution whose cell half-diagonal (the reference's initial radius rule,
    gather_flux.cl:51-56) is closest to `radius` for a cubic scene of side `extent`."""
    g = int(round(math.sqrt(3.0) * extent / (2.0 * radius)))
    return max(1, min(255, g))


WHITE = (0.75, 0.75, 0.75)
RED = (0.63, 0.06, 0.06)
GREEN = (0.14, 0.45, 0.14)


def _cornell_room(b: SceneBuilder, white: int, red: int, green: int, light_mat: int, boxes: bool = True) -> None:
    # 5 walls of the unit box (open at z = 1): floor, ceiling, back, left (red), right (green)
    b.add_quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0), white)   # floor, normal +y
    b.add_quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1), white)   # ceiling, normal -y
    b.add_quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), white)   # back wall, normal +z
    b.add_quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1), red)     # left wall, normal +x
    b.add_quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0), green)   # right wall, normal -x
    if boxes:
        b.add_box((0.18, 0.0, 0.18), (0.48, 0.6, 0.48), white)      # tall box
        b.add_box((0.52, 0.0, 0.50), (0.82, 0.3, 0.80), white)      # short box
    # 0.25 × 0.25 area light just below the ceiling, facing down
    y = 0.999
    b.add_quad((0.375, y, 0.375), (0.625, y, 0.375), (0.625, y, 0.625), (0.375, y, 0.625), light_mat)


def cornell_box(width: int = 512, height: int = 512) -> Scene:
    """Procedural Cornell box: 5 walls + 2 axis-aligned boxes (34 triangles) and a
    0.25-unit quad light of emission 17 (2 more triangles, also in the BVH)."""
    b = SceneBuilder()
    white = b.material(default_material(WHITE))
    red = b.material(default_material(RED))
    green = b.material(default_material(GREEN))
    light = b.material(default_material(WHITE, emissive=(17.0, 17.0, 17.0)))
    _cornell_room(b, white, red, green, light)
    cam = make_camera((0.5, 0.5, 2.2), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0), 45.0, width / height)
    return b.build(cam, name="cornell")


def caustic_scene(width: int = 1024, height: int = 1024, level: int = 5) -> Scene:
    """Cornell box plus two tessellated spheres (20480 triangles each at level 5):
    a GGX dielectric (transmission 1, IOR 1.5, roughness 0.05) resting on the short
    box and a GGX conductor (metallic 1, roughness 0.3) on the floor."""
    b = SceneBuilder()
    white = b.material(default_material(WHITE))
    red = b.material(default_material(RED))
    green = b.material(default_material(GREEN))
    light = b.material(default_material(WHITE, emissive=(17.0, 17.0, 17.0)))
    glass = b.material(default_material((1.0, 1.0, 1.0), transmission=1.0, ior=1.5, roughness=0.05))
    metal = b.material(default_material((0.95, 0.64, 0.54), metallic=1.0, roughness=0.3))
    _cornell_room(b, white, red, green, light)
    b.add_sphere((0.67, 0.3 + 0.14, 0.65), 0.14, glass, level)
    b.add_sphere((0.30, 0.14, 0.75), 0.14, metal, level)
    cam = make_camera((0.5, 0.5, 2.2), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0), 45.0, width / height)
    return b.build(cam, name="caustic")


def _cylinder(center, radius, height, seg, rings):
    """Open cylinder along +z (z-up scenes), seg·rings·2 triangles, outward vertex normals."""
    cx, cy, z0 = center
    ang = np.linspace(0.0, 2.0 * np.pi, seg + 1)
    zs = np.linspace(z0, z0 + height, rings + 1)
    tris, norms = [], []
    for r in range(rings):
        for s in range(seg):
            a0, a1 = ang[s], ang[s + 1]
            p = lambda a, z: (cx + radius * np.cos(a), cy + radius * np.sin(a), z)  # noqa: E731
            n = lambda a: (np.cos(a), np.sin(a), 0.0)  # noqa: E731
            q = [p(a0, zs[r]), p(a1, zs[r]), p(a1, zs[r + 1]), p(a0, zs[r + 1])]
            qn = [n(a0), n(a1), n(a1), n(a0)]
            tris += [[q[0], q[1], q[2]], [q[0], q[2], q[3]]]
            norms += [[qn[0], qn[1], qn[2]], [qn[0], qn[2], qn[3]]]
    return np.asarray(tris, f32), np.asarray(norms, f32)


def _arch(p0, p1, z_base, radius_tube, seg_arc, seg_tube):
    """Half torus from column p0 to column p1 (in the vertical plane through both)."""
    p0 = np.asarray(p0, np.float64)
    p1 = np.asarray(p1, np.float64)
    c = 0.5 * (p0 + p1)
    span = np.linalg.norm(p1 - p0) / 2.0
    ex = (p1 - p0) / (2.0 * span)
    ez = np.array([0.0, 0.0, 1.0])
    ey = np.cross(ez, ex)
    th = np.linspace(np.pi, 0.0, seg_arc + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, seg_tube + 1)

    def pt(t, p):
        ring = c + span * (np.cos(t) * ex) + (z_base + span * np.sin(t)) * ez
        radial = np.cos(t) * ex + np.sin(t) * ez
        n = np.cos(p) * radial + np.sin(p) * ey
        return ring + radius_tube * n, n

    tris, norms = [], []
    for i in range(seg_arc):
        for j in range(seg_tube):
            a, na = pt(th[i], ph[j])
            bb, nb = pt(th[i + 1], ph[j])
            cc, nc = pt(th[i + 1], ph[j + 1])
            d, nd = pt(th[i], ph[j + 1])
            tris += [[a, bb, cc], [a, cc, d]]
            norms += [[na, nb, nc], [na, nc, nd]]
    return np.asarray(tris, f32), np.asarray(norms, f32)


def sky_envmap(width: int = 2048, height: int = 1024, sun_dir=(0.35, 0.25, 0.9), sun_radiance: float = 5e4,
               sun_angle_deg: float = 0.53) -> np.ndarray:
    """Synthetic fp32 HDR sky for envmap.h's z-up equirectangular lookup: a
    horizon-to-zenith gradient plus a sun disk of the given radiance."""
    v = (np.arange(height, dtype=np.float64) + 0.5) / height      # v = 1 - θ/π (row 0 = nadir)
    u = (np.arange(width, dtype=np.float64) + 0.5) / width
    theta = (1.0 - v) * np.pi
    phi = u * 2.0 * np.pi - np.pi
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    d = np.stack([st * np.cos(phi)[None, :], st * np.sin(phi)[None, :], np.broadcast_to(ct, (height, width))], -1)
    up = np.clip(d[..., 2], 0.0, 1.0)
    zenith = np.array([0.25, 0.45, 1.0])
    horizon = np.array([0.9, 0.95, 1.0])
    ground = np.array([0.15, 0.13, 0.12])
    sky = horizon[None, None, :] * (1.0 - up[..., None]) + zenith[None, None, :] * up[..., None]
    img = np.where(d[..., 2:3] >= 0.0, sky, ground[None, None, :])
    s = np.asarray(sun_dir, np.float64)
    s /= np.linalg.norm(s)
    cosang = (d * s[None, None, :]).sum(-1)
    img[cosang > np.cos(np.radians(sun_angle_deg / 2.0))] = sun_radiance
    return np.ascontiguousarray(img.astype(f32))


def atrium_scene(width: int = 1920, height: int = 1080, target_tris: int = 300_000, seed: int = 1,
                 env_size: tuple[int, int] = (2048, 1024)) -> Scene:
    """Procedural ~300k-triangle atrium (z up): a jittered grid of tessellated columns
    joined by half-torus arches over a floor, per-object albedo U[0.2, 0.8] and GGX
    roughness U[0.05, 0.9]; a sky envmap with a sun disk (camera misses) and a
    matching directional sun light (the envmap emits no photons in the reference)."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    nx = ny = 12
    pitch = 2.0
    col_r, col_h = 0.25, 4.0
    seg, rings = 32, 14            # 896 triangles per column
    arc_seg, tube_seg = 40, 16     # 1280 triangles per arch
    floor = b.material(default_material((0.6, 0.6, 0.6), roughness=0.8))
    ext = pitch * (nx - 1)
    b.add_quad((-2, -2, 0), (ext + 2, -2, 0), (ext + 2, ext + 2, 0), (-2, ext + 2, 0), floor)
    pts = {}
    for i in range(nx):
        for j in range(ny):
            jit = rng.uniform(-0.25, 0.25, size=2)
            pts[(i, j)] = (i * pitch + jit[0], j * pitch + jit[1])
    budget = target_tris - 2
    objects = []
    for i in range(nx):
        for j in range(ny):
            objects.append(("col", (i, j)))
    for i in range(nx - 1):
        for j in range(ny):
            objects.append(("archx", (i, j)))
    for i in range(nx):
        for j in range(ny - 1):
            objects.append(("archy", (i, j)))
    for kind, (i, j) in objects:
        if budget <= 0:
            break
        mat = b.material(default_material(tuple(rng.uniform(0.2, 0.8, size=3)),
                                          roughness=float(rng.uniform(0.05, 0.9))))
        if kind == "col":
            tris, norms = _cylinder((*pts[(i, j)], 0.0), col_r, col_h, seg, rings)
        else:
            q = pts[(i + 1, j)] if kind == "archx" else pts[(i, j + 1)]
            tris, norms = _arch((*pts[(i, j)], 0.0), (*q, 0.0), col_h, 0.18, arc_seg, tube_seg)
        if tris.shape[0] > budget:
            tris, norms = tris[:budget], norms[:budget]
        budget -= tris.shape[0]
        b.add_mesh(tris, mat, normals=norms)
    sun = np.array([0.35, 0.25, 0.9])
    sun /= np.linalg.norm(sun)
    b.add_directional_light(-sun.astype(f32), (3.0, 2.9, 2.7))
    env = sky_envmap(env_size[0], env_size[1], sun_dir=sun)
    cam = make_camera((-1.0, -1.5, 1.7), (ext * 0.6, ext * 0.5, 2.2), (0.0, 0.0, 1.0), 70.0, width / height)
    return b.build(cam, name="atrium", envmap=env)


def sphere_room(width: int, height: int, target_tris: int, rng: np.random.Generator, sky: bool = False) -> Scene:
    """Cornell room (5 walls, 2 boxes, quad light) filled with icospheres up to
    about `target_tris` triangles: diffuse (albedo U[0.2, 0.8], roughness
    U[0.05, 0.9]), conductor or dielectric.  With `sky`, misses through the open
    front see a sky envmap and a weak directional light joins the area light."""
    b = SceneBuilder()
    white = b.material(default_material(WHITE))
    red = b.material(default_material(RED))
    green = b.material(default_material(GREEN))
    light = b.material(default_material(WHITE, emissive=(17.0, 17.0, 17.0)))
    _cornell_room(b, white, red, green, light)
    budget = target_tris - b.n_tris
    while budget >= 20:
        level = 4
        while level > 0 and 20 * 4 ** level > budget:
            level -= 1
        u = rng.uniform()
        if u < 0.7:
            m = default_material(tuple(rng.uniform(0.2, 0.8, size=3)), roughness=float(rng.uniform(0.05, 0.9)))
        elif u < 0.85:
            m = default_material(tuple(rng.uniform(0.5, 0.95, size=3)), metallic=1.0,
                                 roughness=float(rng.uniform(0.05, 0.5)))
        else:
            m = default_material((1.0, 1.0, 1.0), transmission=1.0, ior=1.5, roughness=0.05)
        r = float(rng.uniform(0.02, 0.09))
        c = rng.uniform(r, 1.0 - r, size=3)
        c[1] = rng.uniform(r, 0.8)
        b.add_sphere(tuple(c), r, b.material(m), level)
        budget -= 20 * 4 ** level
    env = None
    if sky:
        b.add_directional_light((0.2, -0.9, -0.35), (0.4, 0.4, 0.38))
        env = sky_envmap(256, 128, sun_radiance=50.0)
    cam = make_camera((0.5, 0.5, 2.2), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0), 45.0, width / height)
    return b.build(cam, name="sphere_room", envmap=env)


def adjoint_scene(width: int = 512, height: int = 512, seed: int = 1) -> Scene:
    """Eight-material room for gradient checks: floor, ceiling + back wall,
    left wall, right wall, tall box, short box, a GGX conductor sphere and the
    quad light; hidden albedo and roughness U[0.1, 0.9] drawn from PCG64(seed).
    Material order = parameter order."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = SceneBuilder()

    def mat(**kw):
        return b.material(default_material(tuple(rng.uniform(0.1, 0.9, size=3)),
                                           roughness=float(rng.uniform(0.1, 0.9)), **kw))

    floor, back, left, right, tall, short = (mat() for _ in range(6))
    sphere = mat(metallic=1.0)
    light = b.material(default_material((0.8, 0.8, 0.8), emissive=(17.0, 17.0, 17.0)))
    b.add_quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0), floor)
    b.add_quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1), back)
    b.add_quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), back)
    b.add_quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1), left)
    b.add_quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0), right)
    b.add_box((0.18, 0.0, 0.18), (0.48, 0.6, 0.48), tall)
    b.add_box((0.52, 0.0, 0.50), (0.82, 0.3, 0.80), short)
    b.add_sphere((0.68, 0.42, 0.65), 0.12, sphere, 3)
    y = 0.999
    b.add_quad((0.375, y, 0.375), (0.625, y, 0.375), (0.625, y, 0.625), (0.375, y, 0.625), light)
    cam = make_camera((0.5, 0.5, 2.2), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0), 45.0, width / height)
    return b.build(cam, name="adjoint8")


def batch_workload(n_scenes: int = 64, seed: int = 1):
    """The batched benchmark: n scenes with triangle counts log-uniform in
    [2k, 200k], square resolutions from {256, 512, 1024}, photons per pass
    log-uniform in [2^16, 2^21]; reference defaults otherwise (grid 64,
    16 NEE samples, 16 bounces, α 0.7).  Scene i draws from PCG64(seed + i).
    Returns [(PassConfig, Scene)]."""
    from . import PassConfig

    out = []
    for i in range(n_scenes):
        rng = np.random.Generator(np.random.PCG64(seed + i))
        tris = int(round(math.exp(rng.uniform(math.log(2000), math.log(200000)))))
        res = int(rng.choice([256, 512, 1024]))
        photons = int(round(math.exp(rng.uniform(math.log(1 << 16), math.log(1 << 21)))))
        scene = sphere_room(res, res, tris, rng, sky=(i % 4 == 3))
        cfg = PassConfig(width=res, height=res, photons_per_round=photons, max_photons_per_batch=1 << 20,
                         max_photon_bounces=16, grid_res=64, direct_samples=16, seed=seed + i, sppm_alpha=0.7)
        out.append((cfg, scene))
    return out


# Configurations of the benchmark workloads (sizes of the published configs)
CONFIGS = {
    "cornell": dict(scene=cornell_box, width=512, height=512, photons_per_round=1 << 18, rounds=16,
                    max_photon_bounces=8, radius=0.02, sppm_alpha=0.7, seed=1),
    "caustic": dict(scene=caustic_scene, width=1024, height=1024, photons_per_round=1 << 21, rounds=64,
                    max_photon_bounces=16, radius=0.01, sppm_alpha=0.7, seed=1),
    "atrium": dict(scene=atrium_scene, width=1920, height=1080, photons_per_round=1 << 22, rounds=32,
                   max_photon_bounces=16, radius=None, sppm_alpha=0.7, seed=1),
}
