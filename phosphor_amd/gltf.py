"""glTF 2.0 scene import: ``read_gltf_scene`` (src/parsers/scene.cpp:206-233).

The reference reads glTF through assimp v6.0.5 (CMakeLists.txt:55-61;
aiProcess_Triangulate | JoinIdenticalVertices | CalcTangentSpace) and walks
assimp's node tree; this module reads the glTF document itself (``.glb``
binary container or ``.gltf`` JSON with external or data-URI buffers) and
applies the mapping that assimp's glTF 2.0 importer publishes, then the
reference's own parse functions:

* nodes (scene.cpp:183-204): the default scene's root nodes depth-first,
  children in order; at each node its camera, then its light, then its mesh's
  primitives (assimp makes one mesh per primitive, in order); the transform
  is parent · local, local = ``matrix`` or T · R · S.
* meshes (scene.cpp:36-131): triangles of every primitive (TRIANGLES, and
  TRIANGLE_STRIP / TRIANGLE_FAN split as the glTF specification defines
  them; point and line primitives are skipped); positions by the transform,
  normals and tangents by its inverse transpose, normalised; face normals
  where NORMAL is absent; where TANGENT is absent the reference's fallback
  (tangent from the UV derivatives, (1, 0, 0) for degenerate UVs).  assimp
  flips texture V (v ↦ 1 − v), so the same is done here.  Emissive meshes
  become LIGHT_TEXTURED lights with power π·E·area (scene.cpp:115-130).
* materials (material.cpp): one per mesh, from the reference's defaults
  (material.cpp:9-30) overwritten by what the file sets: baseColorFactor,
  metallicFactor and roughnessFactor (glTF defaults 1, 1, 1 where the
  material exists; a primitive without material gets the glTF default
  material), emissiveFactor × KHR_materials_emissive_strength,
  KHR_materials_transmission, KHR_materials_ior, KHR_materials_volume
  (attenuation colour w = 1 when given), texture slots and
  KHR_texture_transform.
* textures (texture.cpp): decoded to 8-bit RGB (stb_image with 3 channels in
  the reference, Pillow here), named by their file name or, for embedded
  images, ``*<index among embedded images>`` as assimp names them, loaded
  once each in the order material.cpp's parse_textures asks for them, and
  packed into one atlas (shading.cpp:24-35).
* lights (scene.cpp:151-181, KHR_lights_punctual): power = colour ·
  intensity / 683 (LUMINOUS_EFFICACY, constants.h:18); spot cones in
  radians; directions = rotation of (0, 0, −1).
* cameras (scene.cpp:133-149): perspective only; assimp's horizontal field
  of view 2·atan(tan(yfov/2)·aspectRatio); position, look direction (0, 0,
  −1) and up (0, 1, 0) through the node transform; the first camera is the
  one rendered, at the size ``resolve_render_size`` gives (output.cpp:13-37),
  focused with the defocus angle and focus distance (camera.cpp:15-23).

A scene without triangles or without a camera is refused (main.cpp:39-47).
"""
from __future__ import annotations

import base64
import io
import json
import math
import struct
from pathlib import Path

import numpy as np

from .abi import LIGHT_DTYPE, LIGHT_TEXTURED, MATERIAL_DTYPE, NO_TEXTURE, TEX_META_DTYPE, UV_BLOCK_DTYPE
from .scenes import Scene, as_float, default_material, make_camera, make_tbn

f32 = np.float32
PI = f32(3.141592654)                 # constants.h:14
EPS = f32(1e-5)                       # constants.h:11
LUMINOUS_EFFICACY = f32(683.0)        # constants.h:18
DEFAULT_CAMERA_ASPECT = f32(16.0 / 9.0)  # constants.h:28
DEFAULT_SIZE = 1024                   # cmd_args.hpp:11-12

_GLB_MAGIC, _CHUNK_JSON, _CHUNK_BIN = 0x46546C67, 0x4E4F534A, 0x004E4942
_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_WIDTH = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}
# texture slots in the order parse_textures loads them (texture.cpp:80-91):
# diffuse and base colour are both glTF's baseColorTexture, metalness and
# roughness both its metallicRoughnessTexture
_TEXTURE_ORDER = ("base", "base", "emissive", "normal", "occlusion", "metal_rough", "metal_rough", "transmission")


class GltfError(ValueError):
    """The file is not a glTF 2.0 scene this importer (or the reference) can render."""


class _Document:
    def __init__(self, path: Path):
        self.dir = path.parent
        data = path.read_bytes()
        self.bin = None
        if data[:4] == b"glTF":
            magic, version, length = struct.unpack_from("<III", data, 0)
            if magic != _GLB_MAGIC or version != 2:
                raise GltfError(f"{path}: not a glTF 2.0 binary container")
            off, js = 12, None
            while off + 8 <= min(length, len(data)):
                clen, ctype = struct.unpack_from("<II", data, off)
                chunk = data[off + 8:off + 8 + clen]
                if ctype == _CHUNK_JSON and js is None:
                    js = json.loads(chunk.decode("utf-8"))
                elif ctype == _CHUNK_BIN and self.bin is None:
                    self.bin = chunk
                off += 8 + clen
            if js is None:
                raise GltfError(f"{path}: no JSON chunk")
            self.doc = js
        else:
            self.doc = json.loads(data.decode("utf-8"))
        if not str(self.doc.get("asset", {}).get("version", "")).startswith("2"):
            raise GltfError(f"{path}: asset.version is not 2.x")
        self._buffers: dict[int, bytes] = {}

    def _uri_bytes(self, uri: str) -> bytes:
        if uri.startswith("data:"):
            head, _, payload = uri.partition(",")
            if not head.endswith(";base64"):
                raise GltfError("data URI without base64 payload")
            return base64.b64decode(payload)
        return (self.dir / uri).read_bytes()

    def buffer(self, i: int) -> bytes:
        if i not in self._buffers:
            b = self.doc["buffers"][i]
            if "uri" in b:
                self._buffers[i] = self._uri_bytes(b["uri"])
            elif self.bin is not None:
                self._buffers[i] = self.bin
            else:
                raise GltfError(f"buffer {i} has no data")
        return self._buffers[i]

    def view_bytes(self, i: int) -> tuple[bytes, int, int | None]:
        v = self.doc["bufferViews"][i]
        buf = self.buffer(v["buffer"])
        off = v.get("byteOffset", 0)
        return buf[off:off + v["byteLength"]], v["byteLength"], v.get("byteStride")

    def accessor(self, i: int) -> np.ndarray:
        """(count, width) array of accessor i; normalised integers become
        floats as the glTF specification defines them."""
        a = self.doc["accessors"][i]
        if "sparse" in a:
            raise GltfError(f"accessor {i}: sparse accessors are not supported")
        dt = np.dtype(_COMPONENT[a["componentType"]]).newbyteorder("<")
        width, count = _WIDTH[a["type"]], a["count"]
        if "bufferView" not in a:
            out = np.zeros((count, width), dt)
        else:
            raw, _, stride = self.view_bytes(a["bufferView"])
            elem = dt.itemsize * width
            stride = stride or elem
            off = a.get("byteOffset", 0)
            if count and off + stride * (count - 1) + elem > len(raw):
                raise GltfError(f"accessor {i} reaches beyond its buffer view")
            strided = np.lib.stride_tricks.as_strided(np.frombuffer(raw, np.uint8, offset=off),
                                                      shape=(count, elem), strides=(stride, 1))
            out = np.ascontiguousarray(strided).view(dt).reshape(count, width)
        if a.get("normalized") and dt.kind in "iu":
            scale = float(np.iinfo(dt).max)
            out = np.maximum(out.astype(np.float64) / scale, -1.0).astype(f32)
        return out


def _quat_matrix(q) -> np.ndarray:
    x, y, z, w = (f32(c) for c in q)
    two = f32(2)
    return np.array([[f32(1) - two * (y * y + z * z), two * (x * y - z * w), two * (x * z + y * w)],
                     [two * (x * y + z * w), f32(1) - two * (x * x + z * z), two * (y * z - x * w)],
                     [two * (x * z - y * w), two * (y * z + x * w), f32(1) - two * (x * x + y * y)]], f32)


def _local_matrix(node: dict) -> np.ndarray:
    if "matrix" in node:
        return np.array(node["matrix"], f32).reshape(4, 4).T     # column-major in the file
    m = np.eye(4, dtype=f32)
    if "translation" in node:
        t = np.eye(4, dtype=f32)
        t[:3, 3] = node["translation"]
        m = (m @ t).astype(f32)
    if "rotation" in node:
        r = np.eye(4, dtype=f32)
        r[:3, :3] = _quat_matrix(node["rotation"])
        m = (m @ r).astype(f32)
    if "scale" in node:
        s = np.diag([*node["scale"], 1.0]).astype(f32)
        m = (m @ s).astype(f32)
    return m


def _normalize(v: np.ndarray) -> np.ndarray:
    """glm::normalize in float32: v · (1 / sqrt(dot(v, v))), dot = (x² + y²) + z²."""
    v = v.astype(f32)
    d = (v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2]
    return (v * (f32(1) / np.sqrt(d))[..., None]).astype(f32)


def _cross(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return np.stack([a[..., 1] * b[..., 2] - b[..., 1] * a[..., 2],
                     a[..., 2] * b[..., 0] - b[..., 2] * a[..., 0],
                     a[..., 0] * b[..., 1] - b[..., 0] * a[..., 1]], axis=-1).astype(f32)


def _point(g: np.ndarray, p) -> np.ndarray:
    return (g[:3, :3] @ np.asarray(p, f32) + g[:3, 3]).astype(f32)


def _faces(indices: np.ndarray, mode: int) -> np.ndarray:
    n = indices.shape[0]
    if mode == 4:
        return indices[:n - n % 3].reshape(-1, 3)
    if mode == 5:   # strip: triangle i = (i, i+1+i%2, i+2-i%2)
        i = np.arange(max(n - 2, 0))
        return np.stack([indices[i], indices[i + 1 + i % 2], indices[i + 2 - i % 2]], axis=1)
    if mode == 6:   # fan: triangle i = (i+1, i+2, 0)
        i = np.arange(max(n - 2, 0))
        return np.stack([indices[i + 1], indices[i + 2], np.full_like(i, indices[0])], axis=1)
    return np.zeros((0, 3), np.int64)


class _Importer:
    def __init__(self, path: Path):
        self.d = _Document(path)
        self.doc = self.d.doc
        self.verts, self.normals, self.tangents, self.uvs, self.mat_index = [], [], [], [], []
        self.etris: list[tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]] = []
        self.n_etris = 0
        self.mats: list[np.ndarray] = []
        self.uvt: list[np.ndarray] = []
        self.lights: list[np.ndarray] = []
        self.cameras: list[tuple] = []
        self.tex_names: list[str] = []
        self.tex_pixels: list[tuple[int, int, bytes]] = []
        self._image_names = self._name_images()
        self._load_textures()

    # ---- textures (texture.cpp) ----
    def _name_images(self) -> list[str]:
        names, embedded = [], 0
        for im in self.doc.get("images", []):
            uri = im.get("uri")
            if uri is not None and not uri.startswith("data:"):
                names.append(Path(uri).name)
            else:
                names.append(f"*{embedded}")
                embedded += 1
        return names

    def _texture_image(self, info: dict | None) -> int | None:
        if not info or "index" not in info:
            return None
        tex = self.doc.get("textures", [])[info["index"]]
        return tex.get("source")

    def _slots(self, mat: dict) -> dict:
        pbr = mat.get("pbrMetallicRoughness", {})
        ext = mat.get("extensions", {})
        return {"base": pbr.get("baseColorTexture"), "metal_rough": pbr.get("metallicRoughnessTexture"),
                "emissive": mat.get("emissiveTexture"), "normal": mat.get("normalTexture"),
                "occlusion": mat.get("occlusionTexture"),
                "transmission": ext.get("KHR_materials_transmission", {}).get("transmissionTexture")}

    def _decode(self, image: int) -> tuple[int, int, bytes]:
        from PIL import Image

        im = self.doc["images"][image]
        if "bufferView" in im:
            data = self.d.view_bytes(im["bufferView"])[0]
        else:
            data = self.d._uri_bytes(im["uri"])
        with Image.open(io.BytesIO(data)) as pic:
            rgb = pic.convert("RGB")
            return rgb.width, rgb.height, rgb.tobytes()

    def _load_textures(self) -> None:
        """parse_textures (texture.cpp:80-91): every material, every slot, each
        texture name loaded once."""
        for mat in self.doc.get("materials", []):
            slots = self._slots(mat)
            for slot in _TEXTURE_ORDER:
                img = self._texture_image(slots[slot])
                if img is None:
                    continue
                name = self._image_names[img]
                if name in self.tex_names:
                    continue
                self.tex_names.append(name)
                self.tex_pixels.append(self._decode(img))

    def _find_texture(self, info: dict | None) -> int:
        img = self._texture_image(info)
        if img is None:
            return NO_TEXTURE
        name = self._image_names[img]
        return self.tex_names.index(name) if name in self.tex_names else NO_TEXTURE

    # ---- materials (material.cpp) ----
    @staticmethod
    def _uv_transform(info: dict | None):
        """KHR_texture_transform as assimp stores it (rotation negated, offset
        moved to assimp's texture-space origin)."""
        t = (info or {}).get("extensions", {}).get("KHR_texture_transform")
        if t is None:
            return None
        sx, sy = (f32(v) for v in t.get("scale", (1.0, 1.0)))
        rot = -f32(t.get("rotation", 0.0))
        ox, oy = (f32(v) for v in t.get("offset", (0.0, 0.0)))
        c, s = f32(math.cos(-rot)), f32(math.sin(-rot))
        tx = (f32(0.5) * sx) * (-c + s + f32(1)) + ox
        ty = ((f32(0.5) * sy) * (s + c - f32(1))) + f32(1) - sy - oy
        return (tx, ty), (sx, sy), rot

    def _material(self, mat: dict | None) -> int:
        m = default_material()
        uv = np.zeros(1, UV_BLOCK_DTYPE)[0]
        for k in UV_BLOCK_DTYPE.names:
            uv[k]["uv_scale"] = 1.0
        if mat is None:
            mat = {}  # the glTF default material: metallic and roughness 1
        pbr = mat.get("pbrMetallicRoughness", {})
        ext = mat.get("extensions", {})
        m["base_color"] = pbr.get("baseColorFactor", (1.0, 1.0, 1.0, 1.0))
        m["metallic"] = pbr.get("metallicFactor", 1.0)
        m["roughness"] = pbr.get("roughnessFactor", 1.0)
        strength = f32(ext.get("KHR_materials_emissive_strength", {}).get("emissiveStrength", 1.0))
        m["emissive"] = (*(np.asarray(mat.get("emissiveFactor", (0.0, 0.0, 0.0)), f32) * strength), 0.0)
        if "KHR_materials_transmission" in ext:
            m["transmission"] = ext["KHR_materials_transmission"].get("transmissionFactor", 0.0)
        if "KHR_materials_ior" in ext:
            m["ior"] = ext["KHR_materials_ior"].get("ior", 1.5)
        if "KHR_materials_volume" in ext:
            vol = ext["KHR_materials_volume"]
            m["thickness"] = vol.get("thicknessFactor", 0.0)
            if "attenuationDistance" in vol:
                m["att_dist"] = vol["attenuationDistance"]
            if "attenuationColor" in vol:
                m["att_color"] = (*vol["attenuationColor"], 1.0)
        slots = self._slots(mat)
        for field, slot in (("diff_index", "base"), ("emis_index", "emissive"), ("norm_index", "normal"),
                            ("occlusion_index", "occlusion"), ("metal_rough_index", "metal_rough"),
                            ("transmission_index", "transmission")):
            m[field] = self._find_texture(slots[slot])
        for block, slot in (("diff", "base"), ("norm", "normal"), ("emis", "emissive"),
                            ("metal_rough", "metal_rough"), ("occlusion", "occlusion"),
                            ("transmission", "transmission")):
            t = self._uv_transform(slots[slot])
            if t is not None:
                uv[block]["uv_offset"], uv[block]["uv_scale"], uv[block]["uv_rotation"] = t
        self.mats.append(m)
        self.uvt.append(uv)
        return len(self.mats) - 1

    # ---- meshes (scene.cpp:36-131) ----
    def _primitive(self, prim: dict, g: np.ndarray) -> None:
        mode = prim.get("mode", 4)
        if mode not in (4, 5, 6):
            return
        attrs = prim["attributes"]
        pos = self.d.accessor(attrs["POSITION"]).astype(f32)
        idx = (self.d.accessor(prim["indices"]).ravel().astype(np.int64) if "indices" in prim
               else np.arange(pos.shape[0], dtype=np.int64))
        faces = _faces(idx, mode)
        mat_index = self._material(self.doc["materials"][prim["material"]] if "material" in prim else None)
        nf = faces.shape[0]
        if nf == 0:
            return
        nm = np.linalg.inv(g[:3, :3].astype(np.float64)).T.astype(f32)
        p = np.stack([(pos[faces[:, k]] @ g[:3, :3].T + g[:3, 3]).astype(f32) for k in range(3)], axis=1)
        if "TEXCOORD_0" in attrs:
            tc = self.d.accessor(attrs["TEXCOORD_0"]).astype(f32)
            tc = np.stack([tc[:, 0], f32(1) - tc[:, 1]], axis=1)          # assimp flips V
            uv = np.stack([tc[faces[:, k]] for k in range(3)], axis=1)
        else:
            uv = np.zeros((nf, 3, 2), f32)
        if "NORMAL" in attrs:
            nrm = self.d.accessor(attrs["NORMAL"]).astype(f32)
            n = np.stack([_normalize(nrm[faces[:, k]] @ nm.T) for k in range(3)], axis=1)
        else:
            fn = _normalize(_cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]))
            n = np.repeat(fn[:, None], 3, axis=1)
        if "TANGENT" in attrs:
            tan = self.d.accessor(attrs["TANGENT"]).astype(f32)[:, :3]
            t = np.stack([_normalize(tan[faces[:, k]] @ nm.T) for k in range(3)], axis=1)
        else:
            e1, e2 = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]
            d1, d2 = uv[:, 1] - uv[:, 0], uv[:, 2] - uv[:, 0]
            det = d1[:, 0] * d2[:, 1] - d2[:, 0] * d1[:, 1]
            ft = np.tile(np.array([1, 0, 0], f32), (nf, 1))
            ok = np.abs(det) >= EPS
            inv = np.where(ok, f32(1) / np.where(ok, det, f32(1)), f32(0)).astype(f32)
            cand = (inv[:, None] * (d2[:, 1:2] * e1 - d1[:, 1:2] * e2)).astype(f32)
            ft[ok] = cand[ok]
            ft = _normalize(ft @ nm.T)
            t = np.repeat(ft[:, None], 3, axis=1)
        v4, n4, t4 = (np.zeros((nf, 3, 4), f32) for _ in range(3))
        v4[..., :3], n4[..., :3], t4[..., :3] = p, n, t
        uv8 = np.zeros((nf, 8), f32)
        uv8[:, :6] = uv.reshape(nf, 6)
        self.verts.append(v4)
        self.normals.append(n4)
        self.tangents.append(t4)
        self.uvs.append(uv8)
        self.mat_index.append(np.full(nf, mat_index, np.uint32))
        em = np.asarray(self.mats[mat_index]["emissive"][:3], f32)
        if np.any(em > EPS):
            total = f32(0)
            for k in range(nf):
                c = _cross(p[k, 1] - p[k, 0], p[k, 2] - p[k, 0])
                total = f32(total + f32(0.5) * f32(np.sqrt(f32(c[0] * c[0] + c[1] * c[1]) + c[2] * c[2])))
            light = np.zeros(1, LIGHT_DTYPE)[0]
            light["kind"] = LIGHT_TEXTURED
            light["power"][:3] = (PI * em).astype(f32) * total
            light["aux"] = (as_float(int(self.mats[mat_index]["emis_index"])), as_float(self.n_etris), as_float(nf),
                            total)
            self.lights.append(light)
            self.etris.append((v4.copy(), n4.copy(), uv8.copy(), np.full(nf, mat_index, np.uint32)))
            self.n_etris += nf

    # ---- lights and cameras (scene.cpp:133-181) ----
    def _light(self, spec: dict, g: np.ndarray) -> None:
        kind = spec.get("type")
        power = (np.asarray(spec.get("color", (1.0, 1.0, 1.0)), f32) * f32(spec.get("intensity", 1.0))
                 / LUMINOUS_EFFICACY).astype(f32)
        light = np.zeros(1, LIGHT_DTYPE)[0]
        light["power"][:3] = power
        rot = g[:3, :3]
        if kind == "point":
            light["kind"] = 0
            light["position"][:3] = _point(g, (0, 0, 0))
        elif kind == "spot":
            spot = spec.get("spot", {})
            light["kind"] = 1
            light["position"][:3] = _point(g, (0, 0, 0))
            light["direction"][:3] = _normalize((rot @ np.array([0, 0, -1], f32)).astype(f32))
            light["aux"] = (spot.get("innerConeAngle", 0.0), spot.get("outerConeAngle", math.pi / 4), 0.0, 0.0)
        elif kind == "directional":
            d = _normalize((rot @ np.array([0, 0, -1], f32)).astype(f32))
            t, b = make_tbn(d)
            light["kind"] = 2
            light["direction"][:3] = d
            light["tangent"][:3] = t
            light["bitangent"][:3] = b
        else:
            return
        self.lights.append(light)

    def _camera(self, spec: dict, g: np.ndarray) -> None:
        if spec.get("type") != "perspective":
            raise GltfError("hfov must be between 0 and 180 degrees exclusive (orthographic camera)")
        per = spec["perspective"]
        aspect = f32(per.get("aspectRatio", 0.0))
        if not aspect > 0:
            raise GltfError("aspect_ratio must be positive")
        yfov = f32(per["yfov"])
        hfov = f32(2.0) * f32(math.atan(f32(math.tan(yfov * f32(0.5))) * aspect))
        if not (f32(0) < hfov < PI):
            raise GltfError("hfov must be between 0 and 180 degrees exclusive")
        position = _point(g, (0, 0, 0))
        look = (g[:3, :3] @ np.array([0, 0, -1], f32)).astype(f32)
        up = _normalize((g[:3, :3] @ np.array([0, 1, 0], f32)).astype(f32))
        self.cameras.append((position, (position + look).astype(f32), up, f32(hfov * (f32(180.0) / PI)), aspect))

    def _node(self, i: int, parent: np.ndarray, depth: int = 0) -> None:
        if depth > 1000:
            raise GltfError("node hierarchy is cyclic or too deep")
        node = self.doc["nodes"][i]
        g = (parent @ _local_matrix(node)).astype(f32)
        if "camera" in node:
            self._camera(self.doc["cameras"][node["camera"]], g)
        light = node.get("extensions", {}).get("KHR_lights_punctual", {}).get("light")
        if light is not None:
            self._light(self.doc["extensions"]["KHR_lights_punctual"]["lights"][light], g)
        if "mesh" in node:
            for prim in self.doc["meshes"][node["mesh"]]["primitives"]:
                self._primitive(prim, g)
        for c in node.get("children", []):
            self._node(c, g, depth + 1)

    def run(self) -> None:
        scenes_ = self.doc.get("scenes", [])
        if scenes_:
            roots = scenes_[self.doc.get("scene", 0)].get("nodes", [])
        else:
            roots = list(range(len(self.doc.get("nodes", []))))
        for r in roots:
            self._node(r, np.eye(4, dtype=f32))


def resolve_render_size(width: int | None, height: int | None, base_aspect: float) -> tuple[int, int, np.float32]:
    """output.cpp:13-37: both or neither given → as given (default 1024²)
    with their aspect; one given → the other from the camera's aspect."""
    base = f32(base_aspect)
    if (width is None) == (height is None):
        w, h = width or DEFAULT_SIZE, height or DEFAULT_SIZE
        return w, h, f32(f32(w) / f32(h))
    if width is not None:
        return width, max(1, int(math.floor(float(f32(width) / base) + 0.5))), base
    return max(1, int(math.floor(float(f32(height) * base) + 0.5))), height, base


def read_gltf_scene(path, width: int | None = None, height: int | None = None, defocus_angle: float = 0.0,
                    focus_distance: float = 1.0, envmap: np.ndarray | None = None) -> Scene:
    """The scene of a .glb / .gltf file, ready for ``Renderer`` (render size
    in ``scene.meta["render_size"]``)."""
    path = Path(path)
    imp = _Importer(path)
    imp.run()
    if not imp.verts or sum(v.shape[0] for v in imp.verts) == 0:
        raise GltfError(f"{path}: empty scene, nothing to render")
    if not imp.cameras:
        raise GltfError(f"{path}: scene without cameras, nothing to render")
    position, look_at, up, hfov, aspect = imp.cameras[0]
    w, h, aspect = resolve_render_size(width, height, aspect)
    cam = make_camera(position, look_at, up, float(hfov), float(aspect), defocus_angle=defocus_angle,
                      focus_distance=focus_distance)
    tv = np.ascontiguousarray(np.concatenate(imp.verts))
    if imp.etris:
        ev, en, euv, em = (np.ascontiguousarray(np.concatenate([e[k] for e in imp.etris])) for k in range(4))
    else:
        ev, en, em = np.zeros((0, 3, 4), f32), np.zeros((0, 3, 4), f32), np.zeros(0, np.uint32)
        euv = np.zeros((0, 8), f32)
    meta = np.zeros(max(len(imp.tex_pixels), 0), TEX_META_DTYPE)
    atlas = bytearray()
    for k, (tw, th, px) in enumerate(imp.tex_pixels):
        meta[k] = (len(atlas), tw, th, 3)
        atlas += px
    lights = np.array(imp.lights, dtype=LIGHT_DTYPE) if imp.lights else np.zeros(0, LIGHT_DTYPE)
    return Scene(tv, np.ascontiguousarray(np.concatenate(imp.normals)),
                 np.ascontiguousarray(np.concatenate(imp.tangents)), np.ascontiguousarray(np.concatenate(imp.uvs)),
                 np.ascontiguousarray(np.concatenate(imp.mat_index)), ev, en, euv, em,
                 np.ascontiguousarray(np.array(imp.mats, dtype=MATERIAL_DTYPE)),
                 np.ascontiguousarray(np.array(imp.uvt, dtype=UV_BLOCK_DTYPE)), np.ascontiguousarray(lights), cam,
                 envmap=envmap, tex_meta=meta,
                 tex_atlas=np.frombuffer(bytes(atlas) or b"\0", np.uint8).copy(),
                 name=path.stem, meta={"render_size": (w, h), "textures": list(imp.tex_names),
                                       "cameras": len(imp.cameras)})
