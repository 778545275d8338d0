"""A minimal glTF 2.0 writer for the importer tests: builds the JSON document
and one binary buffer, and writes a .glb container or a .gltf file with the
buffer in a .bin file or a data URI."""
import base64
import json
import struct
from pathlib import Path

import numpy as np

_CT = {np.dtype(np.int8): 5120, np.dtype(np.uint8): 5121, np.dtype(np.int16): 5122, np.dtype(np.uint16): 5123,
       np.dtype(np.uint32): 5125, np.dtype(np.float32): 5126}
_TYPE = {1: "SCALAR", 2: "VEC2", 3: "VEC3", 4: "VEC4", 16: "MAT4"}


class Writer:
    def __init__(self):
        self.doc = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": [], "meshes": [],
                    "materials": [], "nodes": [], "scenes": [{"nodes": []}], "scene": 0}
        self.bin = bytearray()

    def _pad(self):
        while len(self.bin) % 4:
            self.bin.append(0)

    def view(self, data: bytes, stride: int | None = None) -> int:
        self._pad()
        v = {"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        self.bin += data
        self.doc["bufferViews"].append(v)
        return len(self.doc["bufferViews"]) - 1

    def accessor(self, a: np.ndarray, normalized: bool = False, view: int | None = None, offset: int = 0,
                 count: int | None = None) -> int:
        a = np.ascontiguousarray(a)
        width = 1 if a.ndim == 1 else a.shape[1]
        acc = {"componentType": _CT[a.dtype], "count": int(count if count is not None else a.shape[0]),
               "type": _TYPE[width]}
        acc["bufferView"] = self.view(a.tobytes()) if view is None else view
        if offset:
            acc["byteOffset"] = offset
        if normalized:
            acc["normalized"] = True
        if a.dtype == np.float32 and a.ndim == 2 and width == 3:
            acc["min"] = a.min(0).tolist()
            acc["max"] = a.max(0).tolist()
        self.doc["accessors"].append(acc)
        return len(self.doc["accessors"]) - 1

    def material(self, **m) -> int:
        self.doc["materials"].append(m)
        return len(self.doc["materials"]) - 1

    def mesh(self, primitives: list[dict]) -> int:
        self.doc["meshes"].append({"primitives": primitives})
        return len(self.doc["meshes"]) - 1

    def node(self, root: bool = True, **n) -> int:
        self.doc["nodes"].append(n)
        i = len(self.doc["nodes"]) - 1
        if root:
            self.doc["scenes"][0]["nodes"].append(i)
        return i

    def camera(self, yfov: float, aspect: float | None) -> int:
        per = {"yfov": yfov, "znear": 0.01}
        if aspect is not None:
            per["aspectRatio"] = aspect
        self.doc.setdefault("cameras", []).append({"type": "perspective", "perspective": per})
        return len(self.doc["cameras"]) - 1

    def light(self, **spec) -> int:
        ext = self.doc.setdefault("extensions", {}).setdefault("KHR_lights_punctual", {"lights": []})
        ext["lights"].append(spec)
        self.doc.setdefault("extensionsUsed", [])
        if "KHR_lights_punctual" not in self.doc["extensionsUsed"]:
            self.doc["extensionsUsed"].append("KHR_lights_punctual")
        return len(ext["lights"]) - 1

    def image_embedded(self, png: bytes) -> int:
        self.doc.setdefault("images", []).append({"bufferView": self.view(png), "mimeType": "image/png"})
        self.doc.setdefault("textures", []).append({"source": len(self.doc["images"]) - 1})
        return len(self.doc["textures"]) - 1

    def image_file(self, uri: str) -> int:
        self.doc.setdefault("images", []).append({"uri": uri})
        self.doc.setdefault("textures", []).append({"source": len(self.doc["images"]) - 1})
        return len(self.doc["textures"]) - 1

    def _finish(self, uri: str | None) -> dict:
        self._pad()
        doc = json.loads(json.dumps(self.doc))
        buf = {"byteLength": len(self.bin)}
        if uri is not None:
            buf["uri"] = uri
        doc["buffers"] = [buf]
        return doc

    def write_glb(self, path: Path) -> Path:
        js = json.dumps(self._finish(None)).encode()
        js += b" " * ((4 - len(js) % 4) % 4)
        total = 12 + 8 + len(js) + 8 + len(self.bin)
        with open(path, "wb") as f:
            f.write(struct.pack("<III", 0x46546C67, 2, total))
            f.write(struct.pack("<II", len(js), 0x4E4F534A) + js)
            f.write(struct.pack("<II", len(self.bin), 0x004E4942) + bytes(self.bin))
        return path

    def write_gltf(self, path: Path, data_uri: bool = False) -> Path:
        if data_uri:
            uri = "data:application/octet-stream;base64," + base64.b64encode(bytes(self.bin)).decode()
        else:
            uri = path.stem + ".bin"
            (path.parent / uri).write_bytes(bytes(self.bin))
        path.write_text(json.dumps(self._finish(uri)))
        return path


def png_bytes(rgb: np.ndarray, mode: str = "RGB") -> bytes:
    import io

    from PIL import Image

    out = io.BytesIO()
    Image.fromarray(rgb, mode=mode).save(out, format="PNG")
    return out.getvalue()


def write_scene_glb(scene, path: Path, yfov: float, aspect: float, camera_translation) -> Path:
    """A procedural scene (phosphor_amd.scenes) as a .glb: one mesh per run of
    triangles with one material (non-indexed, per-vertex normals), the
    material's factors (emission above 1 through KHR_materials_emissive_strength),
    and a camera node at `camera_translation` looking down -z."""
    w = Writer()
    tm = scene.tri_mat_index
    start = 0
    while start < len(tm):
        end = start
        while end < len(tm) and tm[end] == tm[start]:
            end += 1
        m = scene.materials[tm[start]]
        e = np.asarray(m["emissive"][:3], np.float64)
        strength = float(e.max()) if e.max() > 1.0 else 1.0
        mat = {"pbrMetallicRoughness": {"baseColorFactor": [float(v) for v in m["base_color"]],
                                        "metallicFactor": float(m["metallic"]),
                                        "roughnessFactor": float(m["roughness"])},
               "emissiveFactor": [float(v) / strength for v in e]}
        if strength != 1.0:
            mat["extensions"] = {"KHR_materials_emissive_strength": {"emissiveStrength": strength}}
        pos = np.ascontiguousarray(scene.tri_verts[start:end, :, :3].reshape(-1, 3))
        nrm = np.ascontiguousarray(scene.tri_normals[start:end, :, :3].reshape(-1, 3))
        prim = {"attributes": {"POSITION": w.accessor(pos), "NORMAL": w.accessor(nrm)}, "material": w.material(**mat)}
        w.node(mesh=w.mesh([prim]))
        start = end
    w.node(camera=w.camera(yfov, aspect), translation=list(camera_translation))
    return w.write_glb(path)
