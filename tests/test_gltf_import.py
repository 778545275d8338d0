"""glTF 2.0 import (phosphor_amd.gltf, src/parsers/scene.cpp + material.cpp +
texture.cpp with assimp v6.0.5's glTF mapping) on files written by
tests/gltf_files.py.  The reference's own models are LFS stubs, so every file
here is synthetic; expected values come from the glTF specification and the
reference's parse functions, computed independently in float64 where a
transform is involved (tolerance 2e-6 relative: float32 rounding of the
transform products), exactly where no arithmetic is."""
import math

import numpy as np
import pytest

from phosphor_amd import gltf
from phosphor_amd.abi import LIGHT_TEXTURED, NO_TEXTURE

from tests.gltf_files import Writer, png_bytes

F = np.float32


def _tri_mesh(w, positions, normals=None, uvs=None, tangents=None, indices=None, material=None, mode=None,
              uv_normalized=False):
    attrs = {"POSITION": w.accessor(np.asarray(positions, F))}
    if normals is not None:
        attrs["NORMAL"] = w.accessor(np.asarray(normals, F))
    if uvs is not None:
        attrs["TEXCOORD_0"] = w.accessor(uvs, normalized=uv_normalized)
    if tangents is not None:
        attrs["TANGENT"] = w.accessor(np.asarray(tangents, F))
    prim = {"attributes": attrs}
    if indices is not None:
        prim["indices"] = w.accessor(indices)
    if material is not None:
        prim["material"] = material
    if mode is not None:
        prim["mode"] = mode
    return w.mesh([prim])


def _camera_node(w, **kw):
    return w.node(camera=w.camera(math.radians(40.0), 1.5), **kw)


QUAD = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], F)


def test_identity_geometry_is_exact_and_defaults_apply(tmp_path):
    w = Writer()
    m = w.material(pbrMetallicRoughness={"baseColorFactor": [0.2, 0.4, 0.6, 1.0], "metallicFactor": 0.0,
                                         "roughnessFactor": 0.35})
    w.node(mesh=_tri_mesh(w, QUAD, indices=np.array([0, 1, 2, 0, 2, 3], np.uint16), material=m))
    w.node(mesh=_tri_mesh(w, QUAD + F(2), indices=np.array([0, 1, 2], np.uint8)))   # no material
    _camera_node(w, translation=[0.5, 0.5, 3.0])
    sc = gltf.read_gltf_scene(w.write_glb(tmp_path / "a.glb"))
    assert sc.n_triangles == 3
    want = np.array([QUAD[[0, 1, 2]], QUAD[[0, 2, 3]], (QUAD + F(2))[[0, 1, 2]]], F)
    assert np.array_equal(sc.tri_verts[:, :, :3], want) and np.all(sc.tri_verts[:, :, 3] == 0)
    # no NORMAL: the face normal; no TANGENT and no UVs: (1, 0, 0)
    assert np.allclose(sc.tri_normals[:, :, :3], [0, 0, 1], atol=1e-7)
    assert np.allclose(sc.tri_tangents[:, :, :3], [1, 0, 0], atol=1e-7)
    assert list(sc.tri_mat_index) == [0, 0, 1]
    m0, m1 = sc.materials
    assert np.allclose(m0["base_color"], [0.2, 0.4, 0.6, 1.0]) and m0["metallic"] == 0 and m0["roughness"] == F(0.35)
    # the glTF default material (assimp's): white, metallic 1, roughness 1
    assert np.array_equal(m1["base_color"], [1, 1, 1, 1]) and m1["metallic"] == 1 and m1["roughness"] == 1
    for m_ in (m0, m1):   # material.cpp:9-30 where the file says nothing
        assert m_["transmission"] == 0 and m_["ior"] == F(1.5) and m_["att_dist"] == F(1e9) and m_["thickness"] == 0
        assert np.array_equal(m_["att_color"], [1, 1, 1, 0]) and m_["diff_index"] == NO_TEXTURE
    assert sc.meta["render_size"] == (1024, 1024) and len(sc.lights) == 0


def test_node_transforms_normals_and_tangents(tmp_path):
    """T·R·S nodes under a matrix parent: positions by the global transform,
    normals and tangents by its inverse transpose, normalised."""
    w = Writer()
    q = np.array([0.1, -0.3, 0.2, 0.9]); q /= np.linalg.norm(q)
    normals = np.tile([0.0, 0.0, 1.0], (4, 1))
    tangents = np.tile([1.0, 0.0, 0.0, 1.0], (4, 1))
    mesh = _tri_mesh(w, QUAD, normals=normals, tangents=tangents, indices=np.array([0, 1, 2, 0, 2, 3], np.uint32))
    child = w.node(root=False, mesh=mesh, translation=[0.3, -1.0, 2.0], rotation=q.tolist(), scale=[2.0, 0.5, 1.5])
    parent = np.array([[1, 0, 0, 5], [0, 0, -1, 1], [0, 1, 0, -2], [0, 0, 0, 1]], float)   # row-major
    w.node(children=[child], matrix=parent.T.ravel().tolist())
    _camera_node(w)
    sc = gltf.read_gltf_scene(w.write_glb(tmp_path / "t.glb"))
    x, y, z, ww = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * ww), 2 * (x * z + y * ww)],
                  [2 * (x * y + z * ww), 1 - 2 * (x * x + z * z), 2 * (y * z - x * ww)],
                  [2 * (x * z - y * ww), 2 * (y * z + x * ww), 1 - 2 * (x * x + y * y)]])
    L = np.eye(4)
    L[:3, :3] = R @ np.diag([2.0, 0.5, 1.5])
    L[:3, 3] = [0.3, -1.0, 2.0]
    G = parent @ L
    tri = QUAD.astype(float)[[0, 1, 2, 0, 2, 3]].reshape(2, 3, 3)
    want = tri @ G[:3, :3].T + G[:3, 3]
    assert np.allclose(sc.tri_verts[:, :, :3], want, rtol=2e-6, atol=2e-6)
    N = np.linalg.inv(G[:3, :3]).T
    n = N @ [0, 0, 1]
    t = N @ [1, 0, 0]
    assert np.allclose(sc.tri_normals[:, :, :3], n / np.linalg.norm(n), atol=2e-6)
    assert np.allclose(sc.tri_tangents[:, :, :3], t / np.linalg.norm(t), atol=2e-6)


def test_strips_fans_index_types_interleaving_and_uv_flip(tmp_path):
    """TRIANGLE_STRIP and TRIANGLE_FAN split as the glTF specification
    defines them; u8/u16/u32 indices; interleaved attributes (byteStride);
    normalised u16 texture coordinates; V flipped as assimp does."""
    w = Writer()
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 2, 0]], F)
    uv16 = np.array([[0, 0], [65535, 0], [0, 32768], [65535, 65535], [16384, 49152]], np.uint16)
    strip = _tri_mesh(w, pos, uvs=uv16, uv_normalized=True, indices=np.arange(5, dtype=np.uint8), mode=5)
    fan = _tri_mesh(w, pos, indices=np.array([0, 1, 3, 4], np.uint32), mode=6)
    # interleaved position + normal, no indices
    inter = np.zeros((3, 6), F)
    inter[:, :3] = [[0, 0, 5], [1, 0, 5], [0, 1, 5]]
    inter[:, 3:] = [0, 0, -1]
    v = w.view(inter.tobytes(), stride=24)
    a_pos = w.accessor(inter[:, :3].copy(), view=v, count=3)
    a_nrm = w.accessor(inter[:, 3:].copy(), view=v, offset=12, count=3)
    il = w.mesh([{"attributes": {"POSITION": a_pos, "NORMAL": a_nrm}}])
    lines = w.mesh([{"attributes": {"POSITION": w.accessor(pos)}, "mode": 1}])
    for mesh in (strip, fan, il, lines):
        w.node(mesh=mesh)
    _camera_node(w)
    sc = gltf.read_gltf_scene(w.write_glb(tmp_path / "s.glb"))
    strip_idx = [(0, 1, 2), (1, 3, 2), (2, 3, 4)]      # p_i = {i, i+1+i%2, i+2-i%2}
    fan_idx = [(1, 3, 0), (3, 4, 0)]                   # p_i = {i+1, i+2, 0} over indices (0, 1, 3, 4)
    want = [pos[list(t)] for t in strip_idx] + [pos[list(t)] for t in fan_idx] + [inter[:, :3]]
    assert sc.n_triangles == 6                         # the line primitive adds nothing
    assert np.array_equal(sc.tri_verts[:, :, :3], np.array(want, F))
    uv = uv16.astype(np.float64) / 65535.0
    uv[:, 1] = 1.0 - uv[:, 1]
    for k, t in enumerate(strip_idx):
        assert np.allclose(sc.tri_uvs[k, :6].reshape(3, 2), uv[list(t)], atol=1e-7)
    assert np.allclose(sc.tri_normals[5, :, :3], [0, 0, -1])


def test_materials_extensions_and_textures(tmp_path):
    """Factors and KHR extensions into the Material fields; textures decoded
    to RGB8, named *k (embedded) or by file name, loaded once each in the
    order material.cpp asks for them, concatenated into the atlas; texture
    slots resolve to those indices; KHR_texture_transform into the uv block."""
    rng = np.random.default_rng(3)
    img_a = rng.integers(0, 256, size=(4, 6, 3), dtype=np.uint8)
    img_b = rng.integers(0, 256, size=(3, 2, 4), dtype=np.uint8)   # RGBA: alpha dropped
    img_c = rng.integers(0, 256, size=(2, 5, 3), dtype=np.uint8)
    (tmp_path / "side.png").write_bytes(png_bytes(img_c))
    w = Writer()
    ta, tb = w.image_embedded(png_bytes(img_a)), w.image_embedded(png_bytes(img_b, "RGBA"))
    tc = w.image_file("side.png")
    m0 = w.material(pbrMetallicRoughness={"baseColorTexture": {"index": tb}, "metallicFactor": 0.25,
                                          "metallicRoughnessTexture": {"index": tc}},
                    normalTexture={"index": ta}, emissiveFactor=[1.0, 0.5, 0.25],
                    extensions={"KHR_materials_emissive_strength": {"emissiveStrength": 4.0},
                                "KHR_materials_transmission": {"transmissionFactor": 0.75},
                                "KHR_materials_ior": {"ior": 1.33},
                                "KHR_materials_volume": {"thicknessFactor": 0.1, "attenuationDistance": 2.5,
                                                         "attenuationColor": [0.9, 0.8, 0.7]}})
    m1 = w.material(pbrMetallicRoughness={"baseColorTexture": {
        "index": ta, "extensions": {"KHR_texture_transform": {"offset": [0.1, 0.2], "rotation": 0.3,
                                                              "scale": [2.0, 3.0]}}}})
    uvs = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], F)
    w.node(mesh=_tri_mesh(w, QUAD, uvs=uvs, indices=np.array([0, 1, 2], np.uint16), material=m0))
    w.node(mesh=_tri_mesh(w, QUAD, uvs=uvs, indices=np.array([0, 2, 3], np.uint16), material=m1))
    _camera_node(w)
    sc = gltf.read_gltf_scene(w.write_glb(tmp_path / "m.glb"))
    # load order: material 0: base (b = *1), normal (a = *0), metal/rough (side.png); material 1 adds nothing
    assert sc.meta["textures"] == ["*1", "*0", "side.png"]
    sizes = [(2, 3), (6, 4), (5, 2)]
    off = 0
    for k, (img, (tw, th)) in enumerate(zip((img_b[..., :3], img_a, img_c), sizes)):
        assert tuple(sc.tex_meta[k]) == (off, tw, th, 3)
        assert np.array_equal(sc.tex_atlas[off:off + img.size], img.reshape(-1))
        off += img.size
    a, b = sc.materials
    assert (a["diff_index"], a["norm_index"], a["metal_rough_index"]) == (0, 1, 2)
    assert (a["emis_index"], a["occlusion_index"], a["transmission_index"]) == (NO_TEXTURE,) * 3
    assert np.allclose(a["emissive"][:3], [4.0, 2.0, 1.0]) and a["metallic"] == F(0.25) and a["roughness"] == 1
    assert a["transmission"] == F(0.75) and a["ior"] == F(1.33) and a["thickness"] == F(0.1)
    assert a["att_dist"] == F(2.5) and np.allclose(a["att_color"], [0.9, 0.8, 0.7, 1.0])
    assert b["diff_index"] == 1
    # KHR_texture_transform as assimp stores it: rotation negated, offset moved
    blk = sc.uv_transforms[1]["diff"]
    rot = -0.3
    c, s = math.cos(-rot), math.sin(-rot)
    assert np.isclose(blk["uv_rotation"], rot)
    assert np.allclose(blk["uv_scale"], [2.0, 3.0])
    assert np.allclose(blk["uv_offset"], [0.5 * 2.0 * (-c + s + 1) + 0.1, 0.5 * 3.0 * (s + c - 1) + 1 - 3.0 - 0.2],
                       atol=1e-6)
    assert np.allclose(sc.uv_transforms[0]["diff"]["uv_scale"], [1, 1])
    # an emissive material makes its mesh a LIGHT_TEXTURED light: power π·E·area
    assert len(sc.lights) == 1 and sc.lights[0]["kind"] == LIGHT_TEXTURED
    assert np.allclose(sc.lights[0]["power"][:3], np.pi * np.array([4.0, 2.0, 1.0]) * 0.5, rtol=1e-6)
    aux = sc.lights[0]["aux"]
    assert aux[:3].view(np.uint32).tolist() == [NO_TEXTURE, 0, 1] and aux[3] == F(0.5)
    assert np.array_equal(sc.etri_verts, sc.tri_verts[:1]) and np.array_equal(sc.etri_uvs, sc.tri_uvs[:1])


def test_lights_and_camera(tmp_path):
    """KHR_lights_punctual: power = colour·intensity/683, positions and
    directions (0, 0, −1) through the node; the first camera with assimp's
    horizontal field of view and the render size rule of output.cpp:13-37."""
    w = Writer()
    w.node(mesh=_tri_mesh(w, QUAD, indices=np.array([0, 1, 2], np.uint8)))
    lp = w.light(type="point", color=[1.0, 0.5, 0.25], intensity=683.0)
    ls = w.light(type="spot", intensity=1366.0, spot={"innerConeAngle": 0.2, "outerConeAngle": 0.5})
    ld = w.light(type="directional", color=[0.1, 0.2, 0.3], intensity=6830.0)
    half = math.sqrt(0.5)
    w.node(translation=[1, 2, 3], extensions={"KHR_lights_punctual": {"light": lp}})
    w.node(translation=[0, 4, 0], rotation=[-half, 0, 0, half], extensions={"KHR_lights_punctual": {"light": ls}})
    w.node(rotation=[0, half, 0, half], extensions={"KHR_lights_punctual": {"light": ld}})
    w.node(camera=w.camera(0.8, 2.0), translation=[0, 1, 5], rotation=[0, math.sin(0.1), 0, math.cos(0.1)])
    w.node(camera=w.camera(0.3, 1.0))                      # a second camera: not the chosen one
    path = w.write_glb(tmp_path / "l.glb")
    sc = gltf.read_gltf_scene(path, width=640, height=320)
    p, s, d = sc.lights
    assert p["kind"] == 0 and np.allclose(p["position"][:3], [1, 2, 3]) and np.allclose(p["power"][:3], [1, .5, .25])
    assert s["kind"] == 1 and np.allclose(s["position"][:3], [0, 4, 0]) and np.allclose(s["power"][:3], [2, 2, 2])
    assert np.allclose(s["direction"][:3], [0, -1, 0], atol=1e-6) and np.allclose(s["aux"], [0.2, 0.5, 0, 0])
    assert d["kind"] == 2 and np.allclose(d["direction"][:3], [-1, 0, 0], atol=1e-6)
    assert np.allclose(d["power"][:3], [1.0, 2.0, 3.0], rtol=1e-6)
    assert abs(np.dot(d["tangent"][:3], d["direction"][:3])) < 1e-6
    cam = sc.camera[0]
    hfov = 2 * math.atan(math.tan(0.4) * 2.0)
    rot = 0.2                                             # quaternion angle about y
    fwd = np.array([-math.sin(rot), 0, -math.cos(rot)])
    assert np.allclose(cam["position"][:3], [0, 1, 5])
    assert sc.meta["render_size"] == (640, 320)
    # camera.cpp:25-38 with aspect 640/320 and focus distance 1
    half_w = math.tan(hfov / 2)
    assert np.isclose(np.linalg.norm(cam["horizontal"][:3]), 2 * half_w, rtol=1e-6)
    assert np.isclose(np.linalg.norm(cam["vertical"][:3]), 2 * half_w / 2.0, rtol=1e-6)
    centre = cam["lower_left_corner"][:3] + 0.5 * cam["horizontal"][:3] + 0.5 * cam["vertical"][:3]
    assert np.allclose(centre - cam["position"][:3], fwd, atol=1e-6)
    # one side given: the other from the camera's aspect (lround)
    assert gltf.read_gltf_scene(path, width=301).meta["render_size"] == (301, 151)
    assert gltf.read_gltf_scene(path, height=101).meta["render_size"] == (202, 101)


def test_gltf_with_external_and_data_uri_buffers_equals_glb(tmp_path):
    w = Writer()
    m = w.material(pbrMetallicRoughness={"baseColorFactor": [0.5, 0.5, 0.5, 1.0]}, emissiveFactor=[2.0, 2.0, 2.0])
    w.node(mesh=_tri_mesh(w, QUAD, normals=np.tile([0, 0, 1.0], (4, 1)), indices=np.array([0, 1, 2, 0, 2, 3],
                                                                                          np.uint16), material=m))
    _camera_node(w, translation=[0, 0, 4])
    a = gltf.read_gltf_scene(w.write_glb(tmp_path / "x.glb"))
    for data_uri in (False, True):
        b = gltf.read_gltf_scene(w.write_gltf(tmp_path / f"x{int(data_uri)}.gltf", data_uri=data_uri))
        for f in ("tri_verts", "tri_normals", "tri_tangents", "tri_uvs", "tri_mat_index", "etri_verts", "lights",
                  "materials", "camera"):
            assert getattr(a, f).tobytes() == getattr(b, f).tobytes(), f    # bits: aux holds NaN patterns


def test_refused_files(tmp_path):
    w = Writer()
    w.node(mesh=_tri_mesh(w, QUAD, indices=np.array([0, 1, 2], np.uint8)))
    with pytest.raises(gltf.GltfError, match="without cameras"):
        gltf.read_gltf_scene(w.write_glb(tmp_path / "nocam.glb"))
    w.node(camera=w.camera(0.8, None))
    with pytest.raises(gltf.GltfError, match="aspect_ratio must be positive"):
        gltf.read_gltf_scene(w.write_glb(tmp_path / "noaspect.glb"))
    e = Writer()
    e.node(camera=e.camera(0.8, 1.0))
    with pytest.raises(gltf.GltfError, match="empty scene"):
        gltf.read_gltf_scene(e.write_glb(tmp_path / "empty.glb"))
    o = Writer()
    o.node(mesh=_tri_mesh(o, QUAD, indices=np.array([0, 1, 2], np.uint8)))
    o.doc.setdefault("cameras", []).append({"type": "orthographic", "orthographic": {"xmag": 1, "ymag": 1,
                                                                                      "znear": 0, "zfar": 1}})
    o.node(camera=0)
    with pytest.raises(gltf.GltfError, match="orthographic"):
        gltf.read_gltf_scene(o.write_glb(tmp_path / "ortho.glb"))


def test_cornell_box_round_trip(tmp_path):
    """The procedural Cornell box written as glTF and read back: the same
    triangles (bit for bit), normals within one rounding, the same material
    per triangle, the same area light, the same camera to float rounding."""
    from phosphor_amd import scenes

    from tests.gltf_files import write_scene_glb

    ref = scenes.cornell_box(64, 64)
    path = write_scene_glb(ref, tmp_path / "cornell.glb", math.radians(45.0), 1.0, (0.5, 0.5, 2.2))
    sc = gltf.read_gltf_scene(path, width=64, height=64)
    assert np.array_equal(sc.tri_verts, ref.tri_verts)
    assert np.allclose(sc.tri_normals, ref.tri_normals, atol=1.2e-7)
    for field in ("base_color", "emissive", "metallic", "roughness", "transmission", "ior", "att_dist"):
        assert np.array_equal(sc.materials[sc.tri_mat_index][field], ref.materials[ref.tri_mat_index][field]), field
    assert sc.lights.tobytes() == ref.lights.tobytes()
    assert np.array_equal(sc.etri_verts, ref.etri_verts)
    for f in ("position", "lower_left_corner", "horizontal", "vertical"):
        assert np.allclose(sc.camera[0][f], ref.camera[0][f], atol=2e-6), f
