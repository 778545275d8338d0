"""Scenes imported from glTF (phosphor_amd.gltf) rendered by the HIP pass
driver against the oracle, bit for bit after every pass: the Cornell box
written as .glb, and a textured scene (base-colour texture under
KHR_texture_transform, an emissive mesh whose emission comes from a texture,
a spot light from KHR_lights_punctual, a transformed camera)."""
import math

import numpy as np
import pytest

from phosphor_amd import PassConfig, Renderer, gltf, scenes

import oracle

from tests.gltf_files import Writer, png_bytes, write_scene_glb
from tests.test_gpu_pass import _bits, _compare_pass

pytestmark = pytest.mark.gpu


def _textured_glb(path):
    rng = np.random.default_rng(11)
    w = Writer()
    tex = w.image_embedded(png_bytes(rng.integers(30, 255, size=(16, 16, 3), dtype=np.uint8)))
    glow = w.image_embedded(png_bytes(rng.integers(100, 255, size=(4, 4, 3), dtype=np.uint8)))
    floor_m = w.material(pbrMetallicRoughness={"metallicFactor": 0.0, "roughnessFactor": 0.8, "baseColorTexture": {
        "index": tex, "extensions": {"KHR_texture_transform": {"scale": [3.0, 2.0], "rotation": 0.4,
                                                               "offset": [0.1, 0.0]}}}})
    wall_m = w.material(pbrMetallicRoughness={"baseColorFactor": [0.6, 0.5, 0.4, 1.0], "metallicFactor": 0.0})
    lamp_m = w.material(pbrMetallicRoughness={"metallicFactor": 0.0}, emissiveFactor=[1.0, 0.9, 0.8],
                        emissiveTexture={"index": glow},
                        extensions={"KHR_materials_emissive_strength": {"emissiveStrength": 12.0}})
    quad = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    idx = np.array([0, 2, 1, 0, 3, 2], np.uint16)                    # facing +y

    def mesh(m):
        return w.mesh([{"attributes": {"POSITION": w.accessor(quad), "TEXCOORD_0": w.accessor(uv),
                                       "NORMAL": w.accessor(np.tile([0, 1, 0], (4, 1)).astype(np.float32))},
                        "indices": w.accessor(idx), "material": m}])

    w.node(mesh=mesh(floor_m), scale=[2.0, 1.0, 2.0])
    half = math.sqrt(0.5)
    w.node(mesh=mesh(wall_m), translation=[0, 1, -2], rotation=[half, 0, 0, half])     # back wall, facing +z
    w.node(mesh=mesh(lamp_m), translation=[0.5, 1.8, 0], rotation=[1, 0, 0, 0], scale=[0.2, 1, 0.2])  # facing -y
    spot = w.light(type="spot", color=[1.0, 0.8, 0.6], intensity=20000.0,
                   spot={"innerConeAngle": 0.3, "outerConeAngle": 0.6})
    w.node(translation=[-0.8, 1.5, 0.5], rotation=[-half, 0, 0, half], extensions={"KHR_lights_punctual": {"light": spot}})
    w.node(camera=w.camera(0.9, 4.0 / 3.0), translation=[0, 1.0, 3.5],
           rotation=[-math.sin(0.1), 0, 0, math.cos(0.1)])
    return w.write_glb(path)


@pytest.mark.parametrize("case", ["cornell", "textured"])
def test_imported_scene_passes_bit_exact(case, tmp_path):
    if case == "cornell":
        path = write_scene_glb(scenes.cornell_box(64, 64), tmp_path / "c.glb", math.radians(45.0), 1.0,
                               (0.5, 0.5, 2.2))
        sc = gltf.read_gltf_scene(path, width=80, height=80)
        cfg = PassConfig(width=80, height=80, photons_per_round=1 << 14, max_photons_per_batch=1 << 14,
                         max_photon_bounces=8, grid_res=43, seed=2, initial_radius=0.02)
    else:
        sc = gltf.read_gltf_scene(_textured_glb(tmp_path / "t.glb"), width=96)
        assert sc.meta["render_size"] == (96, 72) and len(sc.tex_meta) == 2
        cfg = PassConfig(width=96, height=72, photons_per_round=1 << 14, max_photons_per_batch=1 << 14,
                         max_photon_bounces=8, grid_res=50, seed=4, initial_radius=0.03)
    print(f"case gltf {case}: {sc.n_triangles} triangles, {len(sc.lights)} lights", flush=True)
    gpu = Renderer(cfg, sc)
    ref = oracle.RefRenderer(cfg, sc)
    for p in range(2):
        gpu.run_pass()
        gpu.synchronize()
        assert ref.run_pass() == 0
        _compare_pass(gpu, ref, f"gltf {case} pass {p}")
    assert gpu.last_photon_count > 0
    est_g, est_r = gpu.estimate(), ref.estimate()
    assert np.array_equal(_bits(est_g), _bits(est_r)) and float(np.abs(est_r).sum()) > 0
    gpu.close()
