"""GPU parity of whole SPPM passes: the HIP pass driver against the CPU oracle.

Both sides run the same scene, arguments and PCG seed schedule
(renderer.cpp:54,65).  Every stage is deterministic on both sides (canonical
deposit order, stable scatter, IEEE arithmetic with identical elementary
functions), so hitpoints, the sorted photon map, cell ranges and the
per-pixel SPPM state must agree bit for bit after every pass.
"""
import numpy as np
import pytest

from phosphor_amd import PassConfig, PhosphorError, Renderer, scenes

import oracle

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _compare_pass(gpu, ref, tag):
    hp_g, hp_r = gpu.read_hitpoints(), ref.read_hitpoints()
    bad = np.nonzero((_bits(hp_g) != _bits(hp_r)).any(axis=(1, 2)))[0]
    assert bad.size == 0, f"{tag}: {bad.size} pixels with differing hitpoints, first {bad[:5]}"
    assert gpu.last_photon_count == ref.last_photon_count, tag
    pos_g, pw_g, dir_g, cs_g, ce_g = gpu.read_photon_map()
    pos_r, pw_r, dir_r, cs_r, ce_r = ref.read_photon_map()
    assert np.array_equal(cs_g, cs_r) and np.array_equal(ce_g, ce_r), f"{tag}: cell ranges differ"
    assert np.array_equal(_bits(pos_g), _bits(pos_r)), f"{tag}: sorted photon positions differ"
    assert np.array_equal(pw_g, pw_r) and np.array_equal(dir_g, dir_r), f"{tag}: photon power/dir differ"
    for a, b, name in zip(gpu.read_state(), ref.read_state(), ("flux", "weight", "radiance")):
        diff = np.nonzero(_bits(a).reshape(a.shape[0], -1) != _bits(b).reshape(b.shape[0], -1))[0]
        assert diff.size == 0, f"{tag}: {name} differs at {diff.size} pixels, first {diff[:5]}"


@pytest.mark.parametrize("case", ["cornell", "caustic"])
def test_passes_bit_exact(case):
    if case == "cornell":
        scene = scenes.cornell_box(96, 64)
        cfg = PassConfig(width=96, height=64, photons_per_round=1 << 14, max_photons_per_batch=5000,
                         max_photon_bounces=8, grid_res=43, seed=1, initial_radius=0.02)
    else:
        scene = scenes.caustic_scene(64, 64, level=3)
        cfg = PassConfig(width=64, height=64, photons_per_round=1 << 14, max_photons_per_batch=1 << 14,
                         max_photon_bounces=16, grid_res=87, seed=7, initial_radius=0.01)
    print(f"case {case}: {scene.n_triangles} triangles", flush=True)
    gpu = Renderer(cfg, scene)
    ref = oracle.RefRenderer(cfg, scene)
    for p in range(3):
        gpu.run_pass()
        gpu.synchronize()
        assert ref.run_pass() == 0
        _compare_pass(gpu, ref, f"{case} pass {p}")
    est_g, est_r = gpu.estimate(), ref.estimate()
    assert np.array_equal(_bits(est_g), _bits(est_r))
    gpu.close()


def _near_first_cases():
    return [
        ("caustic level 5", lambda: scenes.caustic_scene(128, 96, level=5), 128, 96, 87, 0.01),
        ("atrium 60k", lambda: scenes.atrium_scene(128, 72, target_tris=60_000, env_size=(256, 128)), 128, 72, 64,
         0.0),
        ("sphere room 40k", lambda: scenes.sphere_room(112, 80, 40_000, np.random.default_rng(3)), 112, 80, 64, 0.0),
    ]


@pytest.mark.parametrize("index,leaf", [(0, 1), (1, 1), (2, 1), (0, 2), (0, 4), (1, 4), (2, 3)])
def test_sah_tree_matches_reference_tree(index, leaf):
    """Where the reference's median tree is no deeper than 22 interior levels
    the library traces on its own SAH tree (leaves of up to `leaf`
    triangles), visiting nodes near-first and breaking equal-distance ties by
    the reference tree's leaf order; PHX_REFERENCE_BVH traces on the
    reference tree depth-first.  Hitpoints, photon map and SPPM state are
    identical."""
    name, make, w, h, grid, radius = _near_first_cases()[index]
    scene = make()
    base = PassConfig(width=w, height=h, photons_per_round=1 << 16,
                      max_photons_per_batch=1 << 16, max_photon_bounces=16, grid_res=grid, seed=11,
                      initial_radius=radius, bvh_leaf_size=leaf)
    print(f"case {name}, leaf size {leaf}: {scene.n_triangles} triangles", flush=True)
    near = Renderer(base, scene)
    dfs = Renderer(PassConfig(**{**base.__dict__, "reference_bvh": True}), scene)
    for p in range(2):
        near.run_pass()
        dfs.run_pass()
        near.synchronize()
        dfs.synchronize()
        _compare_pass(near, dfs, f"{name} pass {p}")
    near.close()
    dfs.close()


def _mixed_light_scene(w, h):
    """Cornell room lit by every light kind: the quad light, a 12-triangle
    emissive box, a point light, a spot light and a directional light, so each
    kind's light-sample draw count is exercised."""
    from phosphor_amd import abi
    from phosphor_amd.scenes import SceneBuilder, default_material, make_camera, _cornell_room

    b = SceneBuilder()
    white = b.material(default_material((0.75, 0.75, 0.75)))
    red = b.material(default_material((0.63, 0.06, 0.06)))
    green = b.material(default_material((0.14, 0.45, 0.14)))
    quad = b.material(default_material((1.0, 1.0, 1.0), emissive=(9.0, 9.0, 9.0)))
    glow = b.material(default_material((1.0, 1.0, 1.0), emissive=(2.0, 4.0, 1.0)))
    _cornell_room(b, white, red, green, quad)
    b.add_box((0.7, 0.05, 0.1), (0.8, 0.15, 0.2), glow)
    b.add_point_light((0.3, 0.8, 0.6), (0.8, 0.7, 0.5))
    b.add_directional_light((0.2, -0.9, -0.3), (0.4, 0.4, 0.5))
    spot = np.zeros(1, abi.LIGHT_DTYPE)[0]
    spot["kind"] = abi.LIGHT_SPOT
    spot["position"][:3] = (0.6, 0.9, 0.7)
    spot["direction"][:3] = (0.0, -1.0, 0.0)
    spot["power"][:3] = (1.5, 1.2, 1.0)
    spot["aux"][:2] = (0.3, 0.5)
    b.lights.append(spot)
    cam = make_camera((0.5, 0.5, 2.2), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0), 45.0, w / h)
    return b.build(cam, name="mixed_lights")


def test_mixed_light_kinds_bit_exact():
    """Light samples are drawn by the shadow kernel from PCG states the camera
    kernel records after stepping past each sample's draws: every light kind
    (mesh of 2 and of 12 triangles, point, spot, directional) keeps the
    camera path on the reference's random stream."""
    scene = _mixed_light_scene(64, 48)
    cfg = PassConfig(width=64, height=48, photons_per_round=1 << 13, max_photons_per_batch=1 << 13,
                     max_photon_bounces=8, grid_res=43, seed=3, initial_radius=0.02, direct_samples=7)
    gpu = Renderer(cfg, scene)
    ref = oracle.RefRenderer(cfg, scene)
    for p in range(2):
        gpu.run_pass()
        gpu.synchronize()
        assert ref.run_pass() == 0
        _compare_pass(gpu, ref, f"mixed lights pass {p}")
    gpu.close()


def test_reference_radius_rule_and_determinism():
    """initial_radius 0 → the reference's cell half-diagonal rule (gather_flux.cl:51-56);
    two renders with the same seed are identical."""
    scene = scenes.cornell_box(48, 48)
    cfg = PassConfig(width=48, height=48, photons_per_round=1 << 13, max_photons_per_batch=1 << 13,
                     max_photon_bounces=8, grid_res=64, seed=2137)
    a, b = Renderer(cfg, scene), Renderer(cfg, scene)
    ref = oracle.RefRenderer(cfg, scene)
    for _ in range(2):
        a.run_pass()
        b.run_pass()
        ref.run_pass()
    fa, wa, ra = a.read_state()
    fb, wb, rb = b.read_state()
    fr, wr, rr = ref.read_state()
    assert np.array_equal(_bits(fa), _bits(fb)) and np.array_equal(_bits(ra), _bits(rb))
    assert np.array_equal(_bits(fa), _bits(fr)) and np.array_equal(_bits(wa), _bits(wr))
    hit = wa > 0
    assert hit.any()


def test_graph_replay_bit_exact():
    """PHX_PASS_GRAPH: the pass is captured once and replayed; every replay,
    with host work and synchronisations between replays, equals the oracle."""
    scene = scenes.cornell_box(80, 56)
    cfg = PassConfig(width=80, height=56, photons_per_round=1 << 13, max_photons_per_batch=3000,
                     max_photon_bounces=8, grid_res=43, seed=5, initial_radius=0.02, graph=True)
    gpu = Renderer(cfg, scene)
    ref = oracle.RefRenderer(cfg, scene)
    for p in range(4):
        gpu.run_pass()
        gpu.synchronize()
        assert gpu.pass_ms() > 0.0
        assert ref.run_pass() == 0
        _compare_pass(gpu, ref, f"graph pass {p}")
    gpu.close()


def test_checkpoint_resume_bit_exact():
    scene = scenes.caustic_scene(48, 40, level=2)
    cfg = PassConfig(width=48, height=40, photons_per_round=1 << 12, max_photons_per_batch=1 << 12,
                     max_photon_bounces=12, grid_res=60, seed=9, initial_radius=0.015)
    a = Renderer(cfg, scene)
    a.run(2)
    ckpt = a.checkpoint()
    a.run(2)
    b = Renderer(cfg, scene)
    b.restore(ckpt)
    assert b.rounds_done == 2
    b.run(2)
    for x, y in zip(a.read_state(), b.read_state()):
        assert np.array_equal(_bits(x), _bits(y))
    assert np.array_equal(_bits(a.estimate()), _bits(b.estimate()))
    other = Renderer(PassConfig(width=40, height=40, photons_per_round=1 << 10, max_photon_bounces=4, grid_res=20),
                     scene)
    with pytest.raises(Exception):
        other.restore(ckpt)
    for r in (a, b, other):
        r.close()


def test_pass_without_deposits_is_reported():
    """All-mirror scene: no diffuse event, no deposit; renderer.cpp:117-118 is
    reported by the synchronisation after the pass."""
    from phosphor_amd import PhosphorError, abi

    scene = scenes.cornell_box(32, 32)
    scene.materials["metallic"] = 1.0
    cfg = PassConfig(width=32, height=32, photons_per_round=1 << 10, max_photons_per_batch=1 << 10,
                     max_photon_bounces=4, grid_res=20, seed=3)
    r = Renderer(cfg, scene)
    r.run_pass()
    with pytest.raises(PhosphorError) as ei:
        r.synchronize()
    assert ei.value.code == abi.PHX_ERR_NO_PHOTONS
    assert r.last_photon_count == 0
    r.synchronize()  # reported once
    r.close()


def _floor_and_light(w, h):
    """Four triangles: a floor quad under a quad light, so a tree with leaves
    of four triangles is a single leaf and one of two triangles a single
    pair."""
    from phosphor_amd.scenes import SceneBuilder, default_material, make_camera

    b = SceneBuilder()
    white = b.material(default_material((0.75, 0.75, 0.75)))
    light = b.material(default_material((1.0, 1.0, 1.0), emissive=(17.0, 17.0, 17.0)))
    b.add_quad((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (1.0, 0.0, 1.0), (1.0, 0.0, 0.0), white)  # facing up
    b.add_quad((0.4, 0.9, 0.4), (0.6, 0.9, 0.4), (0.6, 0.9, 0.6), (0.4, 0.9, 0.6), light)  # facing down
    cam = make_camera((0.5, 0.6, 2.0), (0.5, 0.2, 0.5), (0.0, 1.0, 0.0), 45.0, w / h)
    return b.build(cam, name="floor_and_light")


@pytest.mark.parametrize("leaf", [1, 2, 4])
def test_tiny_trees_bit_exact(leaf):
    """A scene whose SAH tree is one leaf (four triangles, leaf size 4), one
    pair of leaves (leaf size 2) or three pairs (leaf size 1) renders the
    oracle's bits."""
    scene = _floor_and_light(32, 24)
    assert scene.n_triangles == 4
    cfg = PassConfig(width=32, height=24, photons_per_round=1 << 12, max_photons_per_batch=1 << 12,
                     max_photon_bounces=8, grid_res=16, seed=5, initial_radius=0.05, bvh_leaf_size=leaf)
    print(f"case floor and light, leaf size {leaf}", flush=True)
    gpu = Renderer(cfg, scene)
    ref = oracle.RefRenderer(cfg, scene)
    for p in range(2):
        gpu.run_pass()
        gpu.synchronize()
        assert ref.run_pass() == 0
        _compare_pass(gpu, ref, f"floor and light, leaf {leaf}, pass {p}")
    gpu.close()


def test_bad_leaf_size_rejected():
    with pytest.raises(PhosphorError, match="bvh_leaf_size"):
        Renderer(PassConfig(width=8, height=8, photons_per_round=64, grid_res=8, bvh_leaf_size=5),
                 scenes.cornell_box(8, 8))


@pytest.mark.parametrize("leaf", [1, 3])
def test_textures_media_and_light_kinds_bit_exact(leaf):
    """A textured floor (bilinear texture with sRGB decoding), an attenuating
    glass box with thickness (Beer-Lambert on shadow rays, camera and photon
    paths) and point, spot and directional lights: the features whose oracle
    parts tests/test_oracle_lights_media.py pins by closed forms, rendered on
    the GPU with the oracle's bits."""
    from tests.test_oracle_lights_media import scene_with_everything

    scene = scene_with_everything()
    cfg = PassConfig(width=48, height=40, photons_per_round=1 << 13, max_photons_per_batch=1 << 13,
                     max_photon_bounces=8, grid_res=24, seed=4, initial_radius=0.04, direct_samples=5,
                     bvh_leaf_size=leaf)
    print(f"case textures/media/lights, leaf size {leaf}", flush=True)
    gpu = Renderer(cfg, scene)
    ref = oracle.RefRenderer(cfg, scene)
    for p in range(2):
        gpu.run_pass()
        gpu.synchronize()
        assert ref.run_pass() == 0
        _compare_pass(gpu, ref, f"textures/media/lights leaf {leaf} pass {p}")
    gpu.close()


@pytest.mark.parametrize("mode", ["pipelined", "serial", "graph"])
def test_back_to_back_passes_bit_exact(mode):
    """Passes enqueued with no host synchronisation between them.  Serial and
    graph passes: the camera trace and the emission run on two streams, joined
    before the gather, and the next pass starts after that gather.  Pipelined
    passes (the default): each gather runs on a third stream beside the next
    pass's trace and emission, on alternating hitpoint / photon-map sets.  The
    state after five passes, and again after two more, is the oracle's."""
    scene = scenes.caustic_scene(64, 48, level=3)
    cfg = PassConfig(width=64, height=48, photons_per_round=1 << 14, max_photons_per_batch=1 << 14,
                     max_photon_bounces=12, grid_res=60, seed=13, initial_radius=0.015, graph=mode == "graph",
                     serial=mode == "serial")
    print(f"case back-to-back, {mode}", flush=True)
    gpu = Renderer(cfg, scene)
    ref = oracle.RefRenderer(cfg, scene)
    for rounds in (5, 2):
        gpu.run(rounds)
        gpu.synchronize()
        for _ in range(rounds):
            assert ref.run_pass() == 0
        _compare_pass(gpu, ref, f"back-to-back {mode}")
    gpu.close()
