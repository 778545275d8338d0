"""The autotuner on a small scene: every exact knob keeps the SPPM state
bit-identical, grid candidates are compared by relative RMSE, and the chosen
configuration is one of the accepted trials."""
import pytest

from phosphor_amd import PassConfig, scenes
from phosphor_amd.tune import autotune

pytestmark = pytest.mark.gpu


def test_autotune_small_scene():
    scene = scenes.caustic_scene(96, 80, level=3)
    cfg = PassConfig(width=96, height=80, photons_per_round=1 << 15, max_photons_per_batch=1 << 15,
                     max_photon_bounces=8, grid_res=87, seed=2, initial_radius=0.01)
    res = autotune(cfg, scene, passes=2, warmup=1, wave_bounces=(1, 2, 4), cell_factors=(1.155, 2.0))
    for knob, value, ms, exact, rmse, ok in res.trials:
        print(knob, value, f"{ms:.3f} ms", "exact" if exact else f"rmse {rmse:.2e}", "accepted" if ok else "rejected")
        if knob in ("trace_mode", "reference_bvh", "fused_trace", "bvh_leaf_size", "emit_wave_bounces",
                    "max_photons_per_batch", "camera_block", "shadow_block", "emit_block", "gather_tile_width"):
            assert exact, (knob, value)
        if knob == "grid_res":
            assert rmse < 1e-3, (value, rmse)  # cells >= one radius find the same photons
    tried = {t[0] for t in res.trials}
    assert {"emit_block", "gather_tile_width", "fused_trace"} <= tried
    assert res.ms_per_pass > 0 and res.baseline_ms > 0
    accepted = [t for t in res.trials if t[5]]
    assert any(abs(t[2] - res.ms_per_pass) < 1e-9 for t in accepted)


@pytest.mark.parametrize("case", ["point_light", "caustic"])
def test_launch_shapes_bit_exact(case):
    """Every block size of the camera, shadow and emission kernels, several
    gather tiles, fused photon bounces (PHX_FUSED_TRACE) and split bounce
    counts give the oracle's pass bit for bit, on a scene lit by one point
    light (one shadow slot per event) and on the caustic scene (area light,
    one slot per sample)."""
    import oracle
    from phosphor_amd import Renderer
    from tests.test_gpu_pass import _compare_pass

    if case == "point_light":
        b = scenes.SceneBuilder()
        white = b.material(scenes.default_material((0.7, 0.7, 0.7)))
        red = b.material(scenes.default_material((0.8, 0.2, 0.2)))
        green = b.material(scenes.default_material((0.2, 0.8, 0.2)))
        scenes._cornell_room(b, white, red, green, white, boxes=True)
        b.lights.clear()                                   # the ceiling quad is not emissive here
        b.add_point_light((0.5, 0.9, 0.5), (3.0, 3.0, 3.0))
        sc = b.build(scenes.make_camera((0.5, 0.5, 2.2), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0), 45.0, 96 / 64))
        assert len(sc.lights) == 1
        base = dict(width=96, height=64, photons_per_round=1 << 14, max_photons_per_batch=1 << 14,
                    max_photon_bounces=8, grid_res=43, seed=6, initial_radius=0.02)
    else:
        sc = scenes.caustic_scene(64, 48, level=3)
        base = dict(width=64, height=48, photons_per_round=1 << 14, max_photons_per_batch=1 << 14,
                    max_photon_bounces=12, grid_res=60, seed=8, initial_radius=0.012)
    shapes = [dict(camera_block=64, shadow_block=256, emit_block=64, gather_tile_width=16),
              dict(camera_block=256, shadow_block=64, emit_block=256, gather_tile_width=64),
              dict(gather_tile_width=1), dict(gather_tile_width=4),
              dict(fused_trace=True), dict(fused_trace=True, emit_block=64, emit_wave_bounces=3),
              dict(emit_wave_bounces=1), dict(emit_wave_bounces=12, emit_block=256)]
    ref = oracle.RefRenderer(PassConfig(**base), sc)
    assert ref.run_pass() == 0
    for shape in shapes:
        gpu = Renderer(PassConfig(**base, **shape), sc)
        gpu.run_pass()
        gpu.synchronize()
        print(case, shape, flush=True)
        _compare_pass(gpu, ref, f"{case} {shape}")
        gpu.close()

