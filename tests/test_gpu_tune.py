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
        if knob in ("trace_mode", "reference_bvh", "bvh_leaf_size", "emit_wave_bounces", "max_photons_per_batch"):
            assert exact, (knob, value)
        if knob == "grid_res":
            assert rmse < 1e-3, (value, rmse)  # cells >= one radius find the same photons
    assert res.ms_per_pass > 0 and res.baseline_ms > 0
    accepted = [t for t in res.trials if t[5]]
    assert any(abs(t[2] - res.ms_per_pass) < 1e-9 for t in accepted)
