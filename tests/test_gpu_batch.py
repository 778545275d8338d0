"""Multi-scene passes: every scene of a batch is bit-identical to rendering it
alone with the one-scene pass driver (same arguments, same seeds), after every
pass: hitpoints, per-scene photon map and cell ranges, SPPM state, estimate.

The scenes differ in size (image widths that are not multiples of 8), bounce
limits (below, at and above the bounce-synchronous launches), grids, NEE sample
counts, radius rule, any-hit eligibility and envmap; group budgets are small so
trace and emission run in several groups."""
import numpy as np
import pytest

from phosphor_amd import Batch, PassConfig, Renderer, scenes

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _cases():
    out = []
    out.append((scenes.cornell_box(48, 40), PassConfig(width=48, height=40, photons_per_round=4096,
                                                       max_photon_bounces=8, grid_res=43, seed=1,
                                                       initial_radius=0.02)))
    out.append((scenes.caustic_scene(64, 48, level=2), PassConfig(width=64, height=48, photons_per_round=6000,
                                                                  max_photon_bounces=12, grid_res=60, seed=2,
                                                                  direct_samples=8, initial_radius=0.015)))
    out.append((scenes.cornell_box(33, 17), PassConfig(width=33, height=17, photons_per_round=1500,
                                                       max_photon_bounces=3, grid_res=20, seed=3,
                                                       direct_samples=4)))
    out.append((scenes.caustic_scene(40, 40, level=1), PassConfig(width=40, height=40, photons_per_round=3000,
                                                                  max_photon_bounces=1, grid_res=30, seed=4,
                                                                  initial_radius=0.02)))
    out.append((scenes.atrium_scene(56, 32, target_tris=4000, env_size=(128, 64)),
                PassConfig(width=56, height=32, photons_per_round=5000, max_photon_bounces=16, grid_res=64, seed=5,
                           direct_samples=2)))
    return out


@pytest.mark.parametrize("mode", ["graph", "pipelined", "serial"])
def test_batch_matches_single_scene_passes(mode):
    cases = _cases()
    print(f"case batch {mode}", flush=True)
    batch = Batch([c for _, c in cases], [s for s, _ in cases], trace_group_pixels=2500,
                  emit_group_photons=7000, graph=mode == "graph", serial=mode == "serial")
    singles = [Renderer(c, s) for s, c in cases]
    for p in range(3):
        batch.run_pass()
        batch.synchronize()
        counts = batch.photon_counts()
        for i, r in enumerate(singles):
            r.run_pass()
            r.synchronize()
            tag = f"pass {p} scene {i}"
            assert counts[i] == r.last_photon_count, tag
            hb, hs = batch.read_hitpoints(i), r.read_hitpoints()
            bad = np.nonzero((_bits(hb) != _bits(hs)).any(axis=(1, 2)))[0]
            assert bad.size == 0, f"{tag}: {bad.size} hitpoints differ, first {bad[:5]}"
            for a, b, name in zip(batch.read_photon_map(i), r.read_photon_map(), ("pos", "power", "dir", "cs", "ce")):
                assert np.array_equal(_bits(a), _bits(b)), f"{tag}: photon map {name} differs"
            for a, b, name in zip(batch.read_state(i), r.read_state(), ("flux", "weight", "radiance")):
                assert np.array_equal(_bits(a), _bits(b)), f"{tag}: {name} differs"
    for i, r in enumerate(singles):
        assert np.array_equal(_bits(batch.estimate(i)), _bits(r.estimate()))
        r.close()
    assert batch.rounds_done == 3
    assert batch.pass_ms() > 0.0
    batch.close()


@pytest.mark.parametrize("mode", ["graph", "pipelined"])
def test_batch_back_to_back_passes(mode):
    """Batched passes enqueued with no host synchronisation between them
    (pipelined: each gather beside the next pass's trace and emission, on
    alternating hitpoint / photon-map sets): after four passes every scene's
    hitpoints, photon map and state are those of its own renderer."""
    cases = _cases()
    print(f"case batch back-to-back {mode}", flush=True)
    batch = Batch([c for _, c in cases], [s for s, _ in cases], trace_group_pixels=2500,
                  emit_group_photons=7000, graph=mode == "graph")
    for _ in range(4):
        batch.run_pass()
    batch.synchronize()
    counts = batch.photon_counts()
    for i, (s, c) in enumerate(cases):
        r = Renderer(c, s)
        r.run(4)
        tag = f"scene {i}"
        assert counts[i] == r.last_photon_count, tag
        assert np.array_equal(_bits(batch.read_hitpoints(i)), _bits(r.read_hitpoints())), f"{tag}: hitpoints"
        for a, b, name in zip(batch.read_photon_map(i), r.read_photon_map(), ("pos", "power", "dir", "cs", "ce")):
            assert np.array_equal(_bits(a), _bits(b)), f"{tag}: photon map {name} differs"
        for a, b, name in zip(batch.read_state(i), r.read_state(), ("flux", "weight", "radiance")):
            assert np.array_equal(_bits(a), _bits(b)), f"{tag}: {name} differs"
        r.close()
    assert batch.rounds_done == 4
    batch.close()
