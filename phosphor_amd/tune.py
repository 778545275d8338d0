"""Autotuner of the pass driver's settings for one scene and configuration.

Two kinds of knobs:

* exact knobs change how the pass is launched, never its results: trace form
  (wavefront or one kernel), BVH (SAH tree visited near-first, or the
  reference's median tree visited depth-first), camera paths continued in
  a second kernel and photon bounces traced with ray replacement and shaded
  apart, or both fused (PHX_FUSED_TRACE), the SAH
  tree's leaf size (1 to 4 triangles; the reference's is 1, bvh.cpp:24),
  emission batch size
  (``--batch-size``, cmd_args.hpp:42), the number of bounce-synchronous
  emission launches, and the launch shapes (the reference's ``--workgroup``,
  cmd_args.hpp:44, per kernel): threads per block of the camera, shadow and
  emission kernels (64, 128, 256) and the width of the gather's 64-pixel
  wave tile (1 … 64 pixels).
  A candidate is accepted only if its SPPM state after the measured passes is
  bit-identical to the state under the given settings.
* the grid resolution (``--grid-res``, cmd_args.hpp:51) changes the hash
  keys; with an explicit initial radius a grid whose cells are at least one
  radius wide finds the same photons, so only the order of the gather's sums
  changes.  Such a candidate is accepted when the estimate's relative RMSE
  against the given settings is within ``rmse_tol``.  With the reference
  radius rule (initial_radius 0) the grid sets the radius, so it is not
  tuned.

The search is coordinate-wise (one knob at a time, keeping the best so far),
each candidate timed over ``passes`` passes after ``warmup`` passes with the
renderer's own pass events.
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np

from . import PassConfig, Renderer
from .output import relative_rmse


def _run(cfg: PassConfig, scene, warmup: int, passes: int):
    """Milliseconds per pass of `passes` passes enqueued back to back (as a
    render runs them: pipelined unless the config says otherwise), after
    `warmup` passes; the state and the estimate at the end."""
    r = Renderer(cfg, scene)
    for _ in range(warmup):
        r.run_pass()
    r.synchronize()
    t0 = time.perf_counter()
    for _ in range(passes):
        r.run_pass()
    r.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / max(passes, 1)
    flux, weight, rad = r.read_state()
    est = r.estimate()
    r.close()
    return ms, (flux, weight, rad), est


def _same_state(a, b) -> bool:
    return all(np.array_equal(np.ascontiguousarray(x).view(np.uint32), np.ascontiguousarray(y).view(np.uint32))
               for x, y in zip(a, b))


def grid_for_cell(extent: float, cell: float) -> int:
    return max(1, min(255, int(math.floor(extent / cell))))


@dataclasses.dataclass
class TuneResult:
    config: PassConfig
    ms_per_pass: float
    baseline_ms: float
    trials: list

    @property
    def speedup(self) -> float:
        return self.baseline_ms / self.ms_per_pass


def autotune(cfg: PassConfig, scene, passes: int = 3, warmup: int = 1, wave_bounces=(1, 2, 3, 4, 6),
             batch_sizes=None, trace_modes=(0, 1), leaf_sizes=(1, 2, 3, 4), cell_factors=(1.0, 1.155, 1.5, 2.0),
             block_sizes=(64, 128, 256), tile_widths=(4, 8, 16, 64), rmse_tol: float = 1e-4) -> TuneResult:
    """Returns the fastest accepted configuration and every trial
    (knob, value, ms per pass, exact, relative RMSE, accepted)."""
    base_ms, base_state, base_est = _run(cfg, scene, warmup, passes)
    best, best_ms = dataclasses.replace(cfg), base_ms
    trials = [("baseline", None, base_ms, True, 0.0, True)]

    def consider(knob, values, exact):
        nonlocal best, best_ms
        for v in values:
            cand = dataclasses.replace(best, **{knob: v})
            if cand == best:
                continue
            ms, state, est = _run(cand, scene, warmup, passes)
            rm = relative_rmse(est, base_est)
            same = _same_state(state, base_state)
            ok = same if exact else rm <= rmse_tol
            trials.append((knob, v, ms, same, rm, ok))
            if ok and ms < best_ms:
                best, best_ms = cand, ms

    consider("trace_mode", trace_modes, True)
    consider("reference_bvh", (False, True), True)
    if not best.reference_bvh:
        consider("fused_trace", (True,), True)
    if not best.reference_bvh:
        consider("bvh_leaf_size", [v for v in leaf_sizes if v != (best.bvh_leaf_size or 3)], True)
    consider("emit_wave_bounces", [w for w in wave_bounces if w <= cfg.max_photon_bounces], True)
    if batch_sizes is None:
        batch_sizes = sorted({min(cfg.photons_per_round, b) for b in (1 << 18, 1 << 20, cfg.photons_per_round)})
    consider("max_photons_per_batch", batch_sizes, True)
    if best.trace_mode == 0:
        consider("camera_block", [b for b in block_sizes if b != (best.camera_block or 128)], True)
        consider("shadow_block", [b for b in block_sizes if b != (best.shadow_block or 128)], True)
    consider("emit_block", [b for b in block_sizes if b != (best.emit_block or 128)], True)
    consider("gather_tile_width", [t for t in tile_widths if t != (best.gather_tile_width or 8)], True)
    if cfg.initial_radius > 0.0:
        lo, hi = scene.bounds()
        extent = float((hi - lo).max()) + 2e-3
        grids = [grid_for_cell(extent, f * cfg.initial_radius) for f in cell_factors]
        consider("grid_res", sorted(set(grids)), False)
    return TuneResult(best, best_ms, base_ms, trials)


def main(argv=None) -> int:
    """python -m phosphor_amd.tune --workload caustic: tunes a named workload
    and prints every trial and the chosen configuration."""
    import argparse

    from . import config_for

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--workload", default="caustic")
    ap.add_argument("--passes", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    args = ap.parse_args(argv)
    cfg, scene, _ = config_for(args.workload)
    res = autotune(cfg, scene, passes=args.passes, warmup=args.warmup, wave_bounces=(1, 2, 3, 4, 5, 6, 8))
    for t in res.trials:
        print(t, flush=True)
    print("best", res.config, res.ms_per_pass, "baseline", res.baseline_ms, "speedup", res.speedup)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
