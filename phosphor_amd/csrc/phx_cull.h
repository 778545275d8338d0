/*
 * phx_cull.h — conservative cell culling for the photon gather, compiled
 * identically for gfx950 device code and for x86-64 host code (the CPU
 * check tests/cpp/gather_cull_check.cpp includes it).
 *
 * The gather (gather_flux.cl:24-77) visits the 27 grid cells around a
 * hitpoint (photon_hash.h:56-103) and accepts a stored photon when its
 * squared distance dsq, computed in float, is below r².  Skipping a cell
 * changes nothing as long as no photon the distance test accepts is stored
 * in it.  The test below skips only cells that lie provably farther than r:
 *
 *   - a photon hashed into cell x has f = fl(fl(p − o)/cs) in [x, x+1)
 *     (photon_hash.h:21-37), so its position lies in [o + x·cs, o + (x+1)·cs]
 *     up to the rounding of f, at most ≈ 1.2e-7·|f|·cs ≤ 1.2e-7·g·cs;
 *   - the box corners o + x·cs and the distances pos − corner are each
 *     computed with a few roundings of magnitudes at most
 *     |o| + |pos| + (g+1)·cs;
 *   - per axis the gap to the box is reduced by the margin
 *     phx_cull_margin = 1e-6·(|o| + |pos| + (g+1)·cs), more than twice the
 *     sum of those rounding bounds (≈ 4.4e-7·(…)), so the gap computed here
 *     is at most the true gap to a box that surely holds the photon;
 *   - an accepted photon has fl(dsq) < r², and dsq ≥ each partial sum of its
 *     non-negative terms, so its true distance is below r·(1 + 3e-7); the
 *     cell is skipped only when the gaps' squared sum is at least
 *     r²·(1 + 1e-5).
 * A fixed margin of a small fraction of a cell is not enough in general:
 * the rounding grows with the grid resolution and with the distance of the
 * scene from the origin, and the CPU check finds accepted photons in cells
 * such a margin skips.  NaN anywhere makes the comparisons false, so nothing
 * is skipped.
 */
#ifndef PHX_CULL_H
#define PHX_CULL_H

#include "phx_math.h"

/* bound, per axis, on the rounding of a cell box edge and of the distance to
 * it */
PHX_HD float phx_cull_margin(float origin, float pos, float cell, uint32_t grid_res) {
    return 1e-6f * (phx_fabsf(origin) + phx_fabsf(pos) + ((float)grid_res + 1.0f) * cell);
}

/* distance from pos to [lo, hi] along one axis, less the margin, at least 0
 * (NaN stays NaN, so that nothing is skipped) */
PHX_HD float phx_cull_gap(float pos, float lo, float hi, float margin) {
    float g = lo - pos > pos - hi ? lo - pos : pos - hi;
    g = g - margin;
    return (g > 0.0f || g != g) ? g : 0.0f;
}

/* true when a box with these per-axis gaps holds no photon within r */
PHX_HD bool phx_cull_far(float gx, float gy, float gz, float r_sq) {
    return gx * gx + gy * gy + gz * gz >= r_sq * 1.00001f;
}

#endif
