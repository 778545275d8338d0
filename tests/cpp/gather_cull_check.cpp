/*
 * CPU check of the gather's conservative cell culling (phosphor_amd/csrc/
 * phx_cull.h): no photon that the gather's float distance test accepts may
 * lie in a cell that the culling skips.
 *
 * Random grids (resolution 4 … 4096, cells from 1e-4 to 512 units, origins
 * anywhere in ±1024), photons placed near cell edges, hitpoints at distance
 * ≈ r (just inside and just outside) in random and in axis directions.  Each
 * photon is hashed with get_hash's arithmetic (photon_hash.h:21-37), the
 * distance is computed as the gather computes it (dd = pos − p per
 * component, ((dx² + dy²) + dz²) + dw²), and for every accepted photon the
 * cell test (phx_cull_far on the cell box with phx_cull_margin) must keep
 * it, and so must the octant test of the gather's octant sort
 * (pass_bodies.h octant_mask: the half-cell box of the photon's octant, whose
 * bit per axis is fraction(f) ≥ 0.5 of get_hash_octant's quotient f, with the
 * half-cell plane l + cs·0.5 and the same margins).  Non-finite hitpoints
 * must keep every cell.
 *
 * The same trials with a fixed margin of 1e-4 cell and no relative slack
 * (mode "fixed") must find violations: that shows the trials reach the
 * rounding-sensitive cases the margin exists for.
 *
 * Usage: gather_cull_check <trials> [fixed]; prints "ok <counts>" (or, in
 * fixed mode, "fixed <violations>"); exit status 0 on the expected outcome.
 * Built with -ffp-contract=off, like the library.
 */
#include <atomic>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "../../phosphor_amd/csrc/phx_cull.h"

struct Grid {
    float o[3], cs[3];
    uint32_t g;
};

/* get_hash's cell of p (photon_hash.h:21-37) and, per axis, whether p lies
 * in the cell's high half (get_hash_octant); false outside the grid */
static bool hash_cell(const Grid &G, const float p[3], uint32_t c[3], bool hi[3]) {
    for (int a = 0; a < 3; a++) {
        float f = (p[a] - G.o[a]) / G.cs[a];
        if (f < 0.0f || f >= (float)G.g)
            return false;
        c[a] = (uint32_t)f;
        hi[a] = f - (float)c[a] >= 0.5f;
    }
    return true;
}

struct Counts {
    uint64_t trials = 0, accepted = 0, near = 0, cell_fail = 0, oct_fail = 0, nan_fail = 0;
};

static void run(uint64_t seed, uint64_t trials, bool fixed, Counts &out) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const uint32_t grids[] = {4, 16, 43, 64, 87, 128, 256, 1024, 4096};
    Counts c;
    Grid G{};
    for (uint64_t t = 0; t < trials; t++) {
        if (t % 256 == 0) { /* a new grid every 256 trials */
            G.g = grids[rng() % 9];
            double extent = std::exp(std::log(1e-4 * G.g) + U(rng) * (std::log(2048.0) - std::log(1e-4 * G.g)));
            if (extent > 2048.0)
                extent = 2048.0;
            for (int a = 0; a < 3; a++) {
                G.cs[a] = (float)(extent / G.g * (0.9 + 0.2 * U(rng)));
                double span = (double)G.cs[a] * G.g;
                double lo = -1024.0, hi = 1024.0 - span;
                G.o[a] = (float)(hi > lo ? lo + (hi - lo) * U(rng) : -span / 2);
            }
        }
        c.trials++;
        /* photon near a cell edge (or a quarter or half of the way in) */
        float p[3];
        for (int a = 0; a < 3; a++) {
            double x = (double)(rng() % G.g);
            double u;
            if (rng() % 2) {
                double e = (double)(rng() % 5) / 4.0;
                double d = std::exp(std::log(1e-9) + U(rng) * (std::log(1e-2) - std::log(1e-9)));
                u = e + ((rng() % 2) ? d : -d);
            } else {
                u = U(rng);
            }
            p[a] = (float)((double)G.o[a] + (x + u) * (double)G.cs[a]);
        }
        uint32_t cell[3];
        bool hi[3];
        if (!hash_cell(G, p, cell, hi))
            continue;
        /* radius from a thousandth of a cell to two cells */
        double cmax = std::fmax(G.cs[0], std::fmax(G.cs[1], G.cs[2]));
        double rr = cmax * std::exp(std::log(1e-3) + U(rng) * (std::log(2.0) - std::log(1e-3)));
        float r_sq = (float)(rr * rr);
        float r = std::sqrt(r_sq);
        /* hitpoint at distance ≈ r: along an axis or a random direction */
        double dir[3];
        if (rng() % 2) {
            int a = (int)(rng() % 3);
            dir[0] = dir[1] = dir[2] = 0.0;
            dir[a] = (rng() % 2) ? 1.0 : -1.0;
        } else {
            double n2 = 0.0;
            std::normal_distribution<double> N(0.0, 1.0);
            for (int a = 0; a < 3; a++) {
                dir[a] = N(rng);
                n2 += dir[a] * dir[a];
            }
            for (int a = 0; a < 3; a++)
                dir[a] /= std::sqrt(n2);
        }
        double tt = (rng() % 4) ? 1.0 + (U(rng) - 0.5) * 1e-5 : U(rng);
        float pos[3];
        for (int a = 0; a < 3; a++)
            pos[a] = (float)((double)p[a] + dir[a] * (double)r * tt);
        /* the gather's distance (dot of the float4 difference, w = 0) */
        float dx = pos[0] - p[0], dy = pos[1] - p[1], dz = pos[2] - p[2], dw = 0.0f - 0.0f;
        float dsq = ((dx * dx + dy * dy) + dz * dz) + dw * dw;
        if (!(dsq < r_sq))
            continue;
        c.accepted++;
        float l[3], m[3];
        for (int a = 0; a < 3; a++) {
            l[a] = G.o[a] + (float)cell[a] * G.cs[a];
            m[a] = fixed ? 1e-4f * G.cs[a] : phx_cull_margin(G.o[a], pos[a], G.cs[a], G.g);
        }
        float gx = phx_cull_gap(pos[0], l[0], l[0] + G.cs[0], m[0]);
        float gy = phx_cull_gap(pos[1], l[1], l[1] + G.cs[1], m[1]);
        float gz = phx_cull_gap(pos[2], l[2], l[2] + G.cs[2], m[2]);
        bool far = fixed ? (gx * gx + gy * gy + gz * gz >= r_sq) : phx_cull_far(gx, gy, gz, r_sq);
        if (gx > 0.0f || gy > 0.0f || gz > 0.0f)
            c.near++;
        if (far)
            c.cell_fail++;
        /* the octant that holds the photon, as octant_mask bounds it */
        float og[3];
        for (int a = 0; a < 3; a++) {
            float mid = l[a] + G.cs[a] * 0.5f;
            og[a] = hi[a] ? phx_cull_gap(pos[a], mid, l[a] + G.cs[a], m[a]) : phx_cull_gap(pos[a], l[a], mid, m[a]);
        }
        bool ofar = fixed ? (og[0] * og[0] + og[1] * og[1] + og[2] * og[2] >= r_sq)
                          : phx_cull_far(og[0], og[1], og[2], r_sq);
        if (ofar)
            c.oct_fail++;
        /* a hitpoint that is not finite keeps every cell */
        if (t % 64 == 0 && !fixed) {
            float nanp = std::nanf("");
            float gn = phx_cull_gap(nanp, l[0], l[0] + G.cs[0], m[0]);
            if (phx_cull_far(gn, gy, gz, r_sq))
                c.nan_fail++;
        }
    }
    out = c;
}

int main(int argc, char **argv) {
    uint64_t trials = argc > 1 ? strtoull(argv[1], nullptr, 10) : 20000000ull;
    bool fixed = argc > 2 && strcmp(argv[2], "fixed") == 0;
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt ? (nt > 16 ? 16 : nt) : 4;
    std::vector<Counts> cs(nt);
    std::vector<std::thread> th;
    for (unsigned i = 0; i < nt; i++)
        th.emplace_back(run, 1234567ull + 7919ull * i, trials / nt, fixed, std::ref(cs[i]));
    for (auto &t : th)
        t.join();
    Counts s;
    for (auto &c : cs) {
        s.trials += c.trials;
        s.accepted += c.accepted;
        s.near += c.near;
        s.cell_fail += c.cell_fail;
        s.oct_fail += c.oct_fail;
        s.nan_fail += c.nan_fail;
    }
    uint64_t fails = s.cell_fail + s.oct_fail + s.nan_fail;
    if (fixed) {
        printf("fixed %" PRIu64 " violations in %" PRIu64 " accepted\n", fails, s.accepted);
        return fails > 0 ? 0 : 1;
    }
    printf("%s trials %" PRIu64 " accepted %" PRIu64 " outside-cell %" PRIu64 " cell_fail %" PRIu64
           " octant_fail %" PRIu64 " nan_fail %" PRIu64 "\n",
           fails ? "FAIL" : "ok", s.trials, s.accepted, s.near, s.cell_fail, s.oct_fail, s.nan_fail);
    return fails ? 1 : 0;
}
