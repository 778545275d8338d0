/*
 * CPU check of the gather's division-free Gaussian weight against the literal
 * expression GAUSS_ALPHA·(1 − (1 − exp(−GAUSS_BETA·dsq / (2r²))) / GAUSS_DENOM)
 * (photon_hash.h:86-90) with IEEE divisions, bit for bit:
 *   1. u / GAUSS_DENOM for u = 0 and every float u in [2^-100, 1] (1 − exp
 *      is 0 or at least 2^-24);
 *   2. the whole weight for random (r², dsq) pairs, r² log-uniform over
 *      [7.5e-7, 1e4] (the gather's radius is at least the half-diagonal of a
 *      0.001 cell) and dsq uniform in [0, r²) or log-uniform down to 1e-45·r²,
 *      and the quotient itself wherever its numerator is at least 2^-100;
 *   3. phx_expf_normal(x) = 1 for |x| < 2^-26 (boundary sweep): numerators
 *      below 2^-100 give quotients below 2^-80 in this range of r², and the
 *      same weight whatever their last bits.
 * Prints "ok <counts>" or the first mismatch; exit status 0 only on success.
 * Built with -ffp-contract=off, like the library.
 */
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "../../phosphor_amd/csrc/phx_math.h"

static const float ALPHA = 1.818f, BETA = 1.953f, DENOM = 0.85815f; /* constants.h:43-45 */
static const float INV_DENOM = 1.0f / DENOM;

static float literal(float dsq, float two_r_sq) {
    return ALPHA * (1.0f - (1.0f - phx_expf(-BETA * dsq / two_r_sq)) / DENOM);
}

static float fast(float dsq, float two_r_sq, float inv) {
    float x = phx_div_recip(-BETA * dsq, two_r_sq, inv);
    float e = phx_expf_normal(x);
    return ALPHA * (1.0f - phx_div_recip(1.0f - e, DENOM, INV_DENOM));
}

int main(int argc, char **argv) {
    long pairs = argc > 1 ? atol(argv[1]) : 20000000;
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt ? (nt > 16 ? 16 : nt) : 4;
    /* 1: exhaustive over [0, 1] */
    const uint32_t lo = phx_as_uint(0x1p-100f), hi = phx_as_uint(1.0f);
    std::vector<uint64_t> bad1(nt, 0);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (uint64_t u = lo - 1 + t; u <= hi; u += nt) {
                float x = u < lo ? 0.0f : phx_as_float((uint32_t)u);
                float a = x / DENOM, b = phx_div_recip(x, DENOM, INV_DENOM);
                if (phx_as_uint(a) != phx_as_uint(b)) {
                    if (!bad1[t])
                        fprintf(stderr, "denominator: u=%a ieee=%a fast=%a\n", x, a, b);
                    bad1[t]++;
                }
            }
        });
    for (auto &x : th)
        x.join();
    th.clear();
    uint64_t b1 = 0;
    for (auto v : bad1)
        b1 += v;
    /* 2: random pairs */
    std::vector<uint64_t> bad2(nt, 0), tiny(nt, 0);
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            std::mt19937_64 rng(1234 + t);
            std::uniform_real_distribution<double> U(0.0, 1.0);
            for (long i = t; i < pairs; i += nt) {
                float r_sq = (float)(7.5e-7 * std::pow(10.0, 10.2 * U(rng)));
                float dsq = (i & 1) ? (float)(U(rng) * r_sq) : (float)std::pow(10.0, -45.0 + 45.0 * U(rng)) * r_sq;
                if (!(dsq < r_sq))
                    continue;
                float two = 2.0f * r_sq, inv = 1.0f / two;
                float wl = literal(dsq, two), wf = fast(dsq, two, inv);
                float ql = -BETA * dsq / two, qf = phx_div_recip(-BETA * dsq, two, inv);
                bool small = std::fabs(-BETA * dsq) < 0x1p-100f;
                bool quotient_bad = !small && phx_as_uint(ql) != phx_as_uint(qf);
                if (small) {
                    tiny[t]++;
                    quotient_bad = std::fabs(ql) >= 0x1p-80f || std::fabs(qf) >= 0x1p-80f;
                }
                if (phx_as_uint(wl) != phx_as_uint(wf) || quotient_bad) {
                    if (!bad2[t])
                        fprintf(stderr, "weight: dsq=%a two_r_sq=%a literal=%a fast=%a q=%a/%a\n", dsq, two, wl, wf,
                                ql, qf);
                    bad2[t]++;
                }
            }
        });
    for (auto &x : th)
        x.join();
    uint64_t b2 = 0, nt2 = 0;
    for (unsigned t = 0; t < nt; t++) {
        b2 += bad2[t];
        nt2 += tiny[t];
    }
    /* 3: exp of tiny arguments */
    uint64_t b3 = 0, n3 = 0;
    for (int sgn = 0; sgn < 2; sgn++)
        for (uint32_t u = 0; u <= phx_as_uint(0x1p-26f); u += (u < 4096 ? 1 : 977)) {
            float x = phx_as_float(u | (sgn ? 0x80000000u : 0u));
            n3++;
            if (phx_expf_normal(x) != 1.0f) {
                if (!b3)
                    fprintf(stderr, "tiny exp: x=%a gives %a\n", x, phx_expf_normal(x));
                b3++;
            }
        }
    printf("%s denominator_mismatch=%" PRIu64 " weight_mismatch=%" PRIu64 " (pairs=%ld, tiny quotients=%" PRIu64
           " with numerators below 2^-100) tiny_exp_mismatch=%" PRIu64 " of %" PRIu64 "\n",
           (b1 || b2 || b3) ? "FAIL" : "ok", b1, b2, pairs, nt2, b3, n3);
    return (b1 || b2 || b3) ? 1 : 0;
}
