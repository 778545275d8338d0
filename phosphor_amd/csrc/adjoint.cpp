/*
 * adjoint.cpp — gradient of an SPPM render with respect to per-material
 * parameters θ[m] = (albedo r, g, b, roughness, emission scale).
 *
 * forward(θ): writes θ into the scene's materials and lights (emission scales
 * a material's emissive colour and the power of the lights built from its
 * triangles), runs `rounds` passes of the one-scene driver from a fresh
 * state, records each pass's per-pixel radius ratio and r², and returns the
 * linear estimate (output.cpp:52-83).
 *
 * backward(∂loss/∂image): ∂loss/∂θ by the detached-sampling estimator of
 * kernels_adjoint.hip, sweeping the passes from last to first.  Each pass's
 * photon map is rebuilt by replaying the forward stages with the pass's round
 * seed; nothing but the two per-pixel floats per pass is kept from the
 * forward.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "internal.h"
#include "phx_math.h"
#include "renderer.h"

using namespace phx;

#define PHX_TRY(expr, what)                                                                                            \
    do {                                                                                                               \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess)                                                                                          \
            return hip_status(e_, what);                                                                               \
    } while (0)

static uint32_t host_random_u32(uint32_t &state) { /* random.h:29-34 */
    uint32_t old = state;
    state = state * 747796405u + 2891336453u;
    return ((old >> ((old >> 28u) + 4u)) ^ old) * 277803737u;
}

/* ---- GGX VNDF rejection (bsdfs.h:27-72) ----
 * The sampler draws h from Heitz's visible-normal density
 * D_v(h) = G1(v)·max(0, v·h)·D(h)/(n·v) and rejects it when the mirror
 * direction falls below the surface, 2(v·h)(n·h) < n·v, up to 8 times, then
 * returns the normal.  With n·v = μ, s = √(1 − μ²) and h at polar angle θ,
 * azimuth φ from the plane of incidence, 2(v·h)cos θ − μ = R·cos(2θ − δ)
 * with tan δ = s·cos φ / μ, so for each φ the rejected half vectors are
 * θ ∈ [(δ + π/2)/2, θ_hi], θ_hi = π/2 (or π/2 + δ where cos φ < 0, the
 * edge of v·h > 0).  The region does not depend on α, so
 *   P(α, μ)  = ∫ D_v dω over it,   ∂P/∂α = ∫ D_v·(∂log G1(v) + ∂log D) dω,
 * both smooth integrals: Gauss–Legendre in θ on the interval, trapezoid in
 * φ (periodic).  An accepted draw has probability density D_v(h)·A(α) with
 * A = Σ_{k<8} P^k (any of the 8 tries), a fallback has probability P^8: the
 * score terms ∂log A/∂α and 8·∂log P/∂α that the half vector's score needs
 * besides ∂log D_v. */
namespace {
constexpr int REJ_PHI = 96, REJ_THETA = 24;

struct GaussLegendre {
    double x[REJ_THETA], w[REJ_THETA];
    GaussLegendre() { /* nodes on [-1, 1] by Newton's method on P_n */
        const int n = REJ_THETA;
        for (int i = 0; i < n; i++) {
            double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
            for (int it = 0; it < 100; it++) {
                double p1 = 1.0, p2 = 0.0;
                for (int j = 1; j <= n; j++) {
                    double p3 = p2;
                    p2 = p1;
                    p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
                }
                pp = n * (z * p1 - p2) / (z * z - 1.0);
                double dz = p1 / pp;
                z -= dz;
                if (std::fabs(dz) < 1e-15)
                    break;
            }
            x[i] = z;
            w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
        }
    }
};

void ggx_rejection(double alpha, double mu, double &P, double &dP) {
    static const GaussLegendre gl;
    mu = std::min(std::max(mu, 0.0), 1.0);
    double s = std::sqrt(std::max(0.0, 1.0 - mu * mu));
    double a2 = alpha * alpha;
    double root = std::sqrt(a2 + (1.0 - a2) * mu * mu);
    double g1_over_mu = 2.0 / (mu + root);                                /* G1(v)/μ */
    double dlog_g1v = root > 0.0 ? -alpha * (1.0 - mu * mu) / (root * (mu + root)) : 0.0;
    P = dP = 0.0;
    for (int i = 0; i < REJ_PHI; i++) {
        double phi = 2.0 * M_PI * (i + 0.5) / REJ_PHI;
        double cphi = std::cos(phi);
        double delta = std::atan2(s * cphi, mu);
        double lo = 0.5 * (delta + 0.5 * M_PI);
        double hi = cphi < 0.0 ? 0.5 * M_PI + delta : 0.5 * M_PI;
        if (!(hi > lo))
            continue;
        double half = 0.5 * (hi - lo), mid = 0.5 * (hi + lo), sp = 0.0, sd = 0.0;
        for (int j = 0; j < REJ_THETA; j++) {
            double th = mid + half * gl.x[j];
            double ct = std::cos(th), st = std::sin(th);
            double vh = s * st * cphi + mu * ct;
            if (vh <= 0.0)
                continue;
            double d = ct * ct * (a2 - 1.0) + 1.0;
            double D = a2 / (M_PI * d * d);
            double dens = g1_over_mu * vh * D * st; /* D_v · sin θ (solid angle) */
            double dlog_d = 2.0 / alpha - 4.0 * alpha * ct * ct / d;
            sp += gl.w[j] * dens;
            sd += gl.w[j] * dens * (dlog_g1v + dlog_d);
        }
        P += half * sp;
        dP += half * sd;
    }
    P *= 2.0 * M_PI / REJ_PHI;
    dP *= 2.0 * M_PI / REJ_PHI;
}
} // namespace

namespace phx {
/* per material: ACC_MU samples over μ = n·v of [∂log A/∂α, 8·∂log P/∂α] */
void rejection_table(float roughness, float *out) {
    double r = std::min(std::max((double)roughness, 0.001), 1.0);
    double alpha = r * r;
    for (int i = 0; i < ACC_MU; i++) {
        double P, dP;
        ggx_rejection(alpha, (double)i / (ACC_MU - 1), P, dP);
        double num = 0.0, den = 0.0, pk = 1.0;
        for (int k = 0; k < 8; k++) { /* A = Σ_{k<8} P^k, A' = Σ k P^(k-1) P' */
            den += pk;
            if (k + 1 < 8)
                num += (k + 1) * pk;
            pk *= P;
        }
        out[2 * i] = (float)(num * dP / den);
        out[2 * i + 1] = P > 0.0 ? (float)(8.0 * dP / P) : 0.0f;
    }
}
} // namespace phx

struct PhxAdjoint {
    PhxRenderer *r = nullptr;
    uint32_t M = 0, max_rounds = 0, rounds = 0, n_px = 0;
    std::vector<PhxMaterial> base_mats;
    std::vector<PhxLight> base_lights;
    std::vector<uint32_t> light_mat;
    std::vector<uint32_t> states; /* PCG state before each pass */
    std::vector<float> theta;
    phx::DevBuf d_ratio, d_r2, d_lam_rad, d_lam_tau, d_R, d_nf, d_adj_sorted, d_adj_dep, d_perm, d_grad, d_am,
        d_light_mat, d_rej;
    phx::DevBuf d_chk_emission, d_chk_direct, d_chk_count;
    bool check = false;

    ~PhxAdjoint() { phx_renderer_destroy(r); }
    int apply(const float *th);
    int forward(const float *th, uint32_t n_rounds, float *image);
    int backward(const float *dl_dimage, float *grad);
};

/* θ → materials, lights, luminance prefix sum and the per-material table of
 * derivative factors */
int PhxAdjoint::apply(const float *th) {
    theta.assign(th, th + 5 * (size_t)M);
    std::vector<PhxMaterial> mats = base_mats;
    std::vector<AdjMaterial> am(M);
    std::vector<bool> emits(M, false);
    for (uint32_t m = 0; m < M; m++) {
        const float *t = th + 5 * m;
        if (!(t[0] > 0.0f && t[1] > 0.0f && t[2] > 0.0f) || !(t[4] >= 0.0f) || !(t[3] >= 0.0f))
            return fail(PHX_ERR_INVALID_ARG, "phx_adjoint: albedo must be > 0, roughness and emission >= 0");
        mats[m].base_color.x = t[0];
        mats[m].base_color.y = t[1];
        mats[m].base_color.z = t[2];
        mats[m].roughness = t[3];
        mats[m].emissive.x = base_mats[m].emissive.x * t[4];
        mats[m].emissive.y = base_mats[m].emissive.y * t[4];
        mats[m].emissive.z = base_mats[m].emissive.z * t[4];
        emits[m] = base_mats[m].emissive.x > 0.0f || base_mats[m].emissive.y > 0.0f || base_mats[m].emissive.z > 0.0f;
        am[m] = AdjMaterial{};
        for (int c = 0; c < 3; c++)
            am[m].inv_bc[c] = 1.0f / t[c];
        /* evaluate_material clamps roughness to [MIN_ROUGHNESS, 1] */
        am[m].rough_gate = (t[3] > 0.001f && t[3] < 1.0f) ? 2.0f * t[3] : 0.0f;
    }
    std::vector<PhxLight> lights = base_lights;
    for (size_t l = 0; l < lights.size(); l++) {
        uint32_t m = light_mat[l];
        if (m == 0xFFFFFFFFu)
            continue;
        emits[m] = true;
        lights[l].power.x = base_lights[l].power.x * th[5 * m + 4];
        lights[l].power.y = base_lights[l].power.y * th[5 * m + 4];
        lights[l].power.z = base_lights[l].power.z * th[5 * m + 4];
    }
    for (uint32_t m = 0; m < M; m++)
        am[m].inv_emis = (emits[m] && th[5 * m + 4] > 0.0f) ? 1.0f / th[5 * m + 4] : 0.0f;
    ShadingBuffers &sh = r->buffers.scene.shading;
    std::vector<float> lum;
    float running = 0.0f;
    for (const PhxLight &l : lights) { /* scene_data.cpp:14-23 */
        float v = l.power.x * 0.2126f + l.power.y * 0.7152f + l.power.z * 0.0722f;
        running += std::max(v, 0.0f);
        lum.push_back(running);
    }
    sh.total_lum = lum.empty() ? 0.0f : lum.back();
    PHX_TRY(hipMemcpy(sh.d_materials.p, mats.data(), M * sizeof(PhxMaterial), hipMemcpyHostToDevice), "materials");
    if (!lights.empty()) {
        PHX_TRY(hipMemcpy(sh.d_lights.p, lights.data(), lights.size() * sizeof(PhxLight), hipMemcpyHostToDevice),
                "lights");
        PHX_TRY(hipMemcpy(sh.d_lum_pref_sum.p, lum.data(), lum.size() * sizeof(float), hipMemcpyHostToDevice),
                "lights");
    }
    PHX_TRY(hipMemcpy(d_am.p, am.data(), M * sizeof(AdjMaterial), hipMemcpyHostToDevice), "adjoint table");
    return PHX_OK;
}

int PhxAdjoint::forward(const float *th, uint32_t n_rounds, float *image) {
    if (n_rounds == 0 || n_rounds > max_rounds)
        return fail(PHX_ERR_INVALID_ARG, "phx_adjoint_forward: rounds outside 1..max_rounds");
    PHX_TRY(r->sync_all(), "adjoint");
    int rc = apply(th);
    if (rc != PHX_OK)
        return rc;
    /* a fresh progressive state (renderer.cpp:9-57) */
    const PhxRenderArgs &a = r->args;
    ImageBuffers &im = r->buffers.image_buf;
    std::vector<phx_float4> f(n_px, phx_float4{0.0f, 0.0f, 0.0f,
                                               a.initial_radius > 0.0f ? a.initial_radius * a.initial_radius : 0.0f});
    PHX_TRY(hipMemcpy(im.d_flux.p, f.data(), n_px * sizeof(phx_float4), hipMemcpyHostToDevice), "reset");
    PHX_TRY(hipMemset(im.d_photon_weight.p, 0, n_px * sizeof(float)), "reset");
    PHX_TRY(hipMemset(im.d_radiance.p, 0, n_px * sizeof(phx_float4)), "reset");
    uint32_t state = host_pcg_seed(a.seed);
    states.clear();
    for (uint32_t p = 0; p < n_rounds; p++) {
        states.push_back(state);
        host_random_u32(state);
    }
    uint32_t rng0[2] = {states[0], 0u};
    PHX_TRY(hipMemcpy(r->d_rng.p, rng0, sizeof(rng0), hipMemcpyHostToDevice), "reset");
    r->rounds = 0;
    r->total_photons_emitted = 0;
    for (uint32_t p = 0; p < n_rounds; p++) {
        r->rec_ratio = d_ratio.as<float>() + (size_t)p * n_px;
        r->rec_r2 = d_r2.as<float>() + (size_t)p * n_px;
        rc = r->run_pass();
        r->rec_ratio = r->rec_r2 = nullptr;
        if (rc != PHX_OK)
            return rc;
    }
    rounds = n_rounds;
    if ((rc = phx_renderer_synchronize(r)) != PHX_OK)
        return rc;
    return image ? phx_renderer_estimate(r, image) : PHX_OK;
}

int PhxAdjoint::backward(const float *dl, float *grad) {
    if (rounds == 0)
        return fail(PHX_ERR_INVALID_ARG, "phx_adjoint_backward: no forward run");
    PHX_TRY(r->sync_all(), "adjoint");
    hipStream_t s = r->stream;
    ImageBuffers &im = r->buffers.image_buf;
    /* λ of the two terms of output.cpp:52-83: L = radiance/rounds + τ·inv_ph/(π r²) */
    std::vector<phx_float4> flux(n_px), lam_rad(n_px), lam_tau(n_px);
    PHX_TRY(hipMemcpy(flux.data(), im.d_flux.p, n_px * sizeof(phx_float4), hipMemcpyDeviceToHost), "adjoint");
    const float PI_F = 3.141592654f;
    float inv_ph = r->total_photons_emitted > 0 ? 1.0f / (float)r->total_photons_emitted : 0.0f;
    float inv_rd = 1.0f / (float)rounds;
    for (uint32_t p = 0; p < n_px; p++) {
        float area = PI_F * flux[p].w;
        float k = area > 0.0f ? inv_ph / area : 0.0f;
        const float *g = dl + 3 * (size_t)p;
        lam_rad[p] = phx_float4{g[0] * inv_rd, g[1] * inv_rd, g[2] * inv_rd, 0.0f};
        lam_tau[p] = phx_float4{g[0] * k, g[1] * k, g[2] * k, 0.0f};
    }
    PHX_TRY(hipMemcpy(d_lam_rad.p, lam_rad.data(), n_px * sizeof(phx_float4), hipMemcpyHostToDevice), "adjoint");
    PHX_TRY(hipMemcpy(d_lam_tau.p, lam_tau.data(), n_px * sizeof(phx_float4), hipMemcpyHostToDevice), "adjoint");
    /* the VNDF rejection terms of each material's roughness */
    std::vector<float> rej((size_t)M * ACC_MU * 2);
    for (uint32_t m = 0; m < M; m++)
        rejection_table(theta[5 * m + 3], rej.data() + (size_t)m * ACC_MU * 2);
    PHX_TRY(copy_to_dev_buf(d_rej, rej.data(), rej.size(), sizeof(float), s), "adjoint");
    PHX_TRY(launch_fill_u32(d_R.as<uint32_t>(), 0x3F800000u, n_px, s), "adjoint"); /* 1.0f */
    PHX_TRY(launch_fill_u32(d_grad.as<uint32_t>(), 0u, 5 * (size_t)M, s), "adjoint");
    const PhxRenderArgs &a = r->args;
    for (uint32_t p = rounds; p-- > 0;) {
        uint32_t rng0[2] = {states[p], 0u};
        PHX_TRY(hipMemcpyAsync(r->d_rng.p, rng0, sizeof(rng0), hipMemcpyHostToDevice, s), "adjoint");
        PHX_TRY(hipStreamSynchronize(s), "adjoint");
        PHX_TRY(launch_next_round_seed(r->d_rng.as<uint32_t>(), r->seed_ptr(), 1, s), "adjoint");
        int rc;
        if ((rc = r->trace_rays()) != PHX_OK || (rc = r->emit_photons()) != PHX_OK)
            return rc;
        r->sort_perm = d_perm.as<uint32_t>();
        rc = r->hash_and_sort_photons();
        r->sort_perm = nullptr;
        if (rc != PHX_OK)
            return rc;
        AdjPass P{};
        P.sc = r->buffers.scene_args();
        P.cam = im.camera;
        P.hash = r->photon_hash_info;
        P.w = a.width;
        P.h = a.height;
        P.n_direct = a.direct_samples;
        P.photons = a.photons_per_round;
        P.max_bounces = a.max_photon_bounces;
        P.M = M;
        P.max_deposits = r->photon_buf_size;
        P.d_seed = r->seed_ptr();
        P.am = d_am.as<AdjMaterial>();
        P.rej = d_rej.as<float>();
        P.light_mat = d_light_mat.as<uint32_t>();
        P.hp_pos = im.d_hp_pos.as<phx_float4>();
        P.hp_normal = im.d_hp_normal.as<phx_float4>();
        P.hp_throughput = im.d_hp_throughput.as<phx_float4>();
        P.hp_base_color = im.d_hp_base_color.as<phx_float4>();
        P.hp_emission = im.d_hp_emission.as<phx_float4>();
        const PhotonBuffers &pb = r->buffers.photon_buf;
        P.ppos = pb.d_photon_pos.as<phx_float4>();
        P.ppow = pb.d_photon_power.as<uint32_t>();
        P.pdir = pb.d_photon_dir.as<uint32_t>();
        P.cell_start = pb.d_cell_start.as<uint32_t>();
        P.cell_end = pb.d_cell_end.as<uint32_t>();
        P.perm = d_perm.as<uint32_t>();
        P.d_count = r->d_batch_size.as<uint32_t>();
        P.emit_ws = r->d_emit_ws.p;
        P.ratio = d_ratio.as<float>() + (size_t)p * n_px;
        P.r2 = d_r2.as<float>() + (size_t)p * n_px;
        P.R = d_R.as<float>();
        P.lam_rad = d_lam_rad.as<phx_float4>();
        P.lam_tau = d_lam_tau.as<phx_float4>();
        P.nf = d_nf.as<phx_float4>();
        P.adj_sorted = d_adj_sorted.as<phx_float4>();
        P.adj_dep = d_adj_dep.as<phx_float4>();
        P.grad = d_grad.as<float>();
        if (check && p == 0) {
            P.chk_emission = d_chk_emission.as<phx_float4>();
            P.chk_direct = d_chk_direct.as<phx_float4>();
            P.chk_count = d_chk_count.as<uint32_t>();
        }
        PHX_TRY(launch_adjoint_pass(P, s), "adjoint pass");
    }
    PHX_TRY(hipStreamSynchronize(s), "adjoint");
    PHX_TRY(hipMemcpy(grad, d_grad.p, 5 * (size_t)M * sizeof(float), hipMemcpyDeviceToHost), "adjoint");
    return PHX_OK;
}

extern "C" {

int phx_adjoint_ggx_rejection(float alpha, float mu, double *out) {
    if (!out || !(alpha > 0.0f) || !(alpha <= 1.0f) || !(mu >= 0.0f) || !(mu <= 1.0f))
        return fail(PHX_ERR_INVALID_ARG, "phx_adjoint_ggx_rejection: alpha outside (0, 1] or mu outside [0, 1]");
    ggx_rejection(alpha, mu, out[0], out[1]);
    return PHX_OK;
}

int phx_adjoint_create(const PhxRenderArgs *args, const PhxSceneDesc *scene, uint32_t max_rounds,
                       PhxAdjoint **out) {
    if (!args || !scene || !out || max_rounds == 0)
        return fail(PHX_ERR_INVALID_ARG, "phx_adjoint_create: null argument or zero rounds");
    *out = nullptr;
    if (check_scene_desc(*scene, "phx_adjoint_create") != PHX_OK)
        return PHX_ERR_INVALID_ARG;
    if (scene->n_materials > 16)
        return fail(PHX_ERR_INVALID_ARG, "phx_adjoint_create: at most 16 materials");
    for (uint32_t m = 0; m < scene->n_materials; m++) {
        const PhxMaterial &mt = scene->materials[m];
        uint32_t tex[6] = {mt.diff_index, mt.emis_index, mt.norm_index, mt.occlusion_index, mt.metal_rough_index,
                           mt.transmission_index};
        for (uint32_t t : tex)
            if (t != 0xFFFFFFFFu)
                return fail(PHX_ERR_INVALID_ARG, "phx_adjoint_create: textured materials are not differentiated");
    }
    PhxRenderArgs ra = *args;
    ra.flags = 0;
    ra.trace_mode = 0;
    ra.max_photons_per_batch = ra.photons_per_round; /* one emission sequence: deposit offsets per photon */
    std::unique_ptr<PhxAdjoint> A(new PhxAdjoint());
    int rc = phx_renderer_create(&ra, scene, &A->r);
    if (rc != PHX_OK)
        return rc;
    PhxRenderer *r = A->r;
    if (r->emit_chunk != ra.photons_per_round) {
        r->emit_chunk = ra.photons_per_round;
        PHX_TRY(r->d_emit_ws.alloc(emit_workspace_bytes(ra.photons_per_round, ra.max_photon_bounces)),
                "adjoint emit workspace");
    }
    A->M = scene->n_materials;
    A->max_rounds = max_rounds;
    A->n_px = ra.width * ra.height;
    A->base_mats.assign(scene->materials, scene->materials + A->M);
    if (scene->n_lights)
        A->base_lights.assign(scene->lights, scene->lights + scene->n_lights);
    for (const PhxLight &l : A->base_lights) {
        uint32_t m = 0xFFFFFFFFu;
        if (l.kind == PHX_LIGHT_TEXTURED) {
            uint32_t start = phx_as_uint(l.aux.y);
            if (scene->etri_mat_index && start < scene->n_e_triangles)
                m = scene->etri_mat_index[start];
        }
        A->light_mat.push_back(m < A->M ? m : 0xFFFFFFFFu);
    }
    size_t np = A->n_px, nd = r->photon_buf_size;
    PHX_TRY(A->d_ratio.alloc(np * max_rounds * sizeof(float)), "adjoint");
    PHX_TRY(A->d_r2.alloc(np * max_rounds * sizeof(float)), "adjoint");
    PHX_TRY(A->d_lam_rad.alloc(np * sizeof(phx_float4)), "adjoint");
    PHX_TRY(A->d_lam_tau.alloc(np * sizeof(phx_float4)), "adjoint");
    PHX_TRY(A->d_R.alloc(np * sizeof(float)), "adjoint");
    PHX_TRY(A->d_nf.alloc(np * sizeof(phx_float4)), "adjoint");
    PHX_TRY(A->d_adj_sorted.alloc(nd * sizeof(phx_float4)), "adjoint");
    PHX_TRY(A->d_adj_dep.alloc(nd * sizeof(phx_float4)), "adjoint");
    PHX_TRY(A->d_perm.alloc(nd * sizeof(uint32_t)), "adjoint");
    PHX_TRY(A->d_grad.alloc(5 * (size_t)A->M * sizeof(float)), "adjoint");
    PHX_TRY(A->d_am.alloc(A->M * sizeof(AdjMaterial)), "adjoint");
    PHX_TRY(copy_to_dev_buf(A->d_light_mat, A->light_mat.data(), A->light_mat.size(), 4, r->stream), "adjoint");
    PHX_TRY(A->d_chk_emission.alloc(np * sizeof(phx_float4)), "adjoint");
    PHX_TRY(A->d_chk_direct.alloc(np * sizeof(phx_float4)), "adjoint");
    PHX_TRY(A->d_chk_count.alloc((size_t)ra.photons_per_round * sizeof(uint32_t)), "adjoint");
    PHX_TRY(hipStreamSynchronize(r->stream), "adjoint");
    *out = A.release();
    return PHX_OK;
}

void phx_adjoint_destroy(PhxAdjoint *a) { delete a; }

uint32_t phx_adjoint_num_params(PhxAdjoint *a) { return a ? 5 * a->M : 0u; }

int phx_adjoint_forward(PhxAdjoint *a, const float *theta, uint32_t rounds, float *image_out) {
    if (!a || !theta)
        return fail(PHX_ERR_INVALID_ARG, "phx_adjoint_forward: null argument");
    return a->forward(theta, rounds, image_out);
}

int phx_adjoint_backward(PhxAdjoint *a, const float *dloss_dimage, float *grad_out) {
    if (!a || !dloss_dimage || !grad_out)
        return fail(PHX_ERR_INVALID_ARG, "phx_adjoint_backward: null argument");
    return a->backward(dloss_dimage, grad_out);
}

int phx_adjoint_set_replay_check(PhxAdjoint *a, uint32_t on) {
    if (!a)
        return fail(PHX_ERR_INVALID_ARG, "null adjoint");
    a->check = on != 0;
    return PHX_OK;
}

/* what the last backward's replay kernels recomputed for the first pass
 * (replay check on): per pixel emission and direct, per photon its number of
 * deposits; forward_counts receives the forward emission's counts of the same
 * replayed pass, and phx_adjoint_renderer's hitpoints hold its trace */
int phx_adjoint_read_replay(PhxAdjoint *a, phx_float4 *emission, phx_float4 *direct, uint32_t *counts,
                            uint32_t *forward_counts) {
    if (!a)
        return fail(PHX_ERR_INVALID_ARG, "null adjoint");
    size_t np = a->n_px, nph = a->r->args.photons_per_round;
    if (emission)
        PHX_TRY(hipMemcpy(emission, a->d_chk_emission.p, np * 16, hipMemcpyDeviceToHost), "read");
    if (direct)
        PHX_TRY(hipMemcpy(direct, a->d_chk_direct.p, np * 16, hipMemcpyDeviceToHost), "read");
    if (counts)
        PHX_TRY(hipMemcpy(counts, a->d_chk_count.p, nph * 4, hipMemcpyDeviceToHost), "read");
    if (forward_counts)
        PHX_TRY(hipMemcpy(forward_counts, adjoint_emit_counts(a->r->d_emit_ws.p, (uint32_t)nph,
                                                              a->r->args.max_photon_bounces),
                          nph * 4, hipMemcpyDeviceToHost),
                "read");
    return PHX_OK;
}

PhxRenderer *phx_adjoint_renderer(PhxAdjoint *a) { return a ? a->r : nullptr; }

} /* extern "C" */
