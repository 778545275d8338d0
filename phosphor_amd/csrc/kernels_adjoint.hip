/*
 * kernels_adjoint.hip — reverse-mode gradient of an SPPM render with respect
 * to per-material albedo (RGB), GGX roughness and emission scale.
 *
 * Estimator: detached sampling.  Every random decision of a pass (light and
 * emissive point, BSDF event, GGX half vector, hemisphere direction, Russian
 * roulette) is replayed from the same PCG stream and held fixed; the
 * gradient differentiates the weights.  For a sampled event the weight is
 * f/p with p held constant, so its log-derivative is that of the integrand:
 *   albedo   : 1/base_color for diffuse and transmission events, the
 *              log-derivative of Schlick's F(f0 = base_color) for conductors;
 *   roughness: the score of the sampled half vector under the visible-normal
 *              distribution, ∂log[G1(v) D(h)], plus ∂log G1(l) for the
 *              specular events (conductor, dielectric reflection,
 *              transmission); NEE's GGX term adds ∂log[D G1(l) G1(v)];
 *   emission : 1/scale for light power and emissive radiance.
 * Russian-roulette probabilities are treated as constants, which keeps the
 * estimate unbiased.  Photon powers are differentiated in float (the RGBE
 * code is piecewise constant).  The found counts and radii of the gather do
 * not depend on the parameters, so a pixel's final flux is Σ_p Φ_p·R_p with
 * R_p the product of the recorded radius ratios of passes p..P-1.
 *
 * Per pass, last to first: the photon map is rebuilt by the forward stages,
 * the gather adjoint scatters ∂loss/∂(photon power) to the records it finds,
 * and two replay kernels (photon paths, camera paths) walk the paths again,
 * keeping per lane the log-derivatives of the path throughput per material
 * (LDS, [entry][lane]) and per-lane gradient sums that are reduced once per
 * block.  No path is stored.
 */
#include "pass_bodies.h"

namespace phx {

constexpr int ADJ_BLOCK = 64;

/* ∂log G1(c; α)/∂α of smith_g1_ggx (bsdfs.h) */
PHX_DEV f32 dlog_g1(f32 c, f32 alpha) {
    if (c <= 0.0f)
        return 0.0f;
    f32 a2 = alpha * alpha;
    f32 s = phx_sqrtf(a2 + (1.0f - a2) * c * c);
    return -alpha * (1.0f - c * c) / (s * (c + s));
}

/* ∂log D(n·h; α)/∂α of ggx_ndf */
PHX_DEV f32 dlog_ndf(f32 ndh, f32 alpha) {
    f32 a2 = alpha * alpha;
    f32 d = ndh * ndh * (a2 - 1.0f) + 1.0f;
    return 2.0f / alpha - 4.0f * alpha * ndh * ndh / d;
}

/* ∂log F_c/∂f0_c of fresnel4(f0, i, n) per channel */
PHX_DEV F4 dlog_fresnel_f0(F4 f0, F4 i, F4 n) {
    f32 c = phx_fabsf(dot(i, n));
    f32 x = 1.0f - c;
    f32 p = (x * x) * (x * x) * x;
    return mk((1.0f - p) / (f0.x + (1.0f - f0.x) * p), (1.0f - p) / (f0.y + (1.0f - f0.y) * p),
              (1.0f - p) / (f0.z + (1.0f - f0.z) * p), 0.0f);
}

/* the rejection loop's share of the half vector's score at n·v = mu: the
 * material's table (ACC_MU samples over [0, 1], linear interpolation) of
 * ∂log A/∂α for an accepted draw, 8·∂log P/∂α for the fallback */
PHX_DEV f32 rejection_score(const f32 *rej, f32 mu, u32 fell_back) {
    f32 x = phx_clampf(mu, 0.0f, 1.0f) * (f32)(ACC_MU - 1);
    u32 i = (u32)x;
    i = i < (u32)(ACC_MU - 2) ? i : (u32)(ACC_MU - 2);
    f32 f = x - (f32)i;
    f32 a = rej[2 * i + fell_back], b = rej[2 * (i + 1) + fell_back];
    return a + (b - a) * f;
}

/* log-derivative of the sampled event's weight (albedo per channel in xyz,
 * roughness in w) for a hit with material m; `score` receives the roughness
 * part that belongs to choosing the event at all (the half vector's score:
 * ∂log of its density under the VNDF and of the rejection loop's
 * acceptance), which is what a contribution conditional on the event
 * carries.  A fallback draw (h = the normal) has no VNDF density term. */
PHX_DEV F4 event_dlog(const AdjMaterial &am, const f32 *rej, const Ctx &c, const Bsdf &bs, F4 view, F4 sn,
                      f32 &score_out) {
    f32 alpha = c.roughness * c.roughness;
    f32 ndv = phx_fmaxf(dot(view, sn), 0.0f);
    f32 score = bs.fell_back ? 0.0f : dlog_g1(ndv, alpha) + dlog_ndf(dot(bs.h, sn), alpha);
    score += rejection_score(rej, ndv, bs.fell_back);
    score_out = score * am.rough_gate;
    F4 r = mk(0.0f, 0.0f, 0.0f, 0.0f);
    f32 rough = score;
    if (bs.event == BSDF_DIFFUSE) {
        r = mk(am.inv_bc[0], am.inv_bc[1], am.inv_bc[2], 0.0f);
    } else if (bs.event == BSDF_METALLIC) {
        r = dlog_fresnel_f0(c.base_color, -view, bs.h);
        rough += dlog_g1(phx_fmaxf(0.0f, dot(bs.dir, sn)), alpha);
    } else if (bs.event == BSDF_FRESNEL) {
        rough += dlog_g1(phx_fmaxf(0.0f, dot(bs.dir, sn)), alpha);
    } else {
        r = mk(am.inv_bc[0], am.inv_bc[1], am.inv_bc[2], 0.0f);
        rough += dlog_g1(phx_fmaxf(0.0f, phx_fabsf(dot(bs.dir, sn))), alpha);
    }
    r.w = rough * am.rough_gate;
    return r;
}

/* per-lane state in LDS: S[m][4] path log-derivatives, G[m][5] gradient sums */
struct AdjLane {
    f32 *S; /* S[(m * 4 + k) * ADJ_BLOCK] */
    f32 *G; /* G[(m * 5 + k) * ADJ_BLOCK] */
    u32 M;

    __device__ __forceinline__ f32 &s(u32 m, u32 k) const { return S[(m * 4 + k) * ADJ_BLOCK]; }
    __device__ __forceinline__ f32 &g(u32 m, u32 k) const { return G[(m * 5 + k) * ADJ_BLOCK]; }
    __device__ __forceinline__ void clear_path() const {
        for (u32 m = 0; m < M; m++)
            for (u32 k = 0; k < 4; k++)
                s(m, k) = 0.0f;
    }
    /* a contribution lt = λ ⊙ T carried by the current path */
    __device__ __forceinline__ void add_path(F4 lt) const {
        f32 sum = lt.x + lt.y + lt.z;
        for (u32 m = 0; m < M; m++) {
            g(m, 0) += lt.x * s(m, 0);
            g(m, 1) += lt.y * s(m, 1);
            g(m, 2) += lt.z * s(m, 2);
            g(m, 3) += sum * s(m, 3);
        }
    }
    /* a contribution's own log-derivative d (albedo xyz, roughness w) for material m */
    __device__ __forceinline__ void add_local(u32 m, F4 lt, F4 d) const {
        g(m, 0) += lt.x * d.x;
        g(m, 1) += lt.y * d.y;
        g(m, 2) += lt.z * d.z;
        g(m, 3) += (lt.x + lt.y + lt.z) * d.w;
    }
    __device__ __forceinline__ void advance(u32 m, F4 d) const {
        s(m, 0) += d.x;
        s(m, 1) += d.y;
        s(m, 2) += d.z;
        s(m, 3) += d.w;
    }
};

/* wave sum of the per-lane gradients (a block is one wave), one atomic per
 * entry and block; every lane of the block must call it */
PHX_DEV void flush_grads(const AdjLane &L, f32 *grad) {
    u32 n = L.M * 5;
    for (u32 e = 0; e < n; e++) {
        f32 v = L.G[e * ADJ_BLOCK];
        for (int o = 32; o >= 1; o >>= 1)
            v += __shfl_xor(v, o, 64);
        if (threadIdx.x == 0)
            atomicAdd(&grad[e], v);
    }
}

/* ---------------- gather adjoint ---------------- */
struct AdjGatherArgs {
    const phx_float4 *hp_pos, *hp_normal, *hp_throughput, *hp_base_color, *hp_emission;
    u32 n_pixels;
    const phx_float4 *ppos;
    const u32 *ppow, *pdir, *cell_start, *cell_end;
    PhxPhotonHashData d;
    const f32 *r2;          /* the pass's r² per pixel */
    const phx_float4 *lam_tau;
    const f32 *R;
    phx_float4 *adj;        /* per sorted photon record: ∂loss/∂power (xyz) */
    phx_float4 *nf;         /* per pixel: Σ w·power of the pass */
};

/* the 27-cell query of gather_pixel with the pass's radius: the per-pixel
 * photon sum (for the camera side) and ∂loss/∂power_k = λ_Φ ⊙ thr ⊙ bc/π · w_k
 * scattered to every record found */
__global__ __launch_bounds__(256) void k_adj_gather(AdjGatherArgs a) {
    u32 tid = blockIdx.x * 256 + threadIdx.x;
    if (tid >= a.n_pixels)
        return;
    F4 ge = ld(a.hp_emission + tid);
    if (ge.w == 0.0f) {
        st(a.nf + tid, BLACK);
        return;
    }
    const PhxPhotonHashData &d = a.d;
    f32 r_sq = a.r2[tid];
    F4 lam = ld(a.lam_tau + tid) * a.R[tid];
    F4 coef = lam * ld(a.hp_throughput + tid) * (ld(a.hp_base_color + tid) / PI);
    bool scatter = (coef.x != 0.0f) || (coef.y != 0.0f) || (coef.z != 0.0f);
    F4 pos = ld(a.hp_pos + tid), sn = ld(a.hp_normal + tid);
    F4 nf = BLACK;
    f32 two_r_sq = 2.0f * r_sq;
    for (u32 i = 0; i < 27; i++) {
        i32 cx = (i32)(i % 3) - 1, cy = (i32)((i / 3) % 3) - 1, cz = (i32)((i / 9) % 3) - 1;
        u32 h = get_hash(pos.x + (f32)cx * d.cell_sizes.x, pos.y + (f32)cy * d.cell_sizes.y,
                         pos.z + (f32)cz * d.cell_sizes.z, d);
        if (h == 0)
            continue;
        u32 e0 = a.cell_end[h];
        for (u32 k = a.cell_start[h]; k < e0; k++) {
            F4 dd = pos - ld(a.ppos + k);
            f32 dsq = dot(dd, dd);
            if (dsq >= r_sq)
                continue;
            if (dot(decode_oct(a.pdir[k]), sn) > 0.0f)
                continue;
            f32 wgt = GAUSS_ALPHA * (1.0f - (1.0f - phx_expf(-GAUSS_BETA * dsq / two_r_sq)) / GAUSS_DENOM);
            nf = nf + wgt * decode_rgbe(a.ppow[k]);
            if (scatter) {
                atomicAdd(&a.adj[k].x, coef.x * wgt);
                atomicAdd(&a.adj[k].y, coef.y * wgt);
                atomicAdd(&a.adj[k].z, coef.z * wgt);
            }
        }
    }
    st(a.nf + tid, nf);
}

/* sorted → deposit order: adj_dep[perm[s]] = adj_sorted[s] */
__global__ __launch_bounds__(256) void k_adj_unsort(const phx_float4 *__restrict__ adj_sorted,
                                                    const u32 *__restrict__ perm, const u32 *__restrict__ d_n,
                                                    u32 max_n, phx_float4 *__restrict__ adj_dep) {
    u32 n = *d_n;
    n = n < max_n ? n : max_n;
    for (u32 s = blockIdx.x * 256 + threadIdx.x; s < n; s += gridDim.x * 256)
        adj_dep[perm[s]] = adj_sorted[s];
}

/* R[px] *= ratio[px] */
__global__ __launch_bounds__(256) void k_adj_ratio(f32 *__restrict__ R, const f32 *__restrict__ ratio, u32 n) {
    u32 i = blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        R[i] *= ratio[i];
}

/* ---------------- photon-path replay ---------------- */
struct AdjPhotonArgs {
    SceneView sc;
    u32 photons, max_bounces;
    const u32 *d_seed;
    const AdjMaterial *am;
    const f32 *rej;
    const u32 *light_mat; /* material of each light (0xFFFFFFFF: none) */
    u32 M;
    const u32 *offs;         /* first deposit of each photon (exclusive scan of counts) */
    const phx_float4 *adj;   /* per deposit (deposit order) */
    u32 max_deposits;
    f32 *grad;
    u32 *chk_count;          /* optional: deposits of each photon (replay check) */
};

/* dynamic LDS of the replay kernels: S then G, M·9 floats per lane */
extern __shared__ f32 s_adj_dyn[];

__global__ __launch_bounds__(ADJ_BLOCK) void k_adj_photons(AdjPhotonArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * ADJ_BLOCK];
    __shared__ f32 s_ior[IOR_STACK_SIZE * ADJ_BLOCK];
    u32 t = threadIdx.x;
    AdjLane L{s_adj_dyn + t, s_adj_dyn + a.M * 4 * ADJ_BLOCK + t, a.M};
    for (u32 e = 0; e < a.M * 5; e++)
        L.G[e * ADJ_BLOCK] = 0.0f;
    L.clear_path();
    u32 li = blockIdx.x * ADJ_BLOCK + t;
    const SceneView &sc = a.sc;
    if (li < a.photons) {
        u32 *stk = s_stack + t;
        u32 rng = pcg_seed(*a.d_seed + li);
        F4 origin, dir, power;
        u32 light;
        sample_light(rng, sc, origin, dir, power, &light);
        u32 em = (light != 0xFFFFFFFFu) ? a.light_mat[light] : 0xFFFFFFFFu;
        f32 inv_e = (em != 0xFFFFFFFFu) ? a.am[em].inv_emis : 0.0f;
        F4 throughput = mk(1.0f, 1.0f, 1.0f, 0.0f);
        IorStack ior{s_ior + t, (u32)ADJ_BLOCK, 0u};
        init_ior_stack(ior);
        u32 n = 0, base = a.offs[li];
        /* emit_photons.cl:49-113 */
        for (u32 depth = 0; depth < a.max_bounces; depth++) {
            HitRecord rec;
            if (!scene_intersect(sc, origin, dir, EPS, INF, rec, stk, ADJ_BLOCK))
                break;
            SurfaceHit sh = process_hit(sc, rec, origin, dir);
            const PhxMaterial &mat = sc.materials[sh.mat_index];
            f32 curr_ior = current_ior(ior);
            if (curr_ior != AIR_IOR && mat.thickness > EPS && mat.att_dist > EPS) {
                F4 vol = beer_lambert(from(mat.att_color), mat.att_dist, sh.t);
                power = power * vol;
                throughput = throughput * vol;
            }
            Ctx c = evaluate_material(sc, mat, sh.mat_index, rec, sh);
            if (depth >= RR_MIN_PHOTON_DEPTH) {
                f32 q = phx_fmaxf(throughput.x, phx_fmaxf(throughput.y, throughput.z));
                q = phx_clampf(q, RR_MIN_Q, RR_MAX_Q);
                if (random_float(rng) > q)
                    break;
                f32 iq = 1.0f / q;
                throughput = throughput * iq;
                power = power * iq;
            }
            F4 view = -dir;
            Bsdf bs = sample_bsdf(rng, c, c.shading_normal, sh.normal, view, ior, sh.front_face);
            u32 m = sh.mat_index < a.M ? sh.mat_index : 0u;
            f32 score;
            F4 ev = event_dlog(a.am[m], a.rej + (size_t)m * ACC_MU * 2, c, bs, view, c.shading_normal, score);
            if (bs.event == BSDF_DIFFUSE && dot(dir, sh.normal) < 0.0f) {
                /* the deposit carries the path so far and the choice of the
                   diffuse event here (the half-vector score) */
                F4 ad = (base + n < a.max_deposits) ? ld(a.adj + base + n) : BLACK;
                F4 lt = mk(ad.x * power.x, ad.y * power.y, ad.z * power.z, 0.0f);
                L.add_path(lt);
                L.add_local(m, lt, mk(0.0f, 0.0f, 0.0f, score));
                if (em != 0xFFFFFFFFu)
                    L.g(em, 4) += (lt.x + lt.y + lt.z) * inv_e;
                n++;
            }
            L.advance(m, ev);
            power = power * bs.throughput;
            throughput = throughput * bs.throughput;
            F4 side = (dot(bs.dir, sh.normal) > 0.0f) ? sh.normal : -sh.normal;
            origin = sh.position + bs.dir * EPS + side * EPS;
            dir = bs.dir;
        }
        if (a.chk_count)
            a.chk_count[li] = n;
    }
    flush_grads(L, a.grad);
}

/* ---------------- camera-path replay ---------------- */
struct AdjCameraArgs {
    PhxCameraParams cam;
    u32 w, h, n;
    const u32 *d_seed;
    SceneView sc;
    const AdjMaterial *am;
    const f32 *rej;
    const u32 *light_mat;
    u32 M;
    const phx_float4 *lam_rad, *lam_tau, *nf;
    const f32 *R;
    f32 *grad;
    phx_float4 *chk_emission, *chk_direct; /* optional replay check */
};

/* shadow_transmittance (light_sampling.h:302-333) with the materials of the
 * transmissive occluders kept (they scale the result by their base color) */
PHX_DEV F4 shadow_trans_adj(const SceneView &sc, F4 o, F4 d, f32 t_max, u32 *stk, u32 *occ, u32 &n_occ) {
    f32 t_min = EPS;
    F4 trans = WHITE;
    n_occ = 0;
    for (u32 depth = 0; depth < MAX_RAY_BOUNCES; depth++) {
        HitRecord rec;
        if (!scene_intersect(sc, o, d, t_min, t_max, rec, stk, ADJ_BLOCK))
            break;
        const PhxMaterial &mat = sc.materials[rec.mat_index];
        if (mat.transmission < EPS)
            return BLACK;
        F4 bcol = from(mat.base_color);
        if (mat.diff_index != NO_TEXTURE) {
            SurfaceHit s2 = process_hit(sc, rec, o, d);
            bcol = bcol * srgb_to_linear(sample_texture_uv(sc, rec.mat_index, mat.diff_index, s2.uvx, s2.uvy, 0));
        }
        if (mat.att_dist > EPS)
            trans = trans * beer_lambert(from(mat.att_color), mat.att_dist, rec.t - t_min);
        trans = trans * (bcol * mat.transmission);
        occ[n_occ * ADJ_BLOCK] = rec.mat_index;
        n_occ++;
        if (dot(trans, trans) < EPS)
            return BLACK;
        t_min = rec.t + EPS;
    }
    return trans;
}

__global__ __launch_bounds__(ADJ_BLOCK) void k_adj_camera(AdjCameraArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * ADJ_BLOCK];
    __shared__ f32 s_ior[IOR_STACK_SIZE * ADJ_BLOCK];
    __shared__ u32 s_occ[MAX_RAY_BOUNCES * ADJ_BLOCK];
    u32 t = threadIdx.x;
    AdjLane L{s_adj_dyn + t, s_adj_dyn + a.M * 4 * ADJ_BLOCK + t, a.M};
    for (u32 e = 0; e < a.M * 5; e++)
        L.G[e * ADJ_BLOCK] = 0.0f;
    L.clear_path();
    u32 tid = blockIdx.x * ADJ_BLOCK + t;
    const SceneView &sc = a.sc;
    if (tid < a.w * a.h) {
        u32 *stk = s_stack + t;
        u32 *occ = s_occ + t;
        u32 px = tid % a.w, py = tid / a.w;
        u32 rng = pcg_seed(*a.d_seed + tid);
        F4 origin, dir;
        make_ray(a.cam, px, py, a.w, a.h, rng, origin, dir);
        F4 lam_r = ld(a.lam_rad + tid);
        F4 lam_f = ld(a.lam_tau + tid) * a.R[tid];
        F4 throughput = WHITE, emission = BLACK, direct = BLACK;
        IorStack ior{s_ior + t, (u32)ADJ_BLOCK, 0u};
        init_ior_stack(ior);
        f32 inv_n = 1.0f / (f32)a.n;
        for (u32 depth = 0;;) {
            HitRecord rec;
            if (!scene_intersect(sc, origin, dir, EPS, INF, rec, stk, ADJ_BLOCK)) {
                if (sc.has_env) {
                    F4 env = throughput * sample_envmap(sc, dir);
                    emission = emission + env;
                    L.add_path(lam_r * env);
                }
                break;
            }
            SurfaceHit sh = process_hit(sc, rec, origin, dir);
            const PhxMaterial &mat = sc.materials[sh.mat_index];
            u32 m = sh.mat_index < a.M ? sh.mat_index : 0u;
            const AdjMaterial &am = a.am[m];
            f32 curr_ior = current_ior(ior);
            if (curr_ior != AIR_IOR && mat.thickness > EPS && mat.att_dist > EPS)
                throughput = throughput * beer_lambert(from(mat.att_color), mat.att_dist, sh.t);
            Ctx c = evaluate_material(sc, mat, sh.mat_index, rec, sh);
            {
                F4 te = throughput * c.emissive;
                emission = emission + te;
                F4 lt = lam_r * te;
                L.add_path(lt);
                L.g(m, 4) += (lt.x + lt.y + lt.z) * am.inv_emis;
            }
            F4 view = -dir;
            Bsdf bs = sample_bsdf(rng, c, c.shading_normal, sh.normal, view, ior, sh.front_face);
            f32 score;
            F4 ev = event_dlog(am, a.rej + (size_t)m * ACC_MU * 2, c, bs, view, c.shading_normal, score);
            F4 hit_pos = sh.position, hit_n = sh.normal, sn = c.shading_normal;
            if (bs.event != BSDF_TRANSMIT) {
                bool diffuse = bs.event == BSDF_DIFFUSE;
                F4 f0 = ZERO;
                if (bs.event == BSDF_METALLIC) {
                    f0 = c.base_color;
                } else if (bs.event == BSDF_FRESNEL) {
                    f32 ratio = (curr_ior - c.ior) / (curr_ior + c.ior);
                    f32 r0 = ratio * ratio;
                    f0 = mk(r0, r0, r0, 0.0f);
                }
                f32 alpha = c.roughness * c.roughness;
                F4 nee_sum = BLACK;
                /* per-sample terms need the sample's share: weight of the sum
                   that reaches the image, λ ⊙ (throughput·occl/n) */
                F4 lam_sum = diffuse ? lam_r * throughput * c.occlusion * inv_n : lam_r * throughput * inv_n;
                for (u32 d = 0; d < a.n; d++) {
                    LightFragment f;
                    F4 pre;
                    F4 local = mk(0.0f, 0.0f, 0.0f, 0.0f);
                    if (diffuse) {
                        if (!sample_light_fragment(rng, sc, hit_pos, sn, f)) {
                            nee_sum = nee_sum + BLACK;
                            continue;
                        }
                        pre = f.radiance * (c.base_color / PI);
                    } else {
                        f32 ndv = dot(view, sn);
                        if (ndv < EPS) {
                            nee_sum = nee_sum + BLACK;
                            continue;
                        }
                        if (!sample_light_fragment(rng, sc, hit_pos, sn, f) || f.n_dot_dir < EPS) {
                            nee_sum = nee_sum + BLACK;
                            continue;
                        }
                        F4 hv = normalize(f.dir + view);
                        f32 ndh = phx_fmaxf(dot(sn, hv), 0.0f);
                        f32 ggx = ggx_ndf(ndh, alpha);
                        F4 fr = fresnel4(f0, view, hv);
                        f32 sm = smith_g1_ggx(f.n_dot_dir, alpha) * smith_g1_ggx(ndv, alpha);
                        F4 spec = f.radiance * ggx * fr * sm / (4.0f * f.n_dot_dir * ndv);
                        if (dot(spec, spec) < EPS * EPS) {
                            nee_sum = nee_sum + BLACK;
                            continue;
                        }
                        pre = spec;
                        local.w = (dlog_ndf(ndh, alpha) + dlog_g1(f.n_dot_dir, alpha) + dlog_g1(ndv, alpha)) *
                                  am.rough_gate;
                        if (bs.event == BSDF_METALLIC) {
                            F4 df = dlog_fresnel_f0(c.base_color, view, hv);
                            local.x = df.x;
                            local.y = df.y;
                            local.z = df.z;
                        }
                    }
                    u32 n_occ = 0;
                    F4 tr = shadow_trans_adj(sc, hit_pos, f.dir, f.dist - EPS, stk, occ, n_occ);
                    F4 term = pre * tr;
                    nee_sum = nee_sum + term;
                    if (dot(term, term) > 0.0f) {
                        F4 lt = lam_sum * term;
                        L.add_local(m, lt, local);
                        u32 le = a.light_mat[f.light];
                        if (le != 0xFFFFFFFFu)
                            L.g(le, 4) += (lt.x + lt.y + lt.z) * a.am[le].inv_emis;
                        for (u32 k = 0; k < n_occ; k++) {
                            u32 om = occ[k * ADJ_BLOCK];
                            om = om < a.M ? om : 0u;
                            L.add_local(om, lt, mk(a.am[om].inv_bc[0], a.am[om].inv_bc[1], a.am[om].inv_bc[2], 0.0f));
                        }
                    }
                }
                /* the whole NEE sum carries the path and this event's choice
                   (half-vector score); diffuse NEE also scales with albedo */
                F4 lt_sum = lam_sum * nee_sum;
                L.add_path(lt_sum);
                L.add_local(m, lt_sum,
                            mk(diffuse ? am.inv_bc[0] : 0.0f, diffuse ? am.inv_bc[1] : 0.0f,
                               diffuse ? am.inv_bc[2] : 0.0f, score));
                if (diffuse) {
                    /* trace_rays.cl:126-135 */
                    emission = mk(emission.x, emission.y, emission.z, 1.0f);
                    direct = throughput * c.occlusion * (nee_sum / (f32)a.n);
                    /* the flux of this hitpoint: Φ = thr ⊙ bc/π ⊙ Σ w·power */
                    F4 phi = throughput * (c.base_color / PI) * ld(a.nf + tid);
                    F4 lt = lam_f * phi;
                    L.add_path(lt);
                    L.add_local(m, lt, mk(am.inv_bc[0], am.inv_bc[1], am.inv_bc[2], score));
                    break;
                }
                emission = emission + throughput * (nee_sum / (f32)a.n);
            }
            /* trace_rays.cl:149-160 */
            L.advance(m, ev);
            throughput = throughput * bs.throughput;
            if (depth >= RR_MIN_RAY_DEPTH) {
                f32 q = phx_fmaxf(throughput.x, phx_fmaxf(throughput.y, throughput.z));
                q = phx_clampf(q, RR_MIN_Q, RR_MAX_Q);
                if (random_float(rng) > q)
                    break;
                throughput = throughput * (1.0f / q);
            }
            F4 side = (dot(bs.dir, hit_n) > 0.0f) ? hit_n : -hit_n;
            origin = hit_pos + bs.dir * EPS + side * EPS;
            dir = bs.dir;
            depth++;
            if (depth >= MAX_RAY_BOUNCES)
                break;
        }
        if (a.chk_emission) {
            st(a.chk_emission + tid, emission);
            st(a.chk_direct + tid, direct);
        }
    }
    flush_grads(L, a.grad);
}

/* adj[i] = 0 for i < min(*d_n, max_n) */
__global__ __launch_bounds__(256) void k_adj_zero(phx_float4 *__restrict__ adj, const u32 *__restrict__ d_n,
                                                  u32 max_n) {
    u32 n = *d_n;
    n = n < max_n ? n : max_n;
    for (u32 i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        st(adj + i, BLACK);
}

const uint32_t *adjoint_emit_counts(void *emit_ws, u32 photons, u32 max_bounces) {
    return emit_layout(emit_ws, photons, max_bounces).st.count;
}

/* ---------------- one pass of the backward sweep ---------------- */
hipError_t launch_adjoint_pass(const AdjPass &P, hipStream_t s) {
    u32 n_px = P.w * P.h;
    u32 zblocks = (P.max_deposits + 255) / 256;
    zblocks = zblocks > 8192 ? 8192 : (zblocks ? zblocks : 1);
    hipLaunchKernelGGL(k_adj_ratio, dim3((n_px + 255) / 256), dim3(256), 0, s, P.R, P.ratio, n_px);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_adj_zero, dim3(zblocks), dim3(256), 0, s, P.adj_sorted, P.d_count, P.max_deposits);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    AdjGatherArgs ga{P.hp_pos,  P.hp_normal,  P.hp_throughput, P.hp_base_color, P.hp_emission, n_px,
                     P.ppos,    P.ppow,       P.pdir,          P.cell_start,    P.cell_end,    P.hash,
                     P.r2,      P.lam_tau,    P.R,             P.adj_sorted,    P.nf};
    hipLaunchKernelGGL(k_adj_gather, dim3((n_px + 255) / 256), dim3(256), 0, s, ga);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_adj_unsort, dim3(zblocks), dim3(256), 0, s, P.adj_sorted, P.perm, P.d_count,
                       P.max_deposits, P.adj_dep);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    size_t lds = (size_t)P.M * 9 * ADJ_BLOCK * sizeof(f32);
    EmitLayout EL = emit_layout(P.emit_ws, P.photons, P.max_bounces);
    AdjPhotonArgs pa{P.sc,      P.photons,   P.max_bounces, P.d_seed,       P.am,   P.rej,      P.light_mat,
                     P.M,       EL.offs,     P.adj_dep,     P.max_deposits, P.grad, P.chk_count};
    hipLaunchKernelGGL(k_adj_photons, dim3((P.photons + ADJ_BLOCK - 1) / ADJ_BLOCK), dim3(ADJ_BLOCK), lds, s, pa);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    AdjCameraArgs ca{P.cam, P.w,       P.h,       P.n_direct, P.d_seed, P.sc,   P.am,           P.rej,
                     P.light_mat, P.M, P.lam_rad, P.lam_tau,  P.nf,     P.R,    P.grad,         P.chk_emission,
                     P.chk_direct};
    hipLaunchKernelGGL(k_adj_camera, dim3((n_px + ADJ_BLOCK - 1) / ADJ_BLOCK), dim3(ADJ_BLOCK), lds, s, ca);
    return hipGetLastError();
}

} // namespace phx
