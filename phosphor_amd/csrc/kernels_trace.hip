/*
 * kernels_trace.hip — camera-ray tracing to hitpoints and photon emission.
 *
 * trace_rays (kernels/trace_rays.cl:14-169): one lane per pixel.  The
 * reference interleaves four traversal call sites (camera ray, diffuse NEE
 * shadow rays, metallic and Fresnel specular NEE shadow rays); here one loop
 * drives a single inlined traversal and a small state machine decides what
 * the next ray is, so the kernel carries one copy of the traversal code and
 * one LDS stack.  The sequence of random draws, the accumulation order of
 * the NEE sums and every expression are those of the reference.
 *
 * trace_rays also has a wavefront form (camera paths, queued shadow rays,
 * in-order resolve) with the same outputs; the pass driver uses it.
 *
 * emit_photons (kernels/emit_photons.cl:12-114): one launch per bounce over
 * a queue of live photons; the k-th deposit of photon i goes to a depth-major
 * staging slot [k][i] and the per-photon deposit count is stored; a final
 * kernel compacts the staging area in (photon, bounce) order after an
 * exclusive scan of the counts, which makes the deposit order canonical
 * instead of atomic_inc arrival order.
 */
#include "phx_device.h"

namespace phx {

constexpr int TRACE_BLOCK = 128;

struct TraceArgs {
    PhxCameraParams cam;
    u32 w, h, seed, direct_samples;
    SceneView sc;
    phx_float4 *hp_pos, *hp_normal, *hp_throughput, *hp_base_color, *hp_direct, *hp_emission;
};

__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_rays(TraceArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * TRACE_BLOCK];
    __shared__ f32 s_ior[IOR_STACK_SIZE * TRACE_BLOCK];
    const SceneView &sc = a.sc;
    u32 n_pixels = a.w * a.h;
    u32 tid = blockIdx.x * TRACE_BLOCK + threadIdx.x;
    if (tid >= n_pixels)
        return;
    u32 *stk = s_stack + threadIdx.x;

    u32 px = tid % a.w, py = tid / a.w;
    u32 rng = pcg_seed(a.seed + tid);
    F4 origin, dir;
    make_ray(a.cam, px, py, a.w, a.h, rng, origin, dir);

    F4 position = ZERO, normal = ZERO, throughput = WHITE, base_color = BLACK, direct = BLACK, emission = BLACK;
    IorStack ior{s_ior + threadIdx.x, (u32)TRACE_BLOCK, 0u};
    init_ior_stack(ior, AIR_IOR);

    u32 depth = 0;
    /* state of the current hit (kept while its NEE loop runs) */
    bool in_nee = false, nee_diffuse = false;
    u32 nee_i = 0;
    F4 nee_sum = BLACK;
    F4 hit_pos = ZERO, hit_n = ZERO, sn = ZERO, view = ZERO, bc = ZERO, f0 = ZERO;
    f32 rough = 0.0f, occl = 1.0f;
    Bsdf bs;
    bs.dir = ZERO;
    bs.throughput = ZERO;
    bs.event = BSDF_DIFFUSE;
    /* state of the current shadow ray (light_sampling.h:302-333) */
    F4 sh_dir = ZERO, sh_pre = ZERO, sh_trans = WHITE;
    f32 sh_tmin = EPS, sh_tmax = 0.0f;
    u32 sh_depth = 0;

    for (;;) {
        F4 ro = in_nee ? hit_pos : origin;
        F4 rd = in_nee ? sh_dir : dir;
        f32 tmin = in_nee ? sh_tmin : EPS;
        f32 tmax = in_nee ? sh_tmax : INF;
        HitRecord rec;
        bool hit = scene_intersect(sc, ro, rd, tmin, tmax, rec, stk, TRACE_BLOCK);
        bool next_sample = false;
        bool advance = false;

        if (!in_nee) {
            if (!hit) {
                if (sc.has_env)
                    emission = emission + throughput * sample_envmap(sc, dir);
                break;
            }
            SurfaceHit sh = process_hit(sc, rec, origin, dir);
            const PhxMaterial &mat = sc.materials[sh.mat_index];
            f32 curr_ior = current_ior(ior);
            if (curr_ior != AIR_IOR && mat.thickness > EPS && mat.att_dist > EPS)
                throughput = throughput * beer_lambert(from(mat.att_color), mat.att_dist, sh.t);
            Ctx c = evaluate_material(sc, mat, sh.mat_index, rec, sh);
            emission = emission + throughput * c.emissive;
            view = -dir;
            bs = sample_bsdf(rng, c, c.shading_normal, sh.normal, view, ior, sh.front_face);
            hit_pos = sh.position;
            hit_n = sh.normal;
            if (bs.event == BSDF_TRANSMIT) {
                advance = true;
            } else {
                in_nee = true;
                nee_diffuse = bs.event == BSDF_DIFFUSE;
                nee_i = 0;
                nee_sum = BLACK;
                sn = c.shading_normal;
                bc = c.base_color;
                rough = c.roughness;
                occl = c.occlusion;
                if (bs.event == BSDF_METALLIC) {
                    f0 = c.base_color;
                } else if (bs.event == BSDF_FRESNEL) {
                    f32 ratio = (curr_ior - c.ior) / (curr_ior + c.ior);
                    f32 r0 = ratio * ratio;
                    f0 = mk(r0, r0, r0, 0.0f);
                }
                next_sample = true;
            }
        } else {
            /* one step of shadow_transmittance */
            bool finished = false;
            F4 tr_res = sh_trans;
            if (!hit) {
                finished = true;
            } else {
                const PhxMaterial &m = sc.materials[rec.mat_index];
                if (m.transmission < EPS) {
                    tr_res = BLACK;
                    finished = true;
                } else {
                    F4 bcol = from(m.base_color);
                    if (m.diff_index != NO_TEXTURE) {
                        SurfaceHit s2 = process_hit(sc, rec, hit_pos, sh_dir);
                        bcol = bcol * srgb_to_linear(sample_texture_uv(sc, rec.mat_index, m.diff_index, s2.uvx,
                                                                       s2.uvy, 0));
                    }
                    if (m.att_dist > EPS)
                        sh_trans = sh_trans * beer_lambert(from(m.att_color), m.att_dist, rec.t - sh_tmin);
                    sh_trans = sh_trans * (bcol * m.transmission);
                    if (dot(sh_trans, sh_trans) < EPS) {
                        tr_res = BLACK;
                        finished = true;
                    } else {
                        sh_tmin = rec.t + EPS;
                        sh_depth++;
                        if (sh_depth >= MAX_RAY_BOUNCES) {
                            tr_res = sh_trans;
                            finished = true;
                        }
                    }
                }
            }
            if (finished) {
                nee_sum = nee_sum + sh_pre * tr_res;
                nee_i++;
                next_sample = true;
            }
        }

        if (next_sample) {
            /* direct_lighting_{diffuse,specular} up to the shadow ray */
            bool launched = false;
            while (nee_i < a.direct_samples) {
                LightFragment f;
                if (nee_diffuse) {
                    if (!sample_light_fragment(rng, sc, hit_pos, sn, f)) {
                        nee_sum = nee_sum + BLACK;
                        nee_i++;
                        continue;
                    }
                    sh_pre = f.radiance * (bc / PI);
                } else {
                    f32 ndv = dot(view, sn);
                    if (ndv < EPS) {
                        nee_sum = nee_sum + BLACK;
                        nee_i++;
                        continue;
                    }
                    if (!sample_light_fragment(rng, sc, hit_pos, sn, f) || f.n_dot_dir < EPS) {
                        nee_sum = nee_sum + BLACK;
                        nee_i++;
                        continue;
                    }
                    f32 alpha = rough * rough;
                    F4 hv = normalize(f.dir + view);
                    f32 ndh = phx_fmaxf(dot(sn, hv), 0.0f);
                    f32 ggx = ggx_ndf(ndh, alpha);
                    F4 fr = fresnel4(f0, view, hv);
                    f32 sm = smith_g1_ggx(f.n_dot_dir, alpha) * smith_g1_ggx(ndv, alpha);
                    F4 spec = f.radiance * ggx * fr * sm / (4.0f * f.n_dot_dir * ndv);
                    if (dot(spec, spec) < EPS * EPS) {
                        nee_sum = nee_sum + BLACK;
                        nee_i++;
                        continue;
                    }
                    sh_pre = spec;
                }
                sh_dir = f.dir;
                sh_tmin = EPS;
                sh_tmax = f.dist - EPS;
                sh_trans = WHITE;
                sh_depth = 0;
                launched = true;
                break;
            }
            if (!launched) {
                in_nee = false;
                if (nee_diffuse) {
                    emission = mk(emission.x, emission.y, emission.z, 1.0f);
                    position = hit_pos;
                    normal = hit_n;
                    base_color = bc;
                    direct = throughput * occl * (nee_sum / (f32)a.direct_samples);
                    break;
                }
                emission = emission + throughput * (nee_sum / (f32)a.direct_samples);
                advance = true;
            }
        }

        if (advance) {
            /* trace_rays.cl:149-160 */
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
    }

    st(a.hp_emission + tid, emission);
    st(a.hp_pos + tid, position);
    st(a.hp_normal + tid, normal);
    st(a.hp_throughput + tid, throughput);
    st(a.hp_base_color + tid, base_color);
    st(a.hp_direct + tid, direct);
}

hipError_t launch_trace(const PhxCameraParams &cam, u32 w, u32 h, u32 seed, u32 direct_samples, const SceneArgs &sc,
                        phx_float4 *hp_pos, phx_float4 *hp_normal, phx_float4 *hp_throughput,
                        phx_float4 *hp_base_color, phx_float4 *hp_direct, phx_float4 *hp_emission, hipStream_t s) {
    u32 n = w * h;
    if (n == 0)
        return hipSuccess;
    TraceArgs a;
    a.cam = cam;
    a.w = w;
    a.h = h;
    a.seed = seed;
    a.direct_samples = direct_samples;
    a.sc = sc;
    a.hp_pos = hp_pos;
    a.hp_normal = hp_normal;
    a.hp_throughput = hp_throughput;
    a.hp_base_color = hp_base_color;
    a.hp_direct = hp_direct;
    a.hp_emission = hp_emission;
    u32 blocks = (n + TRACE_BLOCK - 1) / TRACE_BLOCK;
    hipLaunchKernelGGL(k_trace_rays, dim3(blocks), dim3(TRACE_BLOCK), 0, s, a);
    return hipGetLastError();
}

/* ----------------------------------------------------------------------
 * Wavefront form of trace_rays (same outputs, bit for bit).
 *
 * The random draws of the NEE loops (light selection, emissive point) never
 * depend on shadow visibility, and the Russian roulette after a specular NEE
 * only reads throughput, so a camera-path kernel can draw all light fragments
 * of an NEE event and keep walking the path.  Shadow rays go to a queue (the
 * direct_samples rays of one event are contiguous: same origin, coherent
 * directions) traced by a lean kernel; a resolve kernel then forms
 * emission and direct exactly in the reference's order of additions:
 *   emission = ((BLACK + a_0) + s_0) + a_1 + ... (+ envmap on a miss),
 * a_k = throughput_k·emissive_k, s_k = throughput_k·(Σ_d pre_d·trans_d / n)
 * for specular events, and direct = (throughput·occlusion)·(Σ / n) at the
 * first diffuse event (trace_rays.cl:97-147).
 * ---------------------------------------------------------------------- */
constexpr int WF_CAM_BLOCK = 128;
constexpr int WF_SHADOW_BLOCK = 256;
constexpr u32 NEE_NONE = 0, NEE_DIFFUSE = 1, NEE_SPECULAR = 2;

struct WfState {
    phx_float4 *rec_a, *rec_thr; /* [pixel · 5 + bounce] */
    u32 *rec_base;               /* first shadow slot of the NEE event of that bounce */
    u32 *meta;                   /* per pixel: bits 0-2 records, 3-12 NEE kinds, 15 envmap term */
    phx_float4 *env;             /* per pixel: throughput · envmap on a miss */
    f32 *occl;                   /* per pixel: occlusion at the diffuse hit */
    u32 *counter;                /* shadow slots handed out */
    phx_float4 *sh_o;            /* origin.xyz, t_max in w */
    phx_float4 *sh_d;            /* direction */
    phx_float4 *sh_pre;          /* NEE weight; w = 0 valid, w = -1 no shadow ray (adds BLACK) */
    phx_float4 *sh_tr;           /* shadow transmittance */
    u32 cap;
};

struct WfCamArgs {
    PhxCameraParams cam;
    u32 w, h, seed, n;
    SceneView sc;
    WfState st;
    phx_float4 *hp_pos, *hp_normal, *hp_throughput, *hp_base_color;
};

__global__ __launch_bounds__(WF_CAM_BLOCK) void k_trace_camera(WfCamArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * WF_CAM_BLOCK];
    __shared__ f32 s_ior[IOR_STACK_SIZE * WF_CAM_BLOCK];
    const SceneView &sc = a.sc;
    const WfState &S = a.st;
    u32 n_pixels = a.w * a.h;
    u32 tid = blockIdx.x * WF_CAM_BLOCK + threadIdx.x;
    bool live = tid < n_pixels;
    u32 *stk = s_stack + threadIdx.x;
    u32 px = live ? tid % a.w : 0, py = live ? tid / a.w : 0;
    u32 rng = pcg_seed(a.seed + tid);
    F4 origin = ZERO, dir = ZERO;
    if (live)
        make_ray(a.cam, px, py, a.w, a.h, rng, origin, dir);
    F4 throughput = WHITE;
    IorStack ior{s_ior + threadIdx.x, (u32)WF_CAM_BLOCK, 0u};
    init_ior_stack(ior, AIR_IOR);
    u32 depth = 0, kinds = 0, env_flag = 0;
    bool diffuse = false;
    size_t rbase = (size_t)tid * MAX_RAY_BOUNCES;
    /* the loop runs while any lane of the wave is live, so the wave-wide slot
       allocation below sees every lane */
    bool active = live;
    while (__any(active)) {
        bool nee = false, spec = false;
        F4 hit_pos = ZERO, sn = ZERO, bc = ZERO, f0 = ZERO, view = ZERO;
        f32 rough = 0.0f;
        Bsdf bs;
        bs.dir = ZERO;
        bs.throughput = ZERO;
        bs.event = BSDF_TRANSMIT;
        F4 hit_n = ZERO;
        if (active) {
            HitRecord rec;
            bool hit = scene_intersect(sc, origin, dir, EPS, INF, rec, stk, WF_CAM_BLOCK);
            if (!hit) {
                if (sc.has_env) {
                    st(S.env + tid, throughput * sample_envmap(sc, dir));
                    env_flag = 1;
                }
                active = false;
            } else {
                SurfaceHit sh = process_hit(sc, rec, origin, dir);
                const PhxMaterial &mat = sc.materials[sh.mat_index];
                f32 curr_ior = current_ior(ior);
                if (curr_ior != AIR_IOR && mat.thickness > EPS && mat.att_dist > EPS)
                    throughput = throughput * beer_lambert(from(mat.att_color), mat.att_dist, sh.t);
                Ctx c = evaluate_material(sc, mat, sh.mat_index, rec, sh);
                st(S.rec_a + rbase + depth, throughput * c.emissive);
                view = -dir;
                bs = sample_bsdf(rng, c, c.shading_normal, sh.normal, view, ior, sh.front_face);
                hit_pos = sh.position;
                hit_n = sh.normal;
                sn = c.shading_normal;
                bc = c.base_color;
                rough = c.roughness;
                if (bs.event == BSDF_DIFFUSE) {
                    nee = true;
                    diffuse = true;
                    kinds |= NEE_DIFFUSE << (2 * depth);
                    st(a.hp_pos + tid, sh.position);
                    st(a.hp_normal + tid, sh.normal);
                    st(a.hp_base_color + tid, c.base_color);
                    st(a.hp_throughput + tid, throughput);
                    S.occl[tid] = c.occlusion;
                } else if (bs.event == BSDF_METALLIC || bs.event == BSDF_FRESNEL) {
                    nee = true;
                    spec = true;
                    kinds |= NEE_SPECULAR << (2 * depth);
                    if (bs.event == BSDF_METALLIC) {
                        f0 = c.base_color;
                    } else {
                        f32 ratio = (curr_ior - c.ior) / (curr_ior + c.ior);
                        f32 r0 = ratio * ratio;
                        f0 = mk(r0, r0, r0, 0.0f);
                    }
                }
                if (nee)
                    st(S.rec_thr + rbase + depth, throughput);
            }
        }
        /* one atomic per wave for the shadow slots of this bounce */
        unsigned long long want = __ballot(nee);
        u32 lane = threadIdx.x & 63u;
        u32 rank = (u32)__popcll(want & ((1ull << lane) - 1ull));
        u32 wave_base = 0;
        if (want) {
            u32 total = (u32)__popcll(want) * a.n;
            u32 first = (u32)__ffsll((long long)want) - 1u;
            if (lane == first)
                wave_base = atomicAdd(S.counter, total);
            wave_base = __shfl(wave_base, (int)first, 64);
        }
        if (nee) {
            u32 base = wave_base + rank * a.n;
            S.rec_base[rbase + depth] = base;
            for (u32 d = 0; d < a.n; d++) {
                LightFragment f;
                bool valid = false;
                F4 pre = BLACK;
                if (!spec) {
                    if (sample_light_fragment(rng, sc, hit_pos, sn, f)) {
                        valid = true;
                        pre = f.radiance * (bc / PI);
                    }
                } else {
                    f32 ndv = dot(view, sn);
                    /* direct_lighting_specular: each early return adds BLACK */
                    if (!(ndv < EPS) && sample_light_fragment(rng, sc, hit_pos, sn, f) && !(f.n_dot_dir < EPS)) {
                        f32 alpha = rough * rough;
                        F4 hv = normalize(f.dir + view);
                        f32 ndh = phx_fmaxf(dot(sn, hv), 0.0f);
                        f32 ggx = ggx_ndf(ndh, alpha);
                        F4 fr = fresnel4(f0, view, hv);
                        f32 sm = smith_g1_ggx(f.n_dot_dir, alpha) * smith_g1_ggx(ndv, alpha);
                        F4 spv = f.radiance * ggx * fr * sm / (4.0f * f.n_dot_dir * ndv);
                        if (!(dot(spv, spv) < EPS * EPS)) {
                            valid = true;
                            pre = spv;
                        }
                    }
                }
                size_t slot = (size_t)base + d;
                if (valid) {
                    st(S.sh_o + slot, mk(hit_pos.x, hit_pos.y, hit_pos.z, f.dist - EPS));
                    st(S.sh_d + slot, f.dir);
                    st(S.sh_pre + slot, mk(pre.x, pre.y, pre.z, 0.0f));
                } else {
                    st(S.sh_pre + slot, mk(0.0f, 0.0f, 0.0f, -1.0f));
                }
            }
        }
        if (diffuse)
            active = false;
        if (active) {
            /* trace_rays.cl:149-160 */
            throughput = throughput * bs.throughput;
            bool killed = false;
            if (depth >= RR_MIN_RAY_DEPTH) {
                f32 q = phx_fmaxf(throughput.x, phx_fmaxf(throughput.y, throughput.z));
                q = phx_clampf(q, RR_MIN_Q, RR_MAX_Q);
                if (random_float(rng) > q)
                    killed = true;
                else
                    throughput = throughput * (1.0f / q);
            }
            if (!killed) {
                F4 side = (dot(bs.dir, hit_n) > 0.0f) ? hit_n : -hit_n;
                origin = hit_pos + bs.dir * EPS + side * EPS;
                dir = bs.dir;
            }
            depth++;
            if (killed || depth >= MAX_RAY_BOUNCES)
                active = false;
        }
    }
    if (!live)
        return;
    if (!diffuse) {
        st(a.hp_pos + tid, ZERO);
        st(a.hp_normal + tid, ZERO);
        st(a.hp_base_color + tid, BLACK);
        st(a.hp_throughput + tid, throughput);
    }
    /* records exist for every bounce that hit: 0..depth at a diffuse end,
       0..depth-1 after a miss (nothing recorded for the missing bounce), a
       roulette kill or the bounce limit (depth already advanced) */
    u32 nrec = diffuse ? depth + 1 : depth;
    S.meta[tid] = nrec | (kinds << 3) | (env_flag << 15);
}

struct WfShadowArgs {
    SceneView sc;
    WfState st;
    u32 any_hit;
};

template <bool ANY>
__device__ __forceinline__ void shadow_one(const WfShadowArgs &a, u32 slot, u32 *stk) {
    const SceneView &sc = a.sc;
    const WfState &S = a.st;
    F4 pre = ld(S.sh_pre + slot);
    if (pre.w != 0.0f)
        return;
    F4 ot = ld(S.sh_o + slot);
    F4 o = mk(ot.x, ot.y, ot.z, 0.0f);
    F4 d = ld(S.sh_d + slot);
    f32 t_min = EPS, t_max = ot.w;
    F4 trans = WHITE;
    /* light_sampling.h:302-333 */
    for (u32 depth = 0; depth < MAX_RAY_BOUNCES; depth++) {
        HitRecord rec;
        if (!traverse<ANY>(sc, o, d, t_min, t_max, rec, stk, WF_SHADOW_BLOCK))
            break;
        const PhxMaterial &mat = sc.materials[rec.mat_index];
        if (ANY || mat.transmission < EPS) {
            trans = BLACK;
            break;
        }
        F4 bcol = from(mat.base_color);
        if (mat.diff_index != NO_TEXTURE) {
            SurfaceHit s2 = process_hit(sc, rec, o, d);
            bcol = bcol * srgb_to_linear(sample_texture_uv(sc, rec.mat_index, mat.diff_index, s2.uvx, s2.uvy, 0));
        }
        if (mat.att_dist > EPS)
            trans = trans * beer_lambert(from(mat.att_color), mat.att_dist, rec.t - t_min);
        trans = trans * (bcol * mat.transmission);
        if (dot(trans, trans) < EPS) {
            trans = BLACK;
            break;
        }
        t_min = rec.t + EPS;
    }
    st(S.sh_tr + slot, trans);
}

__global__ __launch_bounds__(WF_SHADOW_BLOCK) void k_trace_shadow(WfShadowArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * WF_SHADOW_BLOCK];
    u32 *stk = s_stack + threadIdx.x;
    u32 count = *a.st.counter;
    for (u32 slot = blockIdx.x * WF_SHADOW_BLOCK + threadIdx.x; slot < count; slot += gridDim.x * WF_SHADOW_BLOCK) {
        if (a.any_hit)
            shadow_one<true>(a, slot, stk);
        else
            shadow_one<false>(a, slot, stk);
    }
}

struct WfResolveArgs {
    WfState st;
    u32 n_pixels, n;
    phx_float4 *hp_direct, *hp_emission;
};

__global__ __launch_bounds__(256) void k_trace_resolve(WfResolveArgs a) {
    u32 tid = blockIdx.x * 256 + threadIdx.x;
    if (tid >= a.n_pixels)
        return;
    const WfState &S = a.st;
    u32 meta = S.meta[tid];
    u32 nrec = meta & 7u;
    size_t rbase = (size_t)tid * MAX_RAY_BOUNCES;
    F4 e = BLACK;
    F4 direct = BLACK;
    for (u32 k = 0; k < nrec; k++) {
        e = e + ld(S.rec_a + rbase + k);
        u32 kind = (meta >> (3 + 2 * k)) & 3u;
        if (kind == NEE_NONE)
            continue;
        size_t base = S.rec_base[rbase + k];
        F4 sum = BLACK;
        for (u32 d = 0; d < a.n; d++) {
            F4 pre = ld(S.sh_pre + base + d);
            if (pre.w != 0.0f) {
                sum = sum + BLACK;
            } else {
                sum = sum + pre * ld(S.sh_tr + base + d);
            }
        }
        F4 thr = ld(S.rec_thr + rbase + k);
        if (kind == NEE_SPECULAR) {
            e = e + thr * (sum / (f32)a.n);
        } else {
            e = mk(e.x, e.y, e.z, 1.0f);
            direct = thr * S.occl[tid] * (sum / (f32)a.n);
        }
    }
    if ((meta >> 15) & 1u)
        e = e + ld(S.env + tid);
    st(a.hp_emission + tid, e);
    st(a.hp_direct + tid, direct);
}

static WfState wf_layout(void *ws, u32 n_pixels, u32 n) {
    size_t rec = (size_t)n_pixels * MAX_RAY_BOUNCES;
    size_t cap = rec * n;
    char *p = (char *)ws;
    WfState S;
    S.counter = (u32 *)p;
    p += 256;
    S.rec_a = (phx_float4 *)p;
    p += align_up(rec * sizeof(phx_float4), 256);
    S.rec_thr = (phx_float4 *)p;
    p += align_up(rec * sizeof(phx_float4), 256);
    S.rec_base = (u32 *)p;
    p += align_up(rec * sizeof(u32), 256);
    S.meta = (u32 *)p;
    p += align_up((size_t)n_pixels * sizeof(u32), 256);
    S.env = (phx_float4 *)p;
    p += align_up((size_t)n_pixels * sizeof(phx_float4), 256);
    S.occl = (f32 *)p;
    p += align_up((size_t)n_pixels * sizeof(f32), 256);
    S.sh_o = (phx_float4 *)p;
    p += align_up(cap * sizeof(phx_float4), 256);
    S.sh_d = (phx_float4 *)p;
    p += align_up(cap * sizeof(phx_float4), 256);
    S.sh_pre = (phx_float4 *)p;
    p += align_up(cap * sizeof(phx_float4), 256);
    S.sh_tr = (phx_float4 *)p;
    p += align_up(cap * sizeof(phx_float4), 256);
    S.cap = (u32)(cap > 0xFFFFFFFFull ? 0xFFFFFFFFull : cap);
    return S;
}

size_t trace_wavefront_workspace_bytes(u32 n_pixels, u32 n) {
    size_t rec = (size_t)n_pixels * MAX_RAY_BOUNCES;
    size_t cap = rec * n;
    return 256 + 2 * align_up(rec * sizeof(phx_float4), 256) + align_up(rec * sizeof(u32), 256) +
           align_up((size_t)n_pixels * sizeof(u32), 256) + align_up((size_t)n_pixels * sizeof(phx_float4), 256) +
           align_up((size_t)n_pixels * sizeof(f32), 256) + 4 * align_up(cap * sizeof(phx_float4), 256);
}

hipError_t launch_trace_wavefront(const PhxCameraParams &cam, u32 w, u32 h, u32 seed, u32 direct_samples,
                                  const SceneArgs &sc, u32 shadow_any_hit, phx_float4 *hp_pos,
                                  phx_float4 *hp_normal, phx_float4 *hp_throughput, phx_float4 *hp_base_color,
                                  phx_float4 *hp_direct, phx_float4 *hp_emission, void *ws, hipStream_t s) {
    u32 n = w * h;
    if (n == 0)
        return hipSuccess;
    WfState S = wf_layout(ws, n, direct_samples);
    hipError_t e = launch_fill_u32(S.counter, 0u, 1, s);
    if (e != hipSuccess)
        return e;
    WfCamArgs ca;
    ca.cam = cam;
    ca.w = w;
    ca.h = h;
    ca.seed = seed;
    ca.n = direct_samples;
    ca.sc = sc;
    ca.st = S;
    ca.hp_pos = hp_pos;
    ca.hp_normal = hp_normal;
    ca.hp_throughput = hp_throughput;
    ca.hp_base_color = hp_base_color;
    hipLaunchKernelGGL(k_trace_camera, dim3((n + WF_CAM_BLOCK - 1) / WF_CAM_BLOCK), dim3(WF_CAM_BLOCK), 0, s, ca);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    WfShadowArgs sa;
    sa.sc = sc;
    sa.st = S;
    sa.any_hit = shadow_any_hit;
    size_t expect = (size_t)n * direct_samples;
    u32 blocks = (u32)((expect + WF_SHADOW_BLOCK - 1) / WF_SHADOW_BLOCK);
    if (blocks > 16384)
        blocks = 16384;
    hipLaunchKernelGGL(k_trace_shadow, dim3(blocks), dim3(WF_SHADOW_BLOCK), 0, s, sa);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    WfResolveArgs ra;
    ra.st = S;
    ra.n_pixels = n;
    ra.n = direct_samples;
    ra.hp_direct = hp_direct;
    ra.hp_emission = hp_emission;
    hipLaunchKernelGGL(k_trace_resolve, dim3((n + 255) / 256), dim3(256), 0, s, ra);
    return hipGetLastError();
}

/* Moves photon li's deposits to base + offset[li] + k (base = the counter
 * value before this batch), dropping those at or beyond max_photons. */
__global__ __launch_bounds__(256) void k_emit_compact(const phx_float4 *__restrict__ st_pos,
                                                      const u32 *__restrict__ st_power, const u32 *__restrict__ st_dir,
                                                      const u32 *__restrict__ count, const u32 *__restrict__ offs,
                                                      u32 photons, const u32 *__restrict__ counter, u32 max_photons,
                                                      phx_float4 *__restrict__ pos, u32 *__restrict__ power,
                                                      u32 *__restrict__ dirs) {
    u32 li = blockIdx.x * 256 + threadIdx.x;
    if (li >= photons)
        return;
    u32 c = count[li];
    uint64_t dst = (uint64_t)*counter + offs[li];
    for (u32 k = 0; k < c; k++, dst++) {
        if (dst >= max_photons)
            break;
        size_t slot = (size_t)k * photons + li;
        pos[dst] = st_pos[slot];
        power[dst] = st_power[slot];
        dirs[dst] = st_dir[slot];
    }
}

__global__ void k_counter_add(u32 *counter, const u32 *total) { *counter += *total; }

/* ----------------------------------------------------------------------
 * Wavefront form of emit_photons (same deposits, bit for bit).  Bounce k of
 * every live photon runs in one launch over a queue of live photons, so the
 * lanes of a wave stay busy however the path lengths differ; a photon's state
 * (origin, direction, power, throughput, PCG state, IOR stack, deposit count)
 * lives in HBM between bounces.  Queue order does not matter: every photon
 * writes only its own state and its own deposit slots [k][photon].
 * ---------------------------------------------------------------------- */
constexpr int EMIT_WF_BLOCK = 128;
constexpr u32 EMIT_WF_BOUNCES = 3; /* bounce-synchronous launches before the finishing launch */

struct EmitWfState {
    phx_float4 *origin, *dir, *power, *thr;
    u32 *rng, *ior_cur, *count;
    f32 *ior; /* [depth][photon] */
    u32 *queue[2];
    u32 *qcount; /* [bounce] live photons entering bounce k */
};

struct EmitWfArgs {
    SceneView sc;
    u32 offset, photons, max_bounces, seed, depth;
    EmitWfState st;
    phx_float4 *st_pos;
    u32 *st_power, *st_dir;
};

__global__ __launch_bounds__(256) void k_emit_wf_start(EmitWfArgs a) {
    u32 li = blockIdx.x * 256 + threadIdx.x;
    if (li >= a.photons)
        return;
    const EmitWfState &S = a.st;
    u32 rng = pcg_seed(a.seed + li + a.offset);
    F4 origin, dir, power;
    sample_light(rng, a.sc, origin, dir, power);
    st(S.origin + li, origin);
    st(S.dir + li, dir);
    st(S.power + li, power);
    st(S.thr + li, mk(1.0f, 1.0f, 1.0f, 0.0f));
    S.rng[li] = rng;
    S.ior_cur[li] = 0;
    S.ior[li] = AIR_IOR;
    S.count[li] = 0;
}

/* One bounce of emit_photons.cl:49-113 on a photon's state; returns false
 * when the path ends (escape, roulette or the bounce limit). */
__device__ __forceinline__ bool emit_step(const EmitWfArgs &a, u32 li, u32 depth, F4 &origin, F4 &dir, F4 &power,
                                          F4 &throughput, u32 &rng, IorStack &ior, u32 &n, u32 *stk) {
    const SceneView &sc = a.sc;
    HitRecord rec;
    if (!scene_intersect(sc, origin, dir, EPS, INF, rec, stk, EMIT_WF_BLOCK))
        return false;
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
            return false;
        f32 iq = 1.0f / q;
        throughput = throughput * iq;
        power = power * iq;
    }
    F4 view = -dir;
    Bsdf bs = sample_bsdf(rng, c, c.shading_normal, sh.normal, view, ior, sh.front_face);
    if (bs.event == BSDF_DIFFUSE && dot(dir, sh.normal) < 0.0f) {
        size_t slot = (size_t)n * a.photons + li;
        st(a.st_pos + slot, sh.position);
        a.st_power[slot] = encode_rgbe(power);
        a.st_dir[slot] = encode_oct(dir);
        n++;
    }
    power = power * bs.throughput;
    throughput = throughput * bs.throughput;
    F4 side = (dot(bs.dir, sh.normal) > 0.0f) ? sh.normal : -sh.normal;
    origin = sh.position + bs.dir * EPS + side * EPS;
    dir = bs.dir;
    return depth + 1 < a.max_bounces;
}

__global__ __launch_bounds__(EMIT_WF_BLOCK) void k_emit_wf_bounce(EmitWfArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * EMIT_WF_BLOCK];
    u32 *stk = s_stack + threadIdx.x;
    const EmitWfState &S = a.st;
    u32 depth = a.depth;
    u32 n_live = depth == 0 ? a.photons : S.qcount[depth];
    const u32 *qin = S.queue[depth & 1];
    u32 *qout = S.queue[(depth + 1) & 1];
    u32 stride = gridDim.x * EMIT_WF_BLOCK;
    u32 rounds = (n_live + stride - 1) / stride; /* same trip count for every lane of the grid */
    u32 lane = threadIdx.x & 63u;
    for (u32 r = 0; r < rounds; r++) {
        if (r * stride + blockIdx.x * EMIT_WF_BLOCK >= n_live)
            break; /* the whole block is past the live photons */
        u32 qi = r * stride + blockIdx.x * EMIT_WF_BLOCK + threadIdx.x;
        bool keep = false;
        u32 li = 0;
        if (qi < n_live) {
            li = depth == 0 ? qi : qin[qi];
            F4 origin = ld(S.origin + li), dir = ld(S.dir + li);
            F4 power = ld(S.power + li), throughput = ld(S.thr + li);
            u32 rng = S.rng[li];
            IorStack ior{S.ior + li, a.photons, S.ior_cur[li]};
            u32 n = S.count[li], n0 = n;
            keep = emit_step(a, li, depth, origin, dir, power, throughput, rng, ior, n, stk);
            if (n != n0)
                S.count[li] = n;
            if (keep) {
                st(S.origin + li, origin);
                st(S.dir + li, dir);
                st(S.power + li, power);
                st(S.thr + li, throughput);
                S.rng[li] = rng;
                S.ior_cur[li] = ior.cur;
            }
        }
        /* append survivors to the next queue, one atomic per wave */
        unsigned long long m = __ballot(keep);
        if (m) {
            u32 first = (u32)__ffsll((long long)m) - 1u;
            u32 base = 0;
            if (lane == first)
                base = atomicAdd(&S.qcount[depth + 1], (u32)__popcll(m));
            base = __shfl(base, (int)first, 64);
            if (keep)
                qout[base + (u32)__popcll(m & ((1ull << lane) - 1ull))] = li;
        }
    }
}

/* Runs every photon still live at bounce a.depth to the end of its path in
 * one launch, so the long tail of the path-length distribution costs one
 * launch instead of one per bounce. */
__global__ __launch_bounds__(EMIT_WF_BLOCK) void k_emit_wf_finish(EmitWfArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * EMIT_WF_BLOCK];
    u32 *stk = s_stack + threadIdx.x;
    const EmitWfState &S = a.st;
    u32 n_live = a.depth == 0 ? a.photons : S.qcount[a.depth];
    const u32 *qin = S.queue[a.depth & 1];
    for (u32 qi = blockIdx.x * EMIT_WF_BLOCK + threadIdx.x; qi < n_live; qi += gridDim.x * EMIT_WF_BLOCK) {
        u32 li = a.depth == 0 ? qi : qin[qi];
        F4 origin = ld(S.origin + li), dir = ld(S.dir + li);
        F4 power = ld(S.power + li), throughput = ld(S.thr + li);
        u32 rng = S.rng[li];
        IorStack ior{S.ior + li, a.photons, S.ior_cur[li]};
        u32 n = S.count[li], n0 = n;
        for (u32 depth = a.depth; depth < a.max_bounces; depth++)
            if (!emit_step(a, li, depth, origin, dir, power, throughput, rng, ior, n, stk))
                break;
        if (n != n0)
            S.count[li] = n;
    }
}

static EmitWfState emit_wf_layout(char *&p, u32 photons, u32 max_bounces) {
    EmitWfState S;
    size_t n = photons;
    S.origin = (phx_float4 *)p;
    p += align_up(n * 16, 256);
    S.dir = (phx_float4 *)p;
    p += align_up(n * 16, 256);
    S.power = (phx_float4 *)p;
    p += align_up(n * 16, 256);
    S.thr = (phx_float4 *)p;
    p += align_up(n * 16, 256);
    S.rng = (u32 *)p;
    p += align_up(n * 4, 256);
    S.ior_cur = (u32 *)p;
    p += align_up(n * 4, 256);
    S.count = (u32 *)p;
    p += align_up(n * 4, 256);
    S.ior = (f32 *)p;
    p += align_up(n * IOR_STACK_SIZE * 4, 256);
    S.queue[0] = (u32 *)p;
    p += align_up(n * 4, 256);
    S.queue[1] = (u32 *)p;
    p += align_up(n * 4, 256);
    S.qcount = (u32 *)p;
    p += align_up(((size_t)max_bounces + 2) * 4, 256);
    return S;
}

size_t emit_workspace_bytes(u32 photons, u32 max_bounces) {
    size_t slots = (size_t)photons * max_bounces;
    size_t n = photons;
    size_t b = 0;
    b += align_up(slots * sizeof(phx_float4), 256); /* staging pos */
    b += align_up(slots * sizeof(u32), 256);        /* staging power */
    b += align_up(slots * sizeof(u32), 256);        /* staging dir */
    b += align_up(n * sizeof(u32), 256);            /* offsets */
    b += 256;                                       /* total */
    b += align_up(scan_workspace_bytes(photons), 256);
    /* wavefront state */
    b += 4 * align_up(n * 16, 256) + 5 * align_up(n * 4, 256) + align_up(n * IOR_STACK_SIZE * 4, 256) +
         align_up(((size_t)max_bounces + 2) * 4, 256);
    return b;
}

hipError_t launch_emit(const SceneArgs &sc, u32 max_photons, u32 offset, u32 photons, u32 max_bounces, u32 seed,
                       phx_float4 *photon_pos, u32 *photon_power, u32 *photon_dir, u32 *n_photons, void *ws,
                       hipStream_t s) {
    if (photons == 0)
        return hipSuccess;
    size_t slots = (size_t)photons * max_bounces;
    char *p = (char *)ws;
    phx_float4 *st_pos = (phx_float4 *)p;
    p += align_up(slots * sizeof(phx_float4), 256);
    u32 *st_power = (u32 *)p;
    p += align_up(slots * sizeof(u32), 256);
    u32 *st_dir = (u32 *)p;
    p += align_up(slots * sizeof(u32), 256);
    u32 *offs = (u32 *)p;
    p += align_up((size_t)photons * sizeof(u32), 256);
    u32 *total = (u32 *)p;
    p += 256;
    void *scan_ws = p;
    p += align_up(scan_workspace_bytes(photons), 256);
    EmitWfState S = emit_wf_layout(p, photons, max_bounces);

    hipError_t e = launch_fill_u32(S.qcount, 0u, (size_t)max_bounces + 2, s);
    if (e != hipSuccess)
        return e;
    EmitWfArgs a;
    a.sc = sc;
    a.offset = offset;
    a.photons = photons;
    a.max_bounces = max_bounces;
    a.seed = seed;
    a.depth = 0;
    a.st = S;
    a.st_pos = st_pos;
    a.st_power = st_power;
    a.st_dir = st_dir;
    hipLaunchKernelGGL(k_emit_wf_start, dim3((photons + 255) / 256), dim3(256), 0, s, a);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    /* enough blocks to fill the GPU; bounces with few live photons retire
       the surplus blocks at once */
    u32 blocks = (photons + EMIT_WF_BLOCK - 1) / EMIT_WF_BLOCK;
    if (blocks > 4096)
        blocks = 4096;
    /* bounce-synchronous launches while most photons are alive, then one
       launch for the rest of every path */
    u32 wave_bounces = max_bounces < EMIT_WF_BOUNCES ? max_bounces : EMIT_WF_BOUNCES;
    for (u32 depth = 0; depth < wave_bounces; depth++) {
        a.depth = depth;
        hipLaunchKernelGGL(k_emit_wf_bounce, dim3(blocks), dim3(EMIT_WF_BLOCK), 0, s, a);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    if (wave_bounces < max_bounces) {
        a.depth = wave_bounces;
        hipLaunchKernelGGL(k_emit_wf_finish, dim3(blocks), dim3(EMIT_WF_BLOCK), 0, s, a);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    e = launch_scan(S.count, photons, offs, nullptr, total, scan_ws, s);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_emit_compact, dim3((photons + 255) / 256), dim3(256), 0, s, st_pos, st_power, st_dir,
                       S.count, offs, photons, n_photons, max_photons, photon_pos, photon_power, photon_dir);
    e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_counter_add, dim3(1), dim3(1), 0, s, n_photons, total);
    return hipGetLastError();
}

} // namespace phx
