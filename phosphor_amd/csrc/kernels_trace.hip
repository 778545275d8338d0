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
 * in-order resolve; bodies in pass_bodies.h) with the same outputs; the pass
 * driver uses it.
 *
 * emit_photons (kernels/emit_photons.cl:12-114): launches over a queue of
 * live photons (bodies in pass_bodies.h); the k-th deposit of photon i goes
 * to a depth-major staging slot [k][i] and the per-photon deposit count is
 * stored; a final kernel compacts the staging area in (photon, bounce) order
 * after an exclusive scan of the counts, which makes the deposit order
 * canonical instead of atomic_inc arrival order.
 *
 * Seeds: every launcher takes the round seed by value or, for passes that
 * run without host synchronisation (graph replay), a device pointer to it.
 */
#include "pass_bodies.h"

namespace phx {

constexpr int TRACE_BLOCK = 128;

struct TraceArgs {
    PhxCameraParams cam;
    u32 w, h, seed, direct_samples;
    const u32 *d_seed; /* when set, the seed is read on the device */
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
    u32 rng = pcg_seed((a.d_seed ? *a.d_seed : a.seed) + tid);
    F4 origin, dir;
    make_ray(a.cam, px, py, a.w, a.h, rng, origin, dir);

    F4 position = ZERO, normal = ZERO, throughput = WHITE, base_color = BLACK, direct = BLACK, emission = BLACK;
    IorStack ior{s_ior + threadIdx.x, (u32)TRACE_BLOCK, 0u};
    init_ior_stack(ior);

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
    bs.h = ZERO;
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
                        phx_float4 *hp_base_color, phx_float4 *hp_direct, phx_float4 *hp_emission, hipStream_t s,
                        const u32 *d_seed) {
    u32 n = w * h;
    if (n == 0)
        return hipSuccess;
    TraceArgs a;
    a.cam = cam;
    a.w = w;
    a.h = h;
    a.seed = seed;
    a.d_seed = d_seed;
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

/* ---------------- wavefront trace (bodies in pass_bodies.h) ---------------- */
/* camera waves over 8×8 pixel tiles (log2 of the tile width; 64 / width
   rows): atrium 5.91 -> 5.81 ms/pass against 64-pixel rows, caustic within
   noise; 4×16 and 16×4 tiles within noise of 8×8 */
constexpr u32 CAM_TW_LOG2 = 3;
struct WfCamArgs {
    PhxCameraParams cam;
    u32 w, h, seed, n;
    u32 tiles_x, tw_log2; /* camera tiles (2^tw_log2 × 64 / 2^tw_log2 pixels) per image row */
    const u32 *d_seed;
    SceneView sc;
    WfState st;
    CamOut o;
};

/* B: threads per block (64, 128 or 256; the LDS stacks are laid out
   [depth][thread] with stride B) */
/* SPLIT: paths that go on after the first bounce are queued for
   k_trace_camera_rest; otherwise each lane follows its path to the end */
template <u32 B, bool SPLIT>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(4))) void k_trace_camera(WfCamArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * B];
    __shared__ f32 s_ior[IOR_STACK_SIZE * B];
    /* one wave per 8×8 pixel tile: the primary rays of a wave leave the
       camera in a narrow cone, so its lanes visit much the same nodes */
    u32 wave = blockIdx.x * (B / 64u) + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
    u32 tw = 1u << a.tw_log2;
    u32 x = (wave % a.tiles_x) * tw + (lane & (tw - 1u)), y = (wave / a.tiles_x) * (64u >> a.tw_log2) + (lane >> a.tw_log2);
    bool live = x < a.w && y < a.h;
    u32 pix = live ? y * a.w + x : 0u;
    u32 seed = a.d_seed ? *a.d_seed : a.seed;
    u32 slots = single_delta_light(a.sc) ? 1u : a.n;
    IorStack ior{s_ior + threadIdx.x, B, 0u};
    CamPath p = camera_path_begin(a.cam, a.w, a.h, seed, live, pix, ior);
    if (!SPLIT) {
        camera_path_run(a.sc, a.n, pix, pix, 0u, a.st, a.o, s_stack + threadIdx.x, B, slots, p, ior, MAX_RAY_BOUNCES);
        camera_path_end(a.st, a.o, live, pix, pix, p);
        return;
    }
    /* The first bounce here; the paths that go on (specular events, a few
       in a hundred on diffuse scenes) continue in k_trace_camera_rest,
       packed densely into waves, instead of holding a wave for the few
       lanes whose paths are still going. */
    camera_path_run(a.sc, a.n, pix, pix, 0u, a.st, a.o, s_stack + threadIdx.x, B, slots, p, ior, 1u);
    unsigned long long m = __ballot(p.active);
    if (m) {
        u32 first = (u32)__ffsll((long long)m) - 1u;
        u32 base = 0;
        if (lane == first)
            base = atomicAdd(a.st.cq_count, (u32)__popcll(m));
        base = __shfl(base, (int)first, 64);
        if (p.active) {
            u32 q = base + (u32)__popcll(m & ((1ull << lane) - 1ull));
            a.st.cq_pix[q] = pix;
            camera_path_save(a.st, q, p, ior);
        }
    }
    if (!p.active)
        camera_path_end(a.st, a.o, live, pix, pix, p);
}

/* The camera paths k_trace_camera queued after their first bounce, 64 to a
 * wave, followed to their ends. */
template <u32 B>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(4))) void k_trace_camera_rest(WfCamArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * B];
    __shared__ f32 s_ior[IOR_STACK_SIZE * B];
    const u32 count = *a.st.cq_count;
    const u32 lane = threadIdx.x & 63u, waves = gridDim.x * (B / 64u);
    u32 slots = single_delta_light(a.sc) ? 1u : a.n;
    for (u32 base = (blockIdx.x * (B / 64u) + (threadIdx.x >> 6)) * 64u; base < count; base += waves * 64u) {
        u32 q = base + lane;
        bool live = q < count;
        IorStack ior{s_ior + threadIdx.x, B, 0u};
        CamPath p;
        u32 pix = 0;
        if (live) {
            pix = a.st.cq_pix[q];
            p = camera_path_load(a.st, q, ior);
        } else {
            p.active = false;
        }
        camera_path_run(a.sc, a.n, pix, pix, 0u, a.st, a.o, s_stack + threadIdx.x, B, slots, p, ior, MAX_RAY_BOUNCES);
        camera_path_end(a.st, a.o, live, pix, pix, p);
    }
}

struct WfShadowArgs {
    SceneView sc;
    WfState st;
    u32 any_hit;
    u32 group; /* n when n divides 64 (each event's sum formed in the wave), else 0 */
};

/* six waves per SIMD (76 VGPRs, no spills) */
template <u32 B>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(6))) void k_trace_shadow(WfShadowArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * B];
    u32 *stk = s_stack + threadIdx.x;
    u32 count = *a.st.counter;
    bool shared = single_delta_light(a.sc); /* one slot per event (camera_path) */
    for (u32 slot = blockIdx.x * B + threadIdx.x; slot < count; slot += gridDim.x * B) {
        F4 v = a.any_hit ? shadow_value<true, B>(a.sc, a.st, slot, stk) : shadow_value<false, B>(a.sc, a.st, slot, stk);
        if (a.group && !shared) {
            /* count is a multiple of n: an event's slots pass this loop together */
            F4 sum = group_sample_sum(v, a.group);
            if ((slot & (a.group - 1u)) == 0u)
                st(a.st.sh_tr + slot, sum);
        } else {
            st(a.st.sh_tr + slot, v);
        }
    }
}

struct WfResolveArgs {
    SceneView sc;
    WfState st;
    u32 n_pixels, n, summed;
    phx_float4 *hp_direct, *hp_emission;
};

__global__ __launch_bounds__(256) void k_trace_resolve(WfResolveArgs a) {
    u32 tid = blockIdx.x * 256 + threadIdx.x;
    if (tid >= a.n_pixels)
        return;
    resolve_pixel(a.st, a.n, tid, tid, a.hp_direct, a.hp_emission, a.summed != 0u, single_delta_light(a.sc));
}

WfState wf_layout(void *ws, u32 n_pixels, u32 n, bool with_scene) {
    size_t rec = (size_t)n_pixels * MAX_RAY_BOUNCES;
    size_t cap = rec * n;
    char *p = (char *)ws;
    WfState S;
    S.n_paths = n_pixels;
    S.counter = (u32 *)p;
    p += 256;
    /* the camera queue's count shares the counter block (one fill clears
       both), 128 bytes from the shadow-slot counter */
    S.cq_count = S.counter + 32;
    S.cq_pix = nullptr;
    S.cq_state = nullptr;
    if (!with_scene) {
        S.cq_pix = (u32 *)p;
        p += align_up((size_t)n_pixels * sizeof(u32), 256);
        S.cq_state = (phx_float4 *)p;
        p += align_up((size_t)n_pixels * 4u * sizeof(phx_float4), 256);
    }
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
    phx_float4 **ev[5] = {&S.ev_pos, &S.ev_n, &S.ev_view, &S.ev_col, &S.ev_misc};
    for (phx_float4 **a : ev) {
        *a = (phx_float4 *)p;
        p += align_up(rec * sizeof(phx_float4), 256);
    }
    S.sh_rng = (u32 *)p;
    p += align_up(cap * sizeof(u32), 256);
    S.sh_ev = (u32 *)p;
    p += align_up(cap * sizeof(u32), 256);
    S.sh_tr = (phx_float4 *)p;
    p += align_up(cap * sizeof(phx_float4), 256);
    S.sh_scene = nullptr;
    if (with_scene) {
        S.sh_scene = (u32 *)p;
        p += align_up(cap * sizeof(u32), 256);
    }
    S.cap = (u32)(cap > 0xFFFFFFFFull ? 0xFFFFFFFFull : cap);
    return S;
}

size_t trace_wavefront_workspace_bytes(u32 n_pixels, u32 n, bool with_scene) {
    size_t rec = (size_t)n_pixels * MAX_RAY_BOUNCES;
    size_t cap = rec * n;
    return 256 +
           (with_scene ? 0
                       : align_up((size_t)n_pixels * sizeof(u32), 256) +
                             align_up((size_t)n_pixels * 4u * sizeof(phx_float4), 256)) +
           2 * align_up(rec * sizeof(phx_float4), 256) + align_up(rec * sizeof(u32), 256) +
           align_up((size_t)n_pixels * sizeof(u32), 256) + align_up((size_t)n_pixels * sizeof(phx_float4), 256) +
           align_up((size_t)n_pixels * sizeof(f32), 256) + 5 * align_up(rec * sizeof(phx_float4), 256) +
           2 * align_up(cap * sizeof(u32), 256) + align_up(cap * sizeof(phx_float4), 256) +
           (with_scene ? align_up(cap * sizeof(u32), 256) : 0);
}

hipError_t launch_trace_wavefront(const PhxCameraParams &cam, u32 w, u32 h, u32 seed, u32 direct_samples,
                                  const SceneArgs &sc, u32 shadow_any_hit, phx_float4 *hp_pos,
                                  phx_float4 *hp_normal, phx_float4 *hp_throughput, phx_float4 *hp_base_color,
                                  phx_float4 *hp_direct, phx_float4 *hp_emission, void *ws, hipStream_t s,
                                  const u32 *d_seed, const LaunchShape *shape) {
    u32 n = w * h;
    if (n == 0)
        return hipSuccess;
    WfState S = wf_layout(ws, n, direct_samples, false);
    /* paths that go on after their first bounce in a kernel of their own */
    const bool split = !shape || shape->split_trace;
    if (!split) {
        S.cq_pix = nullptr;
        S.cq_state = nullptr;
    }
    hipError_t e = launch_fill_u32(S.counter, 0u, 64u, s);
    if (e != hipSuccess)
        return e;
    WfCamArgs ca;
    ca.cam = cam;
    ca.w = w;
    ca.h = h;
    ca.seed = seed;
    ca.n = direct_samples;
    ca.d_seed = d_seed;
    ca.sc = sc;
    ca.st = S;
    ca.o = CamOut{hp_pos, hp_normal, hp_throughput, hp_base_color};
    ca.tw_log2 = CAM_TW_LOG2;
    ca.tiles_x = (w + (1u << ca.tw_log2) - 1u) >> ca.tw_log2;
    u32 cb = shape ? shape->camera : (u32)WF_CAM_BLOCK;
    cb = cb == 64 || cb == 256 ? cb : 128u;
    u32 th = 64u >> ca.tw_log2;
    u32 tiles = ca.tiles_x * ((h + th - 1u) / th), per = cb / 64u;
    dim3 cg((tiles + per - 1) / per);
    void (*cam_kernel)(WfCamArgs);
    if (cb == 64)
        cam_kernel = split ? k_trace_camera<64, true> : k_trace_camera<64, false>;
    else if (cb == 256)
        cam_kernel = split ? k_trace_camera<256, true> : k_trace_camera<256, false>;
    else
        cam_kernel = split ? k_trace_camera<128, true> : k_trace_camera<128, false>;
    hipLaunchKernelGGL(cam_kernel, cg, dim3(cb), 0, s, ca);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    if (split) {
        /* at most 8192 waves, each taking 64 queued paths at a time */
        u32 rb = (u32)(((size_t)n + cb - 1) / cb), cap = 8192u * 64u / cb;
        rb = rb < cap ? rb : cap;
        if (cb == 64)
            hipLaunchKernelGGL(k_trace_camera_rest<64>, dim3(rb), dim3(64), 0, s, ca);
        else if (cb == 256)
            hipLaunchKernelGGL(k_trace_camera_rest<256>, dim3(rb), dim3(256), 0, s, ca);
        else
            hipLaunchKernelGGL(k_trace_camera_rest<128>, dim3(rb), dim3(128), 0, s, ca);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    WfShadowArgs sa;
    sa.sc = sc;
    sa.st = S;
    sa.any_hit = shadow_any_hit;
    /* every slot allocation is a multiple of direct_samples (one per event) */
    sa.group = (direct_samples > 0 && 64u % direct_samples == 0u) ? direct_samples : 0u;
    size_t expect = (size_t)n * direct_samples;
    u32 sb = shape ? shape->shadow : (u32)WF_SHADOW_BLOCK;
    u32 blocks = (u32)((expect + sb - 1) / sb);
    if (blocks > 16384 * 128 / sb)
        blocks = 16384 * 128 / sb;
    if (sb == 64)
        hipLaunchKernelGGL(k_trace_shadow<64>, dim3(blocks), dim3(64), 0, s, sa);
    else if (sb == 256)
        hipLaunchKernelGGL(k_trace_shadow<256>, dim3(blocks), dim3(256), 0, s, sa);
    else
        hipLaunchKernelGGL(k_trace_shadow<128>, dim3(blocks), dim3(128), 0, s, sa);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    WfResolveArgs ra;
    ra.sc = sc;
    ra.st = S;
    ra.n_pixels = n;
    ra.n = direct_samples;
    ra.summed = sa.group ? 1u : 0u;
    ra.hp_direct = hp_direct;
    ra.hp_emission = hp_emission;
    hipLaunchKernelGGL(k_trace_resolve, dim3((n + 255) / 256), dim3(256), 0, s, ra);
    return hipGetLastError();
}

/* ---------------- photon emission ---------------- */
/* Moves photon li's deposits to base + offset[li] + k (base = the counter
 * value before this batch), dropping those at or beyond max_photons: to the
 * reference's three arrays, or, with `rec`, to 32-byte deposit records
 * (position; packed power, octahedral direction, 0, 0), which the sort's
 * last pass reads with one access per record. */
__global__ __launch_bounds__(256) void k_emit_compact(const phx_float4 *__restrict__ st_pos,
                                                      const u32 *__restrict__ st_power, const u32 *__restrict__ st_dir,
                                                      const u32 *__restrict__ count, const u32 *__restrict__ offs,
                                                      u32 photons, const u32 *__restrict__ counter, u32 max_photons,
                                                      phx_float4 *__restrict__ pos, u32 *__restrict__ power,
                                                      u32 *__restrict__ dirs, phx_float4 *__restrict__ rec) {
    u32 li = blockIdx.x * 256 + threadIdx.x;
    if (li >= photons)
        return;
    u32 c = count[li];
    uint64_t dst = (uint64_t)*counter + offs[li];
    for (u32 k = 0; k < c; k++, dst++) {
        if (dst >= max_photons)
            break;
        size_t slot = (size_t)k * photons + li;
        if (rec) {
            rec[2 * dst] = st_pos[slot];
            st(rec + 2 * dst + 1, mk(phx_as_float(st_power[slot]), phx_as_float(st_dir[slot]), 0.0f, 0.0f));
        } else {
            pos[dst] = st_pos[slot];
            power[dst] = st_power[slot];
            dirs[dst] = st_dir[slot];
        }
    }
}

__global__ void k_counter_add(u32 *counter, const u32 *total) { *counter += *total; }

struct EmitWfArgs {
    SceneView sc;
    u32 offset, photons, max_bounces, seed, depth;
    const u32 *d_seed;
    EmitWfState st;
    EmitStaging g;
};

/* Grid-stride over the photons with at most EMIT_START_BLOCKS blocks: in a
 * pipelined pass this kernel starts while the previous pass's gather holds
 * most of the GPU, and fewer blocks that each take several photons need
 * fewer free slots than one short block per 256 photons (atrium, 4096
 * blocks: 6.54 -> 5.83 ms/pass in the run where the unbounded grid was
 * slowest; 256 or 1024 blocks no better). */
constexpr u32 EMIT_START_BLOCKS = 4096;
__global__ __launch_bounds__(256) void k_emit_wf_start(EmitWfArgs a) {
    u32 seed = a.d_seed ? *a.d_seed : a.seed;
    for (u32 li = blockIdx.x * 256 + threadIdx.x; li < a.photons; li += gridDim.x * 256)
        emit_start_one(a.sc, seed, li + a.offset, li, a.st, a.photons);
}

/* Per-entry kernels walk the sub-queues with a grid that is a multiple of
 * EMIT_NQ: block b serves sub-queue b mod EMIT_NQ, the blocks of one
 * sub-queue stride over its entries, and survivors are appended to the same
 * sub-queue of the next bounce.  Every loop bound is uniform over a block. */
struct EmitQueueWalk {
    u32 k, first, step;
};
PHX_DEV EmitQueueWalk emit_queue_walk(u32 B) {
    EmitQueueWalk w;
    w.k = blockIdx.x % EMIT_NQ;
    w.first = (blockIdx.x / EMIT_NQ) * B;
    w.step = (gridDim.x / EMIT_NQ) * B;
    return w;
}

/* Bounce a.depth of every live photon: one lane per queue entry, survivors
 * appended to the next queue.  Queue order does not matter: a photon writes
 * only its own state and its own staging slots. */
/* five waves per SIMD (96 VGPRs and 44 bytes of spills per lane) measured
   faster than four without spills */
template <u32 B>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(5))) void k_emit_wf_bounce(EmitWfArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * B];
    u32 *stk = s_stack + threadIdx.x;
    const EmitWfState &S = a.st;
    const u32 depth = a.depth;
    const EmitQueueWalk w = emit_queue_walk(B);
    const u32 n = emit_qlen(S, depth, w.k, a.photons);
    u32 *cnt_out = emit_qcounter(S.qcount, depth + 1, w.k);
    u32 *qout = S.queue[(depth + 1) & 1] + (size_t)w.k * S.slice;
    for (u32 base = w.first; base < n; base += w.step) {
        u32 i = base + threadIdx.x;
        bool keep = false;
        u32 li = 0;
        if (i < n) {
            li = emit_qentry(S, depth, w.k, i);
            keep = emit_advance<B>(a.sc, S, a.g, li, depth, a.max_bounces, true, stk);
        }
        wave_append(keep, li, cnt_out, qout);
    }
}

/* Runs every photon still live at bounce a.depth to the end of its path in
 * one launch, so the long tail of the path-length distribution costs one
 * launch instead of one per bounce. */
template <u32 B>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(5))) void k_emit_wf_finish(EmitWfArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * B];
    u32 *stk = s_stack + threadIdx.x;
    const EmitWfState &S = a.st;
    const EmitQueueWalk w = emit_queue_walk(B);
    const u32 n = emit_qlen(S, a.depth, w.k, a.photons);
    for (u32 i = w.first + threadIdx.x; i < n; i += w.step) {
        u32 li = emit_qentry(S, a.depth, w.k, i);
        if (S.dense)
            emit_advance<B>(a.sc, S, a.g, li, a.depth, a.max_bounces, false, stk, &S.ps[a.depth & 1],
                            (size_t)w.k * S.slice + i);
        else
            emit_advance<B>(a.sc, S, a.g, li, a.depth, a.max_bounces, false, stk);
    }
}

/* Split bounce, first half: the closest hit of every live photon's ray on
 * the pair tree.  Persistent waves, each serving sub-queue (wave index mod
 * EMIT_NQ), take entries from that sub-queue's ticket EMIT_TRACE_CHUNK at a
 * time into a wave-private pool, and a lane whose search ends stores its hit
 * and takes the next entry of the pool once EMIT_TRACE_REFILL lanes of its
 * wave are idle: photon rays need from a handful to hundreds of steps, and a
 * wave that waited for its slowest ray would leave most lanes idle (measured
 * on the atrium: 17.5 steps per ray, 99 per wave).  Every wave ends: its
 * ticket only grows, a wave stops fetching once the ticket has passed the
 * sub-queue's length, and every search ends. */
constexpr u32 EMIT_TRACE_CHUNK = 64;
#ifndef EMIT_TRACE_REFILL
#define EMIT_TRACE_REFILL 16u
#endif
/* a round of triangle tests runs when waiting lanes × weight exceed those
   waiting for a pair step × theirs */
#ifndef EMIT_TRACE_TRI_WEIGHT
#define EMIT_TRACE_TRI_WEIGHT 1u
#endif
#ifndef EMIT_TRACE_NODE_WEIGHT
#define EMIT_TRACE_NODE_WEIGHT 1u
#endif

/* The persistent search loop of k_emit_trace: the calling
 * wave serves one sub-queue of n entries whose ticket is `ticket`; src.ray(i,
 * o, d) gives entry i's ray, src.hit(i, h) stores its hit (t, u, v, position
 * in the tree's triangle order, EMIT_MISS for none).  Every lane of the wave
 * calls it. */
template <u32 B, typename Src>
PHX_DEV void replace_trace(const SceneView &sc, u32 n, u32 *ticket, u32 *stk, Src &src) {
    const u32 lane = threadIdx.x & 63u;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    PairSearch sr;
    bool act = false;
    u32 qi = 0;
    /* the wave's pool [pnext, pend) of entries; dry: the ticket has passed n
       (both wave-uniform) */
    u32 pnext = 0, pend = 0;
    bool dry = n == 0;
    for (;;) {
        unsigned long long idle = __ballot(!act);
        u32 cnt = (u32)__popcll(idle);
        if (cnt >= EMIT_TRACE_REFILL && (pnext < pend || !dry)) {
            if (pnext >= pend) {
                u32 base = 0;
                u32 first = (u32)__ffsll((long long)idle) - 1u;
                if (lane == first)
                    base = atomicAdd(ticket, EMIT_TRACE_CHUNK);
                base = __builtin_amdgcn_readfirstlane(__shfl(base, (int)first, 64));
                if (base >= n) {
                    dry = true;
                } else {
                    pnext = base;
                    pend = n - base < EMIT_TRACE_CHUNK ? n : base + EMIT_TRACE_CHUNK;
                    if (pend == n)
                        dry = true;
                }
            }
            u32 rank = (u32)__popcll(idle & lt_mask);
            u32 avail = pend - pnext;
            if (!act && rank < avail) {
                qi = pnext + rank;
                F4 o, d;
                src.ray(qi, o, d);
                pair_search_begin(sr, sc, o, d, EPS, INF);
                act = true;
            }
            pnext += cnt < avail ? cnt : avail;
        }
        /* one round of pair steps or one round of triangle tests, whichever
           more lanes are waiting for */
        const bool at_leaf = act && (sr.ref & PHX_PAIR_LEAF) != 0;
        const u32 n_act = (u32)__popcll(__ballot(act));
        if (n_act == 0) {
            /* every lane idle: done once the ticket and the pool are used up */
            if (dry && pnext >= pend)
                break;
            continue;
        }
        const u32 n_tri = (u32)__popcll(__ballot(at_leaf));
        const bool tri_round = n_tri * EMIT_TRACE_TRI_WEIGHT > (n_act - n_tri) * EMIT_TRACE_NODE_WEIGHT;
        bool more = true;
        if (tri_round) {
            if (at_leaf)
                more = pair_search_tri(sr, sc, stk, B);
        } else if (act && !at_leaf) {
            more = pair_search_node(sr, sc, stk, B);
        }
        if (act && !more) {
            src.hit(qi, mk(sr.rec.t, sr.rec.u, sr.rec.v, phx_as_float(sr.found ? sr.rec.tri_index : EMIT_MISS)));
            act = false;
        }
    }
}

struct EmitRaySrc {
    const EmitWfState &S;
    u32 depth, k;
    phx_float4 *hits;
    /* the split bounces keep the ray state dense: entry qi's at slot
       k·slice + qi of set depth & 1 */
    __device__ void ray(u32 qi, F4 &o, F4 &d) const {
        const PhotonSet &ps = S.ps[depth & 1];
        size_t at = (size_t)k * S.slice + qi;
        o = ld(ps.origin + at);
        d = ld(ps.dir + at);
    }
    __device__ void hit(u32 qi, F4 h) const { st(hits + qi, h); }
};

template <u32 B> __global__ __launch_bounds__(B) void k_emit_trace(EmitWfArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * B];
    const EmitWfState &S = a.st;
    const u32 k = (blockIdx.x * (B / 64u) + (threadIdx.x >> 6)) % EMIT_NQ;
    EmitRaySrc src{S, a.depth, k, S.hit + (size_t)k * S.slice};
    replace_trace<B>(a.sc, emit_qlen(S, a.depth, k, a.photons), emit_qcounter(S.ticket, a.depth, k),
                     s_stack + threadIdx.x, src);
}

/* Split bounce, second half: one lane per queue entry shades its photon at
 * the hit k_emit_trace found, stores the deposit and the new state, and
 * appends survivors to the next queue (emit_after_hit). */
/* five waves per SIMD (96 VGPRs) against four (120 VGPRs): atrium 5.72 ->
   5.67 ms/pass, caustic within noise, when the cap still cost 28 bytes of
   spills per lane; with the photon state read in queue order it spills
   nothing */
template <u32 B>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(5))) void k_emit_shade(EmitWfArgs a) {
    const EmitWfState &S = a.st;
    const u32 depth = a.depth;
    const EmitQueueWalk w = emit_queue_walk(B);
    const u32 n = emit_qlen(S, depth, w.k, a.photons);
    u32 *cnt_out = emit_qcounter(S.qcount, depth + 1, w.k);
    u32 *qout = S.queue[(depth + 1) & 1] + (size_t)w.k * S.slice;
    const phx_float4 *hits = S.hit + (size_t)w.k * S.slice;
    const size_t row = (size_t)w.k * S.slice;
    const PhotonSet &in = S.ps[depth & 1], &out = S.ps[(depth + 1) & 1];
    for (u32 base = w.first; base < n; base += w.step) {
        u32 i = base + threadIdx.x;
        bool keep = false;
        u32 li = 0;
        PhotonRay r;
        if (i < n) {
            li = emit_qentry(S, depth, w.k, i);
            keep = emit_after_hit(a.sc, S, a.g, li, depth, a.max_bounces, ld(hits + i), in, row + i, r);
        }
        /* the survivor's state goes to the slot of its entry in the next
           queue */
        u32 pos = wave_append(keep, li, cnt_out, qout);
        if (keep) {
            size_t at = row + pos;
            st(out.origin + at, r.origin);
            st(out.dir + at, r.dir);
            st(out.power + at, r.power);
            st(out.thr + at, r.thr);
            out.rng[at] = r.rng;
            out.ior_cur[at] = r.ior_cur;
        }
    }
}

EmitWfState emit_wf_layout(char *&p, u32 photons, u32 max_bounces) {
    EmitWfState S;
    size_t n = photons;
    S.slice = (u32)((n + EMIT_NQ - 1) / EMIT_NQ);
    size_t qn = (size_t)S.slice * EMIT_NQ;
    /* two sets of ray state, one entry per queue position */
    for (int k = 0; k < 2; k++) {
        S.ps[k].origin = (phx_float4 *)p;
        p += align_up(qn * 16, 256);
        S.ps[k].dir = (phx_float4 *)p;
        p += align_up(qn * 16, 256);
        S.ps[k].power = (phx_float4 *)p;
        p += align_up(qn * 16, 256);
        S.ps[k].thr = (phx_float4 *)p;
        p += align_up(qn * 16, 256);
        S.ps[k].rng = (u32 *)p;
        p += align_up(qn * 4, 256);
        S.ps[k].ior_cur = (u32 *)p;
        p += align_up(qn * 4, 256);
    }
    S.origin = S.ps[0].origin;
    S.dir = S.ps[0].dir;
    S.power = S.ps[0].power;
    S.thr = S.ps[0].thr;
    S.rng = S.ps[0].rng;
    S.ior_cur = S.ps[0].ior_cur;
    S.dense = 0;
    S.count = (u32 *)p;
    p += align_up(n * 4, 256);
    S.ior = (f32 *)p;
    p += align_up(n * IOR_STACK_SIZE * 4, 256);
    S.queue[0] = (u32 *)p;
    p += align_up(qn * 4, 256);
    S.queue[1] = (u32 *)p;
    p += align_up(qn * 4, 256);
    S.qcount = (u32 *)p;
    p += align_up(emit_counter_words(max_bounces) * 4, 256);
    S.hit = (phx_float4 *)p;
    p += align_up(qn * 16, 256);
    S.ticket = (u32 *)p;
    p += align_up(emit_counter_words(max_bounces) * 4, 256);
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
    size_t qn = (n + EMIT_NQ - 1) / EMIT_NQ * EMIT_NQ;
    b += 2 * (4 * align_up(qn * 16, 256) + 2 * align_up(qn * 4, 256)); /* two sets of ray state */
    b += align_up(n * 4, 256) + align_up(n * IOR_STACK_SIZE * 4, 256);
    b += 2 * align_up(qn * 4, 256) + align_up(qn * 16, 256);              /* queues, hits */
    b += 2 * align_up(emit_counter_words(max_bounces) * 4, 256);        /* counters, tickets */
    return b;
}

EmitLayout emit_layout(void *ws, u32 photons, u32 max_bounces) {
    size_t slots = (size_t)photons * max_bounces;
    char *p = (char *)ws;
    EmitLayout L;
    L.g.pos = (phx_float4 *)p;
    p += align_up(slots * sizeof(phx_float4), 256);
    L.g.power = (u32 *)p;
    p += align_up(slots * sizeof(u32), 256);
    L.g.dir = (u32 *)p;
    p += align_up(slots * sizeof(u32), 256);
    L.g.stride = photons;
    L.offs = (u32 *)p;
    p += align_up((size_t)photons * sizeof(u32), 256);
    L.total = (u32 *)p;
    p += 256;
    L.scan_ws = p;
    p += align_up(scan_workspace_bytes(photons), 256);
    L.st = emit_wf_layout(p, photons, max_bounces);
    return L;
}

hipError_t launch_emit_compact(const EmitLayout &L, u32 photons, u32 max_photons, phx_float4 *photon_pos,
                               u32 *photon_power, u32 *photon_dir, u32 *n_photons, hipStream_t s,
                               phx_float4 *deposit_rec) {
    hipError_t e = launch_scan(L.st.count, photons, L.offs, nullptr, L.total, L.scan_ws, s);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_emit_compact, dim3((photons + 255) / 256), dim3(256), 0, s, L.g.pos, L.g.power, L.g.dir,
                       L.st.count, L.offs, photons, n_photons, max_photons, photon_pos, photon_power, photon_dir,
                       deposit_rec);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_counter_add, dim3(1), dim3(1), 0, s, n_photons, L.total);
    return hipGetLastError();
}

hipError_t launch_emit(const SceneArgs &sc, u32 max_photons, u32 offset, u32 photons, u32 max_bounces, u32 seed,
                       phx_float4 *photon_pos, u32 *photon_power, u32 *photon_dir, u32 *n_photons, void *ws,
                       hipStream_t s, const u32 *d_seed, u32 wave_bounces, const LaunchShape *shape,
                       phx_float4 *deposit_rec) {
    if (photons == 0)
        return hipSuccess;
    EmitLayout L = emit_layout(ws, photons, max_bounces);
    const EmitWfState &S = L.st;
    hipError_t e = launch_fill_u32(S.qcount, 0u, emit_counter_words(max_bounces), s);
    if (e != hipSuccess)
        return e;
    /* split bounces need the pair tree (the reference tree keeps the fused
       bounce kernel) */
    const bool split = sc.near_first != 0 && (!shape || shape->split_trace);
    if (split && (e = launch_fill_u32(S.ticket, 0u, emit_counter_words(max_bounces), s)) != hipSuccess)
        return e;
    EmitWfArgs a;
    a.sc = sc;
    a.offset = offset;
    a.photons = photons;
    a.max_bounces = max_bounces;
    a.seed = seed;
    a.d_seed = d_seed;
    a.depth = 0;
    a.st = S;
    a.st.dense = split ? 1u : 0u; /* split bounces move each survivor's state to its queue slot */
    a.g = L.g;
    u32 sb = (photons + 255) / 256;
    sb = sb > EMIT_START_BLOCKS ? EMIT_START_BLOCKS : sb;
    hipLaunchKernelGGL(k_emit_wf_start, dim3(sb), dim3(256), 0, s, a);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    /* enough blocks to fill the GPU; bounces with few live photons retire
       the surplus blocks at once */
    u32 eb = shape ? shape->emit : (u32)EMIT_WF_BLOCK;
    if (eb != 64 && eb != 256)
        eb = 128;
    u32 blocks = emit_queue_blocks(eb, S.slice);
    /* bounce-synchronous launches while most photons are alive, then one
       launch for the rest of every path */
    if (wave_bounces == 0)
        wave_bounces = split ? EMIT_SPLIT_BOUNCES : EMIT_WF_BOUNCES;
    wave_bounces = max_bounces < wave_bounces ? max_bounces : wave_bounces;
    /* persistent trace waves: a multiple of EMIT_NQ (every sub-queue is
       served), enough to fill every CU whatever the count */
    u32 tblocks = emit_trace_blocks(eb, photons);
    for (u32 depth = 0; depth < wave_bounces; depth++) {
        a.depth = depth;
        if (split) {
            if (eb == 64) {
                hipLaunchKernelGGL(k_emit_trace<64>, dim3(tblocks), dim3(64), 0, s, a);
                hipLaunchKernelGGL(k_emit_shade<64>, dim3(blocks), dim3(64), 0, s, a);
            } else if (eb == 256) {
                hipLaunchKernelGGL(k_emit_trace<256>, dim3(tblocks), dim3(256), 0, s, a);
                hipLaunchKernelGGL(k_emit_shade<256>, dim3(blocks), dim3(256), 0, s, a);
            } else {
                hipLaunchKernelGGL(k_emit_trace<128>, dim3(tblocks), dim3(128), 0, s, a);
                hipLaunchKernelGGL(k_emit_shade<128>, dim3(blocks), dim3(128), 0, s, a);
            }
            if ((e = hipGetLastError()) != hipSuccess)
                return e;
            continue;
        }
        if (eb == 64)
            hipLaunchKernelGGL(k_emit_wf_bounce<64>, dim3(blocks), dim3(64), 0, s, a);
        else if (eb == 256)
            hipLaunchKernelGGL(k_emit_wf_bounce<256>, dim3(blocks), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL(k_emit_wf_bounce<128>, dim3(blocks), dim3(128), 0, s, a);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    if (wave_bounces < max_bounces) {
        a.depth = wave_bounces;
        if (eb == 64)
            hipLaunchKernelGGL(k_emit_wf_finish<64>, dim3(blocks), dim3(64), 0, s, a);
        else if (eb == 256)
            hipLaunchKernelGGL(k_emit_wf_finish<256>, dim3(blocks), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL(k_emit_wf_finish<128>, dim3(blocks), dim3(128), 0, s, a);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    return launch_emit_compact(L, photons, max_photons, photon_pos, photon_power, photon_dir, n_photons, s,
                               deposit_rec);
}

bool kernel_info_trace(PhxKernelInfo *out, u32 max, u32 &n) {
    PhxKernelInfo k;
    bool ok = true;
    PHX_INFO("k_trace_camera", (k_trace_camera<128, true>), WF_CAM_BLOCK)
    PHX_INFO("k_trace_camera_rest", k_trace_camera_rest<128>, WF_CAM_BLOCK)
    PHX_INFO("k_trace_shadow", k_trace_shadow<128>, WF_SHADOW_BLOCK)
    PHX_INFO("k_trace_resolve", k_trace_resolve, 256)
    PHX_INFO("k_emit_wf_bounce", k_emit_wf_bounce<128>, EMIT_WF_BLOCK)
    PHX_INFO("k_emit_wf_finish", k_emit_wf_finish<128>, EMIT_WF_BLOCK)
    PHX_INFO("k_emit_trace", k_emit_trace<128>, EMIT_WF_BLOCK)
    PHX_INFO("k_emit_shade", k_emit_shade<128>, EMIT_WF_BLOCK)
    return ok;
}

} // namespace phx
