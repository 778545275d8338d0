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
 * emit_photons (kernels/emit_photons.cl:12-114): one lane per photon; the
 * k-th deposit of photon i goes to a depth-major staging slot [k][i] and the
 * per-photon deposit count is stored; a second kernel compacts the staging
 * area in (photon, bounce) order after an exclusive scan of the counts, which
 * makes the deposit order canonical instead of atomic_inc arrival order.
 */
#include "phx_device.h"

namespace phx {

constexpr int TRACE_BLOCK = 128;
constexpr int EMIT_BLOCK = 128;

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

/* ---------------------------------------------------------------------- */

struct EmitArgs {
    SceneView sc;
    u32 offset, photons_to_emit, max_bounces, seed;
    phx_float4 *st_pos; /* [max_bounces][photons_to_emit] */
    u32 *st_power, *st_dir, *count;
};

__global__ __launch_bounds__(EMIT_BLOCK) void k_emit_paths(EmitArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * EMIT_BLOCK];
    __shared__ f32 s_ior[IOR_STACK_SIZE * EMIT_BLOCK];
    const SceneView &sc = a.sc;
    u32 li = blockIdx.x * EMIT_BLOCK + threadIdx.x;
    if (li >= a.photons_to_emit)
        return;
    u32 *stk = s_stack + threadIdx.x;
    u32 tid = li + a.offset;
    u32 rng = pcg_seed(a.seed + tid);
    F4 origin, dir, power;
    sample_light(rng, sc, origin, dir, power);
    IorStack ior{s_ior + threadIdx.x, (u32)EMIT_BLOCK, 0u};
    init_ior_stack(ior, AIR_IOR);
    F4 throughput = mk(1.0f, 1.0f, 1.0f, 0.0f);
    u32 n = 0;
    size_t stride = a.photons_to_emit;
    for (u32 depth = 0; depth < a.max_bounces; depth++) {
        HitRecord rec;
        if (!scene_intersect(sc, origin, dir, EPS, INF, rec, stk, EMIT_BLOCK))
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
        if (bs.event == BSDF_DIFFUSE && dot(dir, sh.normal) < 0.0f) {
            size_t slot = (size_t)n * stride + li;
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
    }
    a.count[li] = n;
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

size_t emit_workspace_bytes(u32 photons, u32 max_bounces) {
    size_t slots = (size_t)photons * max_bounces;
    size_t b = 0;
    b += align_up(slots * sizeof(phx_float4), 256); /* staging pos */
    b += align_up(slots * sizeof(u32), 256);        /* staging power */
    b += align_up(slots * sizeof(u32), 256);        /* staging dir */
    b += align_up((size_t)photons * sizeof(u32), 256); /* counts */
    b += align_up((size_t)photons * sizeof(u32), 256); /* offsets */
    b += 256;                                          /* total */
    b += align_up(scan_workspace_bytes(photons), 256);
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
    u32 *count = (u32 *)p;
    p += align_up((size_t)photons * sizeof(u32), 256);
    u32 *offs = (u32 *)p;
    p += align_up((size_t)photons * sizeof(u32), 256);
    u32 *total = (u32 *)p;
    p += 256;
    void *scan_ws = p;

    EmitArgs a;
    a.sc = sc;
    a.offset = offset;
    a.photons_to_emit = photons;
    a.max_bounces = max_bounces;
    a.seed = seed;
    a.st_pos = st_pos;
    a.st_power = st_power;
    a.st_dir = st_dir;
    a.count = count;
    hipLaunchKernelGGL(k_emit_paths, dim3((photons + EMIT_BLOCK - 1) / EMIT_BLOCK), dim3(EMIT_BLOCK), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    e = launch_scan(count, photons, offs, nullptr, total, scan_ws, s);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_emit_compact, dim3((photons + 255) / 256), dim3(256), 0, s, st_pos, st_power, st_dir, count,
                       offs, photons, n_photons, max_photons, photon_pos, photon_power, photon_dir);
    e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_counter_add, dim3(1), dim3(1), 0, s, n_photons, total);
    return hipGetLastError();
}

} // namespace phx
