/*
 * pass_bodies.h — per-item bodies of the wavefront trace, the wavefront
 * photon emission and the flux gather, shared by the one-scene kernels
 * (kernels_trace.hip, kernels_photon_map.hip) and the multi-scene kernels
 * (kernels_batch.hip), so both run the same code on every pixel and photon.
 *
 * Wavefront trace (same outputs as trace_rays.cl:14-169, bit for bit): the
 * random draws of the NEE loops (light selection, emissive point) never
 * depend on shadow visibility, and the Russian roulette after a specular NEE
 * only reads throughput, so a camera path draws all light fragments of an NEE
 * event and keeps walking.  Shadow rays go to a queue (the direct_samples
 * rays of one event are contiguous: same origin, coherent directions) traced
 * by a lean kernel; the resolve step then forms emission and direct exactly in
 * the reference's order of additions:
 *   emission = ((BLACK + a_0) + s_0) + a_1 + ... (+ envmap on a miss),
 * a_k = throughput_k·emissive_k, s_k = throughput_k·(Σ_d pre_d·trans_d / n)
 * for specular events, and direct = (throughput·occlusion)·(Σ / n) at the
 * first diffuse event (trace_rays.cl:97-147).
 *
 * Wavefront emission (same deposits as emit_photons.cl:12-114): a photon's
 * state (origin, direction, power, throughput, PCG state, IOR stack, deposit
 * count) lives in HBM between bounces; its k-th deposit goes to staging slot
 * [k][photon].
 */
#ifndef PHX_PASS_BODIES_H
#define PHX_PASS_BODIES_H

#include "phx_cull.h"
#include "phx_device.h"

namespace phx {

constexpr int WF_CAM_BLOCK = 128;
constexpr int WF_SHADOW_BLOCK = 128;
constexpr int EMIT_WF_BLOCK = 128;
/* bounce-synchronous launches before the finishing launch: fused bounce
   kernels, and split ones (trace + shade; atrium 6.38 -> 6.19 ms/pass,
   caustic 7.47 -> 7.12 at 4 instead of 5) */
constexpr u32 EMIT_WF_BOUNCES = 5;
constexpr u32 EMIT_SPLIT_BOUNCES = 4;
constexpr u32 NEE_NONE = 0, NEE_DIFFUSE = 1, NEE_SPECULAR = 2;

constexpr u32 NO_EVENT = 0xFFFFFFFFu;
/* hit records of the persistent trace kernels: (t, u, v, position in the
   tree's triangle order), w = EMIT_MISS when the ray leaves the scene */
constexpr u32 EMIT_MISS = 0xFFFFFFFFu;

struct WfState {
    u32 n_paths;                 /* paths the layout holds: the stride of the per-bounce arrays */
    phx_float4 *rec_a, *rec_thr; /* [bounce · n_paths + path] (rec_at) */
    u32 *rec_base;               /* first shadow slot of the NEE event of that bounce */
    u32 *meta;                   /* per pixel: bits 0-2 records, 3-12 NEE kinds, 15 envmap term */
    phx_float4 *env;             /* per pixel: throughput · envmap on a miss */
    f32 *occl;                   /* per pixel: occlusion at the diffuse hit */
    u32 *counter;                /* shadow slots handed out */
    /* NEE events [bounce · n_paths + path]: hit position, shading normal, view
       direction, base colour (diffuse) or F0 (specular), (roughness,
       specular flag) */
    phx_float4 *ev_pos, *ev_n, *ev_view, *ev_col, *ev_misc;
    u32 *sh_rng;                 /* PCG state before the slot's light sample */
    u32 *sh_ev;                  /* event of the slot; NO_EVENT: no sample is drawn (adds BLACK) */
    phx_float4 *sh_tr;           /* NEE weight × shadow transmittance (BLACK without a shadow ray) */
    u32 *sh_scene;               /* scene of each slot (multi-scene queues only, else null) */
    /* camera paths that go on after their first bounce, handed from
       k_trace_camera to k_trace_camera_rest (null: k_trace_camera follows
       every path to its end): their count, pixels and state, four float4
       per path (camera_path_save) */
    u32 *cq_count, *cq_pix;
    phx_float4 *cq_state;
    u32 cap;
};

/* Index of bounce `depth` of path `wf` in the per-bounce arrays: bounce-major,
 * so the lanes of a wave, which hold neighbouring paths at the same bounce,
 * write and read neighbouring entries. */
PHX_DEV size_t rec_at(const WfState &S, u32 wf, u32 depth) { return (size_t)depth * S.n_paths + wf; }

/* hitpoint outputs written by the camera path */
struct CamOut {
    phx_float4 *hp_pos, *hp_normal, *hp_throughput, *hp_base_color;
};

/* The scene's only light is a point, spot or directional light.  Then every
 * light sample of an NEE event selects that light with select_light's one
 * draw and takes no other: its shadow ray, weight and transmittance are the
 * same for all n samples of the event (sample_light_fragment), so one queue
 * slot per event holds the value each sample has. */
PHX_DEV bool single_delta_light(const SceneView &sc) {
    return sc.n_lights == 1 && sc.lights[0].kind != PHX_LIGHT_TEXTURED;
}

/* A camera path of trace_rays.cl between bounces: the ray it follows, its
 * throughput and PCG state, the bounce it is at, the NEE kinds recorded so
 * far (2 bits per bounce), whether the envmap term was stored, whether it
 * ended at a diffuse hit and whether it goes on.  The IOR stack lives in
 * LDS beside it (IorStack). */
struct CamPath {
    F4 origin, dir, throughput;
    u32 rng, depth, kinds, env_flag;
    bool diffuse, active;
};

/* The path of pixel `pix` before its first bounce (trace_rays.cl:47-62):
 * make_ray from the pixel's PCG stream, white throughput, air on the IOR
 * stack. */
PHX_DEV CamPath camera_path_begin(const PhxCameraParams &cam, u32 w, u32 h, u32 seed, bool live, u32 pix,
                                  IorStack &ior) {
    CamPath p;
    u32 px = live ? pix % w : 0, py = live ? pix / w : 0;
    p.rng = pcg_seed(seed + pix);
    p.origin = ZERO;
    p.dir = ZERO;
    if (live)
        make_ray(cam, px, py, w, h, p.rng, p.origin, p.dir);
    p.throughput = WHITE;
    init_ior_stack(ior);
    p.depth = 0;
    p.kinds = 0;
    p.env_flag = 0;
    p.diffuse = false;
    p.active = live;
    return p;
}

/* The bounces of a camera path (trace_rays.cl:64-161) from p.depth until it
 * ends or reaches bounce `stop_depth`, with every NEE event's fragments
 * drawn and its shadow rays queued.  `out` is the path's index in the
 * hitpoint arrays, `wf` its index in the wavefront state.  Every lane of the
 * wave must call this together (the queue slots are allocated wave-wide);
 * lanes whose path is not active only take part in the allocation.  `slots`
 * queue slots per event: n, or 1 for a scene with a single delta light (the
 * slot gets the state before the first sample). */
PHX_DEV void camera_path_run(const SceneView &sc, u32 n, u32 out, u32 wf, u32 tag, const WfState &S, const CamOut &o,
                             u32 *stk, u32 block, u32 slots, CamPath &p, IorStack &ior, u32 stop_depth) {
    F4 &origin = p.origin, &dir = p.dir, &throughput = p.throughput;
    u32 &rng = p.rng, &depth = p.depth, &kinds = p.kinds, &env_flag = p.env_flag;
    bool &diffuse = p.diffuse;
    /* the loop runs while any lane of the wave has a bounce to take, so the
       wave-wide slot allocation below sees every lane */
    for (;;) {
        bool active = p.active && depth < stop_depth;
        if (!__any(active))
            break;
        bool nee = false, spec = false;
        F4 hit_pos = ZERO, sn = ZERO, bc = ZERO, f0 = ZERO, view = ZERO;
        f32 rough = 0.0f;
        Bsdf bs;
        bs.dir = ZERO;
        bs.throughput = ZERO;
        bs.h = ZERO;
        bs.event = BSDF_TRANSMIT;
        F4 hit_n = ZERO;
        if (active) {
            HitRecord rec;
            bool hit = scene_intersect(sc, origin, dir, EPS, INF, rec, stk, block);
            if (!hit) {
                if (sc.has_env) {
                    st(S.env + wf, throughput * sample_envmap(sc, dir));
                    env_flag = 1;
                }
                active = p.active = false;
            } else {
                SurfaceHit sh = process_hit(sc, rec, origin, dir);
                const PhxMaterial &mat = sc.materials[sh.mat_index];
                f32 curr_ior = current_ior(ior);
                if (curr_ior != AIR_IOR && mat.thickness > EPS && mat.att_dist > EPS)
                    throughput = throughput * beer_lambert(from(mat.att_color), mat.att_dist, sh.t);
                Ctx c = evaluate_material(sc, mat, sh.mat_index, rec, sh);
                st(S.rec_a + rec_at(S, wf, depth), throughput * c.emissive);
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
                    st(o.hp_pos + out, sh.position);
                    st(o.hp_normal + out, sh.normal);
                    st(o.hp_base_color + out, c.base_color);
                    st(o.hp_throughput + out, throughput);
                    S.occl[wf] = c.occlusion;
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
                    st(S.rec_thr + rec_at(S, wf, depth), throughput);
            }
        }
        /* one atomic per wave for the shadow slots of this bounce */
        unsigned long long want = __ballot(nee);
        u32 lane = threadIdx.x & 63u;
        u32 rank = (u32)__popcll(want & ((1ull << lane) - 1ull));
        u32 wave_base = 0;
        if (want) {
            u32 total = (u32)__popcll(want) * slots;
            u32 first = (u32)__ffsll((long long)want) - 1u;
            if (lane == first)
                wave_base = atomicAdd(S.counter, total);
            wave_base = __shfl(wave_base, (int)first, 64);
        }
        if (nee) {
            /* The light samples are drawn by the shadow kernel, one lane per
               sample, from the event recorded here and the PCG state before
               each sample; this lane only steps its stream past the draws
               the samples take (skip_light_fragments), so the path goes on
               with the state the reference has after direct lighting. */
            u32 base = wave_base + rank * slots;
            size_t e = rec_at(S, wf, depth);
            S.rec_base[e] = base;
            st(S.ev_pos + e, hit_pos);
            st(S.ev_n + e, sn);
            st(S.ev_view + e, view);
            st(S.ev_col + e, spec ? f0 : bc);
            st(S.ev_misc + e, mk(rough, spec ? 1.0f : 0.0f, 0.0f, 0.0f));
            /* direct_lighting_specular returns BLACK before sampling when
               n·v < EPS: no draws */
            bool draws = !spec || !(dot(view, sn) < EPS);
            for (u32 d = 0; d < slots; d++) {
                size_t slot = (size_t)base + d;
                S.sh_rng[slot] = rng;
                S.sh_ev[slot] = draws ? (u32)e : NO_EVENT;
                if (S.sh_scene)
                    S.sh_scene[slot] = tag;
                if (draws)
                    skip_light_fragments(rng, sc, 1u);
            }
            if (draws && n > slots)
                skip_light_fragments(rng, sc, n - slots);
        }
        if (diffuse)
            active = p.active = false;
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
                active = p.active = false;
        }
    }
}

/* The outputs of a camera path that has ended (trace_rays.cl:163-168): the
 * hitpoint of a path that found no diffuse surface, and the path's record
 * count, NEE kinds and envmap flag. */
PHX_DEV void camera_path_end(const WfState &S, const CamOut &o, bool live, u32 out, u32 wf, const CamPath &p) {
    if (!live)
        return;
    if (!p.diffuse) {
        st(o.hp_pos + out, ZERO);
        st(o.hp_normal + out, ZERO);
        st(o.hp_base_color + out, BLACK);
        st(o.hp_throughput + out, p.throughput);
    }
    /* records exist for every bounce that hit: 0..depth at a diffuse end,
       0..depth-1 after a miss (nothing recorded for the missing bounce), a
       roulette kill or the bounce limit (depth already advanced) */
    u32 nrec = p.diffuse ? p.depth + 1 : p.depth;
    S.meta[wf] = nrec | (p.kinds << 3) | (p.env_flag << 15);
}

/* A path that goes on after its first bounce, as entry q of the camera queue:
 * its ray, throughput, PCG state, NEE kinds and IOR stack.  After one bounce
 * the stack holds air and at most one more entry (one update per bounce). */
PHX_DEV void camera_path_save(const WfState &S, u32 q, const CamPath &p, const IorStack &ior) {
    phx_float4 *d = S.cq_state + (size_t)q * 4u;
    st(d, p.origin);
    st(d + 1, p.dir);
    st(d + 2, p.throughput);
    st(d + 3, mk(phx_as_float(p.rng), phx_as_float(p.kinds), phx_as_float(ior.cur),
                 ior.cur ? ior.ior[ior.stride] : 0.0f));
}
/* the path camera_path_save stored as entry q, at bounce 1 */
PHX_DEV CamPath camera_path_load(const WfState &S, u32 q, IorStack &ior) {
    const phx_float4 *d = S.cq_state + (size_t)q * 4u;
    CamPath p;
    p.origin = ld(d);
    p.dir = ld(d + 1);
    p.throughput = ld(d + 2);
    F4 m = ld(d + 3);
    p.rng = phx_as_uint(m.x);
    p.kinds = phx_as_uint(m.y);
    init_ior_stack(ior);
    ior.cur = phx_as_uint(m.z);
    if (ior.cur)
        ior.ior[ior.stride] = m.w;
    p.depth = 1;
    p.env_flag = 0;
    p.diffuse = false;
    p.active = true;
    return p;
}

/* The whole camera path of pixel `pix` (camera_path_begin, _run, _end). */
PHX_DEV void camera_path(const SceneView &sc, const PhxCameraParams &cam, u32 w, u32 h, u32 seed, u32 n, bool live,
                         u32 pix, u32 out, u32 wf, u32 tag, const WfState &S, const CamOut &o, u32 *stk, f32 *s_ior,
                         u32 block, u32 slots) {
    IorStack ior{s_ior, block, 0u};
    CamPath p = camera_path_begin(cam, w, h, seed, live, pix, ior);
    camera_path_run(sc, n, out, wf, tag, S, o, stk, block, slots, p, ior, MAX_RAY_BOUNCES);
    camera_path_end(S, o, live, out, wf, p);
}

/* The light sample of a shadow slot: draws it from the PCG state the camera
 * kernel recorded and weights it by the BSDF of the event (trace_rays.cl's
 * direct lighting, light_sampling.h:249-262 diffuse, 336-363 specular).
 * False when the sample contributes nothing (no light fragment, or a
 * specular weight below the reference's cut-off). */
PHX_DEV bool shadow_sample(const SceneView &sc, const WfState &S, u32 slot, u32 e, F4 &pre, F4 &o, F4 &d,
                           f32 &t_max) {
    u32 rng = S.sh_rng[slot];
    F4 hit_pos = ld(S.ev_pos + e), sn = ld(S.ev_n + e), view = ld(S.ev_view + e), col = ld(S.ev_col + e);
    F4 misc = ld(S.ev_misc + e);
    LightFragment f;
    bool valid = false;
    pre = BLACK;
    if (misc.y == 0.0f) {
        if (sample_light_fragment(rng, sc, hit_pos, sn, f)) {
            valid = true;
            pre = f.radiance * (col / PI);
        }
    } else {
        f32 ndv = dot(view, sn); /* >= EPS: otherwise the slot has no event */
        if (sample_light_fragment(rng, sc, hit_pos, sn, f) && !(f.n_dot_dir < EPS)) {
            f32 alpha = misc.x * misc.x;
            F4 hv = normalize(f.dir + view);
            f32 ndh = phx_fmaxf(dot(sn, hv), 0.0f);
            f32 ggx = ggx_ndf(ndh, alpha);
            F4 fr = fresnel4(col, view, hv);
            f32 sm = smith_g1_ggx(f.n_dot_dir, alpha) * smith_g1_ggx(ndv, alpha);
            F4 spv = f.radiance * ggx * fr * sm / (4.0f * f.n_dot_dir * ndv);
            if (!(dot(spv, spv) < EPS * EPS)) {
                valid = true;
                pre = spv;
            }
        }
    }
    if (!valid)
        return false;
    pre = mk(pre.x, pre.y, pre.z, 0.0f);
    o = mk(hit_pos.x, hit_pos.y, hit_pos.z, 0.0f);
    d = f.dir;
    t_max = f.dist - EPS;
    return true;
}

/* One step of shadow_transmittance's walk through transmissive surfaces
 * (light_sampling.h:302-333): folds the hit into the transmittance.  False
 * when the walk ends here (opaque surface, or transmittance below EPS: then
 * trans is BLACK). */
template <bool ANY>
PHX_DEV bool shadow_fold(const SceneView &sc, const HitRecord &rec, F4 o, F4 d, f32 &t_min, F4 &trans) {
    const PhxMaterial &mat = sc.materials[rec.mat_index];
    if (ANY || mat.transmission < EPS) {
        trans = BLACK;
        return false;
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
        return false;
    }
    t_min = rec.t + EPS;
    return true;
}

/* shadow_transmittance (light_sampling.h:302-333) of one queued shadow ray,
 * multiplied by the sample's NEE weight (the term resolve_pixel adds), BLACK
 * for a slot without a sample; ANY = any-hit traversal (exact when no
 * material transmits and the stack never overflows). */
template <bool ANY, u32 SB = WF_SHADOW_BLOCK>
PHX_DEV F4 shadow_value(const SceneView &sc, const WfState &S, u32 slot, u32 *stk) {
    u32 e = S.sh_ev[slot];
    F4 pre, o, d;
    f32 t_max;
    if (e == NO_EVENT || !shadow_sample(sc, S, slot, e, pre, o, d, t_max))
        return BLACK;
    f32 t_min = EPS;
    F4 trans = WHITE;
    for (u32 depth = 0; depth < MAX_RAY_BOUNCES; depth++) {
        HitRecord rec;
        if (!(ANY ? scene_any_hit(sc, o, d, t_min, t_max, rec, stk, SB)
                  : scene_intersect(sc, o, d, t_min, t_max, rec, stk, SB)))
            break;
        if (!shadow_fold<ANY>(sc, rec, o, d, t_min, trans))
            break;
    }
    return pre * trans;
}

template <bool ANY> PHX_DEV void shadow_one(const SceneView &sc, const WfState &S, u32 slot, u32 *stk) {
    st(S.sh_tr + slot, shadow_value<ANY>(sc, S, slot, stk));
}

/* The sum over an event's `n` samples in sample order, BLACK + s0 + s1 + ...
 * (light_sampling.h's loop), formed by the wave whose n adjacent lanes hold
 * them: n divides 64 and the event's first slot is a multiple of n, so the
 * group starts at lane (lane & ~(n - 1)).  Every lane of the wave that holds
 * a sample takes part; each gets its group's sum. */
PHX_DEV F4 group_sample_sum(F4 v, u32 n) {
    u32 lane = threadIdx.x & 63u;
    int g0 = (int)(lane & ~(n - 1u));
    F4 acc = BLACK;
    for (u32 d = 0; d < n; d++) {
        int src = g0 + (int)d;
        acc = acc + mk(__shfl(v.x, src, 64), __shfl(v.y, src, 64), __shfl(v.z, src, 64), __shfl(v.w, src, 64));
    }
    return acc;
}

/* emission and direct of one pixel from its records and shadow results;
 * `summed`: the first slot of each event holds the sum over its samples
 * (group_sample_sum), the others are not read; `shared`: each event has one
 * slot, the value of each of its n samples (single_delta_light) */
PHX_DEV void resolve_pixel(const WfState &S, u32 n, u32 wf, u32 out, phx_float4 *hp_direct, phx_float4 *hp_emission,
                           bool summed = false, bool shared = false) {
    u32 meta = S.meta[wf];
    u32 nrec = meta & 7u;
    F4 e = BLACK;
    F4 direct = BLACK;
    for (u32 k = 0; k < nrec; k++) {
        e = e + ld(S.rec_a + rec_at(S, wf, k));
        u32 kind = (meta >> (3 + 2 * k)) & 3u;
        if (kind == NEE_NONE)
            continue;
        size_t base = S.rec_base[rec_at(S, wf, k)];
        /* light_sampling.h: sum of weight × transmittance over the samples,
           BLACK for a sample without a shadow ray (products formed by
           shadow_one) */
        F4 sum = BLACK;
        if (shared) {
            F4 v = ld(S.sh_tr + base);
            for (u32 d = 0; d < n; d++)
                sum = sum + v;
        } else if (summed)
            sum = ld(S.sh_tr + base);
        else
            for (u32 d = 0; d < n; d++)
                sum = sum + ld(S.sh_tr + base + d);
        F4 thr = ld(S.rec_thr + rec_at(S, wf, k));
        if (kind == NEE_SPECULAR) {
            e = e + thr * (sum / (f32)n);
        } else {
            e = mk(e.x, e.y, e.z, 1.0f);
            direct = thr * S.occl[wf] * (sum / (f32)n);
        }
    }
    if ((meta >> 15) & 1u)
        e = e + ld(S.env + wf);
    st(hp_emission + out, e);
    st(hp_direct + out, direct);
}

/* ---------------- photon emission ---------------- */
/* The ray state of live photons: ray, power, throughput, PCG state and IOR
 * stack depth, one entry per slot. */
struct PhotonSet {
    phx_float4 *origin, *dir, *power, *thr;
    u32 *rng, *ior_cur;
};
struct EmitWfState {
    /* photon li's ray state at slot li of set 0 (origin ... ior_cur alias
       ps[0]).  With `dense` (split bounces) the state of the photon at
       entry i of sub-queue k entering bounce d is instead at slot
       k·slice + i of set d & 1: the shading kernel writes each survivor's
       state where it appends the survivor, so every bounce reads and writes
       its state in queue order (contiguous across a wave) */
    phx_float4 *origin, *dir, *power, *thr;
    u32 *rng, *ior_cur;
    PhotonSet ps[2];
    u32 dense;
    u32 *count; /* [photon]: deposits so far */
    f32 *ior;   /* [depth][photon] */
    /* live photons, as EMIT_NQ sub-queues of `slice` entries: photon slot
       li = k·slice + i starts in sub-queue k, and a photon's survivors stay
       in its sub-queue, whose capacity is therefore `slice` */
    u32 *queue[2];
    u32 *qcount; /* sub-queue k's length entering bounce d: emit_qcounter */
    /* split bounces (k_emit_trace, then k_emit_shade): the hit of entry i of
       sub-queue k at [k·slice + i] (t, u, v, position in the tree's triangle
       order, EMIT_MISS when the ray leaves the scene) and, per bounce and
       sub-queue, the ticket from which the trace kernel's waves take
       entries */
    phx_float4 *hit;
    u32 *ticket;
    u32 slice;
};

/* Sub-queues of the live-photon queue.  A queue append or a ticket is a
 * device-scope atomic, and atomics on one address serialise at ≈ 11 ns each
 * on MI355X (65 536 waves appending to one counter take ≈ 0.75 ms), so each
 * sub-queue has its own counter and ticket, one 128-byte line apart. */
constexpr u32 EMIT_NQ = 64;
constexpr u32 EMIT_QSTRIDE = 32; /* u32 between two counters */

PHX_DEV u32 *emit_qcounter(u32 *base, u32 depth, u32 k) {
    return base + ((size_t)depth * EMIT_NQ + k) * EMIT_QSTRIDE;
}
/* entries of sub-queue k entering bounce `depth` */
PHX_DEV u32 emit_qlen(const EmitWfState &S, u32 depth, u32 k, u32 photons) {
    if (depth == 0) {
        u32 lo = k * S.slice;
        return lo >= photons ? 0u : (photons - lo < S.slice ? photons - lo : S.slice);
    }
    return *emit_qcounter(S.qcount, depth, k);
}
/* u32 words of the per-bounce, per-sub-queue counters (and of the tickets) */
inline size_t emit_counter_words(u32 max_bounces) { return ((size_t)max_bounces + 2) * EMIT_NQ * EMIT_QSTRIDE; }
/* blocks of B threads for a per-entry emission kernel: a multiple of
 * EMIT_NQ, at most 512 Ki threads */
inline u32 emit_queue_blocks(u32 B, u32 slice) {
    u32 per_q = (slice + B - 1) / B;
    u32 cap = 4096u * 128u / B / EMIT_NQ;
    per_q = per_q < 1 ? 1 : (per_q > cap ? cap : per_q);
    return per_q * EMIT_NQ;
}
/* blocks of B threads for k_emit_trace: whole groups of EMIT_NQ waves, one
 * wave per 64 photons, at most 8192 waves */
inline u32 emit_trace_blocks(u32 B, u32 photons) {
    u32 group = EMIT_NQ * 64u / B; /* blocks holding EMIT_NQ waves */
    u32 waves = (photons + 63u) / 64u;
    u32 groups = (waves + EMIT_NQ - 1) / EMIT_NQ;
    groups = groups < 1 ? 1 : (groups > 8192u / EMIT_NQ ? 8192u / EMIT_NQ : groups);
    return groups * group;
}
/* photon slot of entry i of sub-queue k at bounce `depth` */
PHX_DEV u32 emit_qentry(const EmitWfState &S, u32 depth, u32 k, u32 i) {
    return depth == 0 ? k * S.slice + i : S.queue[depth & 1][(size_t)k * S.slice + i];
}

/* deposit staging, [k][photon] with `stride` photons per row */
struct EmitStaging {
    phx_float4 *pos;
    u32 *power, *dir;
    u32 stride;
};

/* emit_photons.cl:37-47 for photon `index` of its scene, state slot li */
PHX_DEV void emit_start_one(const SceneView &sc, u32 seed, u32 index, u32 li, const EmitWfState &S, u32 stride) {
    u32 rng = pcg_seed(seed + index);
    F4 origin, dir, power;
    sample_light(rng, sc, origin, dir, power);
    st(S.origin + li, origin);
    st(S.dir + li, dir);
    st(S.power + li, power);
    st(S.thr + li, mk(1.0f, 1.0f, 1.0f, 0.0f));
    S.rng[li] = rng;
    S.ior_cur[li] = 0;
    S.ior[li] = AIR_IOR;
    S.count[li] = 0;
    (void)stride;
}

/* The part of one bounce of emit_photons.cl:49-113 after the ray has hit
 * (rec); returns false when the path ends (roulette or the bounce limit). */
PHX_DEV bool emit_shade(const SceneView &sc, const EmitStaging &g, u32 li, u32 depth, u32 max_bounces,
                        const HitRecord &rec, F4 &origin, F4 &dir, F4 &power, F4 &throughput, u32 &rng, IorStack &ior,
                        u32 &n) {
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
        size_t slot = (size_t)n * g.stride + li;
        st(g.pos + slot, sh.position);
        g.power[slot] = encode_rgbe(power);
        g.dir[slot] = encode_oct(dir);
        n++;
    }
    power = power * bs.throughput;
    throughput = throughput * bs.throughput;
    F4 side = (dot(bs.dir, sh.normal) > 0.0f) ? sh.normal : -sh.normal;
    origin = sh.position + bs.dir * EPS + side * EPS;
    dir = bs.dir;
    return depth + 1 < max_bounces;
}

/* One bounce of emit_photons.cl:49-113 on a photon's state; returns false
 * when the path ends (escape, roulette or the bounce limit).  SB: the block
 * size, the stride of the LDS stack. */
template <u32 SB = EMIT_WF_BLOCK>
PHX_DEV bool emit_step(const SceneView &sc, const EmitStaging &g, u32 li, u32 depth, u32 max_bounces, F4 &origin,
                       F4 &dir, F4 &power, F4 &throughput, u32 &rng, IorStack &ior, u32 &n, u32 *stk) {
    HitRecord rec;
    if (!scene_intersect<true>(sc, origin, dir, EPS, INF, rec, stk, SB))
        return false;
    return emit_shade(sc, g, li, depth, max_bounces, rec, origin, dir, power, throughput, rng, ior, n);
}

/* Loads photon li's state (its ray state from slot `at` of set ps, li's own
 * slot of set 0 unless the state is dense), runs bounces depth .. (depth + 1
 * if one_bounce, else to the end of its path), stores what changed; returns
 * whether the photon is still live after a single bounce. */
template <u32 SB = EMIT_WF_BLOCK>
PHX_DEV bool emit_advance(const SceneView &sc, const EmitWfState &S, const EmitStaging &g, u32 li, u32 depth,
                          u32 max_bounces, bool one_bounce, u32 *stk, const PhotonSet *ps = nullptr,
                          size_t at = 0) {
    if (!ps) {
        ps = &S.ps[0];
        at = li;
    }
    F4 origin = ld(ps->origin + at), dir = ld(ps->dir + at);
    if (one_bounce) {
        /* the rest of the state is loaded after the traversal, which then
           runs with fewer live registers */
        HitRecord rec;
        if (!scene_intersect<true>(sc, origin, dir, EPS, INF, rec, stk, SB))
            return false;
        F4 power = ld(ps->power + at), throughput = ld(ps->thr + at);
        u32 rng = ps->rng[at];
        IorStack ior{S.ior + li, g.stride, ps->ior_cur[at]};
        u32 n = S.count[li], n0 = n;
        bool keep = emit_shade(sc, g, li, depth, max_bounces, rec, origin, dir, power, throughput, rng, ior, n);
        if (keep) {
            st(ps->origin + at, origin);
            st(ps->dir + at, dir);
            st(ps->power + at, power);
            st(ps->thr + at, throughput);
            ps->rng[at] = rng;
            ps->ior_cur[at] = ior.cur;
        }
        if (n != n0)
            S.count[li] = n;
        return keep;
    }
    F4 power = ld(ps->power + at), throughput = ld(ps->thr + at);
    u32 rng = ps->rng[at];
    IorStack ior{S.ior + li, g.stride, ps->ior_cur[at]};
    u32 n = S.count[li], n0 = n;
    for (u32 d = depth; d < max_bounces; d++)
        if (!emit_step<SB>(sc, g, li, d, max_bounces, origin, dir, power, throughput, rng, ior, n, stk))
            break;
    if (n != n0)
        S.count[li] = n;
    return false;
}

/* A surviving photon's ray state between the shading of one bounce and its
 * store at the slot its queue append gives it. */
struct PhotonRay {
    F4 origin, dir, power, thr;
    u32 rng, ior_cur;
};

/* The part of bounce `depth` of photon li after its ray has been traced on
 * the pair tree (k_emit_trace): h = (t, u, v, tree position or EMIT_MISS),
 * the ray state read from slot `at` of set `in`.  Returns whether the photon
 * is still live, its new ray state in r; same state and deposits as
 * emit_advance's single bounce. */
PHX_DEV bool emit_after_hit(const SceneView &sc, const EmitWfState &S, const EmitStaging &g, u32 li, u32 depth,
                            u32 max_bounces, F4 h, const PhotonSet &in, size_t at, PhotonRay &r) {
    u32 j = phx_as_uint(h.w);
    if (j == EMIT_MISS)
        return false;
    HitRecord rec;
    rec.t = h.x;
    rec.u = h.y;
    rec.v = h.z;
    rec.tri_index = sc.tri_order[j];
    rec.mat_index = sc.tmat[rec.tri_index];
    r.origin = ld(in.origin + at);
    r.dir = ld(in.dir + at);
    r.power = ld(in.power + at);
    r.thr = ld(in.thr + at);
    r.rng = in.rng[at];
    IorStack ior{S.ior + li, g.stride, in.ior_cur[at]};
    u32 n = S.count[li], n0 = n;
    bool keep = emit_shade(sc, g, li, depth, max_bounces, rec, r.origin, r.dir, r.power, r.thr, r.rng, ior, n);
    r.ior_cur = ior.cur;
    if (n != n0)
        S.count[li] = n;
    return keep;
}

/* appends li to a queue when keep, one atomic per wave, and returns the
 * position it took (undefined without keep); every lane of the wave must
 * call it */
PHX_DEV u32 wave_append(bool keep, u32 li, u32 *qcount, u32 *queue) {
    u32 lane = threadIdx.x & 63u;
    unsigned long long m = __ballot(keep);
    u32 pos = 0;
    if (m) {
        u32 first = (u32)__ffsll((long long)m) - 1u;
        u32 base = 0;
        if (lane == first)
            base = atomicAdd(qcount, (u32)__popcll(m));
        base = __shfl(base, (int)first, 64);
        pos = base + (u32)__popcll(m & ((1ull << lane) - 1ull));
        if (keep)
            queue[pos] = li;
    }
    return pos;
}

/* ---------------- gather (gather_flux.cl + photon_hash.h:56-103) ---------------- */
struct GatherArgs {
    const phx_float4 *hp_pos, *hp_normal, *hp_throughput, *hp_base_color, *hp_direct, *hp_emission;
    u32 n_pixels;
    const phx_float4 *ppos;
    const u32 *ppow, *pdir, *cell_start, *cell_end;
    PhxPhotonHashData d;
    f32 alpha;
    phx_float4 *flux;
    f32 *photon_weight;
    phx_float4 *radiance;
    /* optional (adjoint forward): the pass's radius ratio (1 where the state
       is not updated) and r² per pixel */
    f32 *rec_ratio, *rec_r2;
    /* optional: every stored record's direction already decoded (decode_oct
       of pdir, the same values) with its packed power (ppow) in w */
    const phx_float4 *pdirpow;
    /* optional (records sorted by cell, then half-cell octant): the first
       record of each octant of every non-empty cell, 8 per cell
       (k_octant_starts); gather_pixel_flat then visits only the octants
       within reach */
    const u32 *oct_start;
};

/* photons of a cell whose records are loaded together before any is used:
 * one memory round trip per batch of photons instead of one or two each */
constexpr u32 GATHER_BATCH = 8;
constexpr u32 GATHER_BATCH_DECODED = 8;

/* Marks a loaded value as used here, so that its load stays with the rest of
 * its batch instead of being moved into the branch that reads it (which would
 * cost a memory round trip of its own on every pass through that branch). */
PHX_DEV void keep_loaded(const F4 &v) { asm volatile("" ::"v"(v.x), "v"(v.y), "v"(v.z), "v"(v.w)); }

/* Gaussian weight of a photon at squared distance dsq < r_sq (photon_hash.h
 * 86-90), GAUSS_ALPHA·(1 − (1 − exp(−GAUSS_BETA·dsq / (2r²))) / GAUSS_DENOM)
 * with the same roundings, inv_two_r_sq = RN(1 / (2r²)):
 *  - the exponent lies in (-1, 0], where phx_expf_normal gives phx_expf's
 *    bits without its range branches;
 *  - both divisions are phx_div_recip, correctly rounded wherever the
 *    numerator is at least 2^-100 in magnitude; 1 − e is 0 or at least
 *    2^-24, and a smaller GAUSS_BETA·dsq (a photon within 1e-15 of the
 *    hitpoint) gives an exponent below 2^-80 (r² is at least 7.5e-7, the
 *    half-diagonal of a 0.001 cell, gather_begin), where exp = 1 whatever
 *    its last bits;
 *  - a NaN distance takes the literal expression (NaN payloads). */
constexpr f32 GAUSS_INV_DENOM = 1.0f / GAUSS_DENOM;

PHX_DEV f32 gauss_weight(f32 dsq, f32 two_r_sq, f32 inv_two_r_sq) {
    if (dsq != dsq)
        return GAUSS_ALPHA * (1.0f - (1.0f - phx_expf(-GAUSS_BETA * dsq / two_r_sq)) / GAUSS_DENOM);
    f32 x = phx_div_recip(-GAUSS_BETA * dsq, two_r_sq, inv_two_r_sq);
    f32 e = phx_expf_normal(x);
    return GAUSS_ALPHA * (1.0f - phx_div_recip(1.0f - e, GAUSS_DENOM, GAUSS_INV_DENOM));
}

/* The photons of one cell, records [s0, e0), that lie within the radius of
 * a hitpoint (photon_hash.h:74-98): Gaussian weights and weighted powers added
 * to nf / found in storage order.  DECODED reads the records' decoded
 * direction (with the packed power) instead of decoding it per visit: the
 * direction test dot(decode_oct(d), sn) is ((x·sn.x + y·sn.y) + z·sn.z) +
 * 0·sn.w, the last term (decode_oct's w is +0) being the lane's constant wz.
 * wz is +0, −0 or NaN, so the test `s + wz > 0` is `s > wz`: for a zero wz
 * the sum is s (or +0 when s is −0, also not above zero), for NaN both are
 * false. */
template <bool DECODED>
PHX_DEV void gather_cell(const GatherArgs &a, u32 s0, u32 e0, F4 pos, F4 sn, f32 r_sq, f32 two_r_sq,
                         f32 inv_two_r_sq, F4 &nf, f32 &found) {
    if (DECODED) {
        const f32 wz = 0.0f * sn.w;
        for (u32 k = s0; k < e0; k += GATHER_BATCH_DECODED) {
            F4 pp[GATHER_BATCH_DECODED], pd[GATHER_BATCH_DECODED];
#pragma unroll
            for (u32 j = 0; j < GATHER_BATCH_DECODED; j++) {
                u32 kk = k + j < e0 ? k + j : e0 - 1; /* in range: k < e0 */
                pp[j] = ld(a.ppos + kk);
                pd[j] = ld(a.pdirpow + kk);
            }
#pragma unroll
            for (u32 j = 0; j < GATHER_BATCH_DECODED; j++)
                keep_loaded(pd[j]);
#pragma unroll
            for (u32 j = 0; j < GATHER_BATCH_DECODED; j++) {
                if (k + j >= e0)
                    break;
                F4 dd = pos - pp[j];
                f32 dsq = dot(dd, dd);
                if (dsq >= r_sq)
                    continue;
                if (((pd[j].x * sn.x + pd[j].y * sn.y) + pd[j].z * sn.z) > wz)
                    continue;
                f32 wgt = gauss_weight(dsq, two_r_sq, inv_two_r_sq);
                nf = nf + wgt * decode_rgbe(phx_as_uint(pd[j].w));
                found += wgt;
            }
        }
        return;
    }
    for (u32 k = s0; k < e0; k += GATHER_BATCH) {
        F4 pp[GATHER_BATCH];
        u32 pw[GATHER_BATCH], pdr[GATHER_BATCH];
#pragma unroll
        for (u32 j = 0; j < GATHER_BATCH; j++) {
            u32 kk = k + j < e0 ? k + j : e0 - 1; /* in range: k < e0 */
            pp[j] = ld(a.ppos + kk);
            pw[j] = a.ppow[kk];
            pdr[j] = a.pdir[kk];
        }
#pragma unroll
        for (u32 j = 0; j < GATHER_BATCH; j++) {
            if (k + j >= e0)
                break;
            F4 dd = pos - pp[j];
            f32 dsq = dot(dd, dd);
            if (dsq >= r_sq)
                continue;
            F4 pd = decode_oct(pdr[j]);
            if (dot(pd, sn) > 0.0f)
                continue;
            f32 wgt = gauss_weight(dsq, two_r_sq, inv_two_r_sq);
            nf = nf + wgt * decode_rgbe(pw[j]);
            found += wgt;
        }
    }
}

/* Cell i of the 27-cell query (photon_hash.h:65-100): its bucket, or 0 when
 * get_hash gives 0 or the cell lies provably beyond the radius, so that no
 * photon stored in it can pass the distance test and the sums are unchanged
 * (phx_cull.h: the gap to the cell's box less the margins m).  The cells are
 * 1.155 radii wide when their half-diagonal is the radius, so about 16 of the
 * 27 survive on average.  l receives the cell's low corner. */
PHX_DEV u32 query_cell(const PhxPhotonHashData &d, F4 pos, f32 r_sq, u32 i, F4 m, F4 &l) {
    i32 cx = (i32)(i % 3) - 1, cy = (i32)((i / 3) % 3) - 1, cz = (i32)((i / 9) % 3) - 1;
    f32 qx = pos.x + (f32)cx * d.cell_sizes.x;
    f32 qy = pos.y + (f32)cy * d.cell_sizes.y;
    f32 qz = pos.z + (f32)cz * d.cell_sizes.z;
    u32 h = get_hash(qx, qy, qz, d);
    if (h == 0)
        return 0;
    /* get_hash's cell indices, recovered from the same quotients */
    u32 ix = (u32)((qx - d.origin.x) / d.cell_sizes.x);
    u32 iy = (u32)((qy - d.origin.y) / d.cell_sizes.y);
    u32 iz = (u32)((qz - d.origin.z) / d.cell_sizes.z);
    l = mk(d.origin.x + (f32)ix * d.cell_sizes.x, d.origin.y + (f32)iy * d.cell_sizes.y,
           d.origin.z + (f32)iz * d.cell_sizes.z, 0.0f);
    f32 gx = phx_cull_gap(pos.x, l.x, l.x + d.cell_sizes.x, m.x);
    f32 gy = phx_cull_gap(pos.y, l.y, l.y + d.cell_sizes.y, m.y);
    f32 gz = phx_cull_gap(pos.z, l.z, l.z + d.cell_sizes.z, m.z);
    return phx_cull_far(gx, gy, gz, r_sq) ? 0u : h;
}

/* Octants of a cell (low corner l) that may hold a photon within the radius:
 * bit x + 2y + 4z for the half x, y, z (0 low, 1 high).  A record is in the
 * high half of an axis when its quotient's fraction is at least 0.5
 * (get_hash_octant), so it lies in [l + cs/2, l + cs] up to the same
 * roundings as the cell's own edges plus one for the half-cell plane, which
 * the per-axis margin m covers as well (phx_cull.h); an octant is skipped only
 * when its box is provably farther than r, so the sums are unchanged. */
PHX_DEV u32 octant_mask(const PhxPhotonHashData &d, F4 pos, f32 r_sq, F4 l, F4 m) {
    f32 mx = l.x + d.cell_sizes.x * 0.5f, my = l.y + d.cell_sizes.y * 0.5f, mz = l.z + d.cell_sizes.z * 0.5f;
    f32 gx0 = phx_cull_gap(pos.x, l.x, mx, m.x), gx1 = phx_cull_gap(pos.x, mx, l.x + d.cell_sizes.x, m.x);
    f32 gy0 = phx_cull_gap(pos.y, l.y, my, m.y), gy1 = phx_cull_gap(pos.y, my, l.y + d.cell_sizes.y, m.y);
    f32 gz0 = phx_cull_gap(pos.z, l.z, mz, m.z), gz1 = phx_cull_gap(pos.z, mz, l.z + d.cell_sizes.z, m.z);
    u32 mask = 0;
#pragma unroll
    for (u32 q = 0; q < 8; q++)
        if (!phx_cull_far((q & 1u) ? gx1 : gx0, (q & 2u) ? gy1 : gy0, (q & 4u) ? gz1 : gz0, r_sq))
            mask |= 1u << q;
    return mask;
}

/* per-axis culling margins for a hitpoint (phx_cull_margin) */
PHX_DEV F4 cull_margins(const PhxPhotonHashData &d, F4 pos) {
    return mk(phx_cull_margin(d.origin.x, pos.x, d.cell_sizes.x, d.grid_res),
              phx_cull_margin(d.origin.y, pos.y, d.cell_sizes.y, d.grid_res),
              phx_cull_margin(d.origin.z, pos.z, d.cell_sizes.z, d.grid_res), 0.0f);
}

template <bool DECODED>
PHX_DEV void gather_cells_lane(const GatherArgs &a, const u32 *cs, const u32 *ce, const PhxPhotonHashData &d, F4 pos,
                               F4 sn, f32 r_sq, f32 two_r_sq, f32 inv_two_r_sq, F4 &nf, f32 &found) {
    const F4 m = cull_margins(d, pos);
    for (u32 i = 0; i < 27; i++) {
        F4 l;
        u32 h = query_cell(d, pos, r_sq, i, m, l);
        if (h == 0)
            continue;
        gather_cell<DECODED>(a, cs[h], ce[h], pos, sn, r_sq, two_r_sq, inv_two_r_sq, nf, found);
    }
}

/* radiance update of gather_flux.cl:24-45 and the pixel's SPPM inputs;
 * returns whether the pixel has a diffuse hitpoint to gather for */
PHX_DEV bool gather_begin(const GatherArgs &a, u32 tid, const PhxPhotonHashData &d, F4 &pos, F4 &sn, F4 &fsf,
                          f32 &wsf, f32 &r_sq) {
    F4 ge = ld(a.hp_emission + tid);
    F4 rsf = ld(a.radiance + tid);
    F4 r = mk(ge.x, ge.y, ge.z, 0.0f);
    if (ge.w == 0.0f) {
        st(a.radiance + tid, rsf + r);
        if (a.rec_ratio)
            a.rec_ratio[tid] = 1.0f;
        return false;
    }
    F4 dr = ld(a.hp_direct + tid);
    r.x += dr.x;
    r.y += dr.y;
    r.z += dr.z;
    st(a.radiance + tid, rsf + r);
    fsf = ld(a.flux + tid);
    wsf = a.photon_weight[tid];
    r_sq = fsf.w;
    if (r_sq < EPS) {
        f32 hx = d.cell_sizes.x * 0.5f, hy = d.cell_sizes.y * 0.5f, hz = d.cell_sizes.z * 0.5f;
        r_sq = hx * hx + hy * hy + hz * hz;
    }
    pos = ld(a.hp_pos + tid);
    sn = ld(a.hp_normal + tid);
    return true;
}

/* SPPM update of gather_flux.cl:65-77 */
PHX_DEV void gather_end(const GatherArgs &a, u32 tid, f32 alpha, F4 fsf, f32 wsf, f32 r_sq, F4 nf, f32 found) {
    f32 ratio = 1.0f;
    if (found > 0.0f) {
        f32 npw = wsf + alpha * found;
        ratio = npw / (wsf + found);
        f32 nr = r_sq * ratio;
        F4 albedo = ld(a.hp_base_color + tid);
        F4 contrib = ld(a.hp_throughput + tid) * (albedo / PI) * nf;
        F4 tmp = (fsf + contrib) * ratio;
        st(a.flux + tid, mk(tmp.x, tmp.y, tmp.z, nr));
        a.photon_weight[tid] = npw;
    }
    if (a.rec_ratio) {
        a.rec_ratio[tid] = ratio;
        a.rec_r2[tid] = r_sq;
    }
}

/* gather_flux.cl:24-77 for pixel tid, one lane on its own; the grid cell h
 * of the pixel's scene is entry bucket_base + h of cell_start / cell_end. */
template <bool DECODED>
PHX_DEV void gather_pixel(const GatherArgs &a, u32 tid, const PhxPhotonHashData &d, f32 alpha, u32 bucket_base) {
    F4 pos, sn, fsf;
    f32 wsf, r_sq;
    if (!gather_begin(a, tid, d, pos, sn, fsf, wsf, r_sq))
        return;
    F4 nf = BLACK;
    f32 found = 0.0f;
    f32 two_r_sq = 2.0f * r_sq;
    gather_cells_lane<DECODED>(a, a.cell_start + bucket_base, a.cell_end + bucket_base, d, pos, sn, r_sq, two_r_sq,
                               1.0f / two_r_sq, nf, found);
    gather_end(a, tid, alpha, fsf, wsf, r_sq, nf, found);
}

/* The 27-cell query of one hitpoint worked out once (photon_hash.h:65-100).
 * The query point of cell (cx, cy, cz) ∈ {−1, 0, 1}³ is pos + (cx·cs.x,
 * cy·cs.y, cz·cs.z), and get_hash's quotient along an axis depends only on
 * that axis' offset, so 9 quotients (instead of 27 × 3) give every cell's
 * indices — idx[a] holds the 3 indices of axis a in 9-bit fields, 511 for a
 * query point outside the grid — and, per axis and offset, the squared gaps of
 * the cell's low and high half (octant_mask's boxes and margins) give the
 * reach of all 27 × 8 octants: bit 8·(i mod 4) + s of word i / 4 is set when
 * octant s of cell i = (cx+1) + 3(cy+1) + 9(cz+1) may hold a photon within
 * r (all 8 bits of a reached cell when there are no octant starts).  Scanning
 * the bits in order visits cells in query order and octants in storage
 * order: the order of the sums. */
struct QueryReach {
    u32 idx[3];
    u32 b0, b1, b2, b3, b4, b5, b6;

    __device__ __forceinline__ u32 word(u32 wi) const {
        u32 w = b6;
        w = wi == 5 ? b5 : w;
        w = wi == 4 ? b4 : w;
        w = wi == 3 ? b3 : w;
        w = wi == 2 ? b2 : w;
        w = wi == 1 ? b1 : w;
        return wi == 0 ? b0 : w;
    }
};

/* one axis of QueryReach: the indices of the three query points' cells and
 * the squared half-cell gaps (+inf outside the grid) */
PHX_DEV void axis_reach(f32 p, f32 o, f32 cs, u32 g, f32 m, u32 &packed, f32 (&sq)[3][2]) {
    packed = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        f32 q = p + (f32)(c - 1) * cs;
        f32 f = (q - o) / cs;
        bool out = f < 0.0f || f >= (f32)g;
        u32 i = out ? 0u : (u32)f;
        f32 l = o + (f32)i * cs;
        f32 mid = l + cs * 0.5f;
        f32 g0 = phx_cull_gap(p, l, mid, m), g1 = phx_cull_gap(p, mid, l + cs, m);
        sq[c][0] = out ? __builtin_inff() : g0 * g0;
        sq[c][1] = out ? __builtin_inff() : g1 * g1;
        packed |= (out ? 511u : i) << (9 * c);
    }
}

PHX_DEV QueryReach query_reach(const PhxPhotonHashData &d, F4 pos, f32 r_sq, bool octants) {
    QueryReach R;
    const F4 m = cull_margins(d, pos);
    f32 sx[3][2], sy[3][2], sz[3][2];
    axis_reach(pos.x, d.origin.x, d.cell_sizes.x, d.grid_res, m.x, R.idx[0], sx);
    axis_reach(pos.y, d.origin.y, d.cell_sizes.y, d.grid_res, m.y, R.idx[1], sy);
    axis_reach(pos.z, d.origin.z, d.cell_sizes.z, d.grid_res, m.z, R.idx[2], sz);
    const f32 lim = r_sq * 1.00001f; /* phx_cull_far */
    u32 w[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (u32 i = 0; i < 27; i++) {
        const u32 cx = i % 3, cy = (i / 3) % 3, cz = i / 9;
        u32 byte = 0;
#pragma unroll
        for (u32 q = 0; q < 8; q++) {
            f32 sum = (sx[cx][q & 1u] + sy[cy][(q >> 1) & 1u]) + sz[cz][q >> 2];
            byte |= (sum >= lim ? 0u : 1u) << q; /* NaN keeps the octant */
        }
        if (!octants)
            byte = byte ? 0xFFu : 0u;
        w[i / 4] |= byte << (8 * (i % 4));
    }
    R.b0 = w[0], R.b1 = w[1], R.b2 = w[2], R.b3 = w[3], R.b4 = w[4], R.b5 = w[5], R.b6 = w[6];
    return R;
}

/* gather_pixel<true> with the cell loop and the record loop flattened into
 * one: each iteration first moves every lane whose record range is used up to
 * its next range, then processes one batch of records of every lane's
 * current range.  A lane that leaves a short range goes on at once instead of
 * idling until the wave's longest one is done (3.45 -> 3.1 ms on the caustic
 * workload).  With octant starts (a.oct_start) the ranges are the runs of
 * consecutive octants of a cell that lie within reach (octant_mask), visited
 * in octant order; without, whole cells.  Either way a pixel's accepted
 * photons are summed in gather_pixel's order (cells in query order, records
 * in storage order; skipped octants hold no photon the distance test
 * accepts): same bits.  Every lane of the wave calls it (valid = false for
 * lanes without a pixel).  Handing pixels out dynamically (a lane that
 * finishes takes another pixel of its block) was measured slower: the lanes
 * of a wave then query unrelated cells and stop sharing record loads. */
template <bool DECODED = true>
PHX_DEV void gather_pixel_flat(const GatherArgs &a, bool valid, u32 tid, const PhxPhotonHashData &d, f32 alpha,
                               u32 bucket_base) {
    const u32 *cs = a.cell_start + bucket_base, *ce = a.cell_end + bucket_base;
    const u32 *os = a.oct_start ? a.oct_start + (size_t)bucket_base * 8u : nullptr;
    F4 pos = ZERO, sn = ZERO, fsf = ZERO, nf = BLACK;
    f32 wsf = 0.0f, r_sq = 0.0f, found = 0.0f;
    /* R: the cells and octants within reach; wi, w: the word of R being
       scanned and its bits not yet visited (more: bits left); cell, h,
       [c0, c1): the cell being visited, its bucket and records; [k, e):
       records of the current range, [k2, e2): the range after it (fetched
       when fewer than a batch of the current one are left, so that one batch
       can take records from both) */
    u32 wi = 0, w = 0, cell = 27, h = 0, c0 = 0, c1 = 0, k = 0, e = 0, k2 = 0, e2 = 0;
    bool more = false;
    if (valid && gather_begin(a, tid, d, pos, sn, fsf, wsf, r_sq))
        more = true;
    const bool px = more;
    const f32 two_r_sq = 2.0f * r_sq, inv_two_r_sq = px ? 1.0f / two_r_sq : 0.0f, wz = 0.0f * sn.w;
    const QueryReach R = query_reach(d, pos, r_sq, os != nullptr);
    const u32 g = d.grid_res;
    w = px ? R.b0 : 0u;
    constexpr u32 B = GATHER_BATCH_DECODED;
    for (;;) {
        for (;;) {
            if (k >= e && k2 < e2) {
                k = k2;
                e = e2;
                k2 = e2 = 0;
            }
            bool need = k >= e || (e - k < B && k2 >= e2);
            if (!need || !more)
                break;
            while (w == 0 && wi < 6) {
                wi++;
                w = R.word(wi);
            }
            if (w == 0) {
                more = false;
                break;
            }
            /* the run of consecutive octants [q0, q1) of cell i from the lowest bit left */
            u32 bit = (u32)__builtin_ctz(w);
            u32 cb = bit >> 3;
            u32 i = wi * 4u + cb;
            u32 q0 = bit & 7u;
            u32 byte = (w >> (cb * 8u)) & 0xFFu;
            u32 q1 = q0 + (u32)__builtin_ctz(~(byte >> q0));
            w &= ~(((1u << (q1 - q0)) - 1u) << bit);
            if (i != cell) {
                cell = i;
                u32 ix = (R.idx[0] >> (9u * (i % 3u))) & 511u;
                u32 iy = (R.idx[1] >> (9u * ((i / 3u) % 3u))) & 511u;
                u32 iz = (R.idx[2] >> (9u * (i / 9u))) & 511u;
                h = g * g * ix + g * iy + iz + 1u; /* get_hash of the query point */
                c0 = cs[h];
                c1 = ce[h];
                if (c0 >= c1) { /* empty: its octant starts are not written */
                    w &= ~(0xFFu << (cb * 8u));
                    continue;
                }
            }
            u32 rk = q0 == 0 ? c0 : os[(size_t)h * 8u + q0];
            u32 re = q1 == 8 ? c1 : os[(size_t)h * 8u + q1];
            if (rk < re) {
                if (k >= e) {
                    k = rk;
                    e = re;
                } else {
                    k2 = rk;
                    e2 = re;
                }
            }
        }
        if (!__any(k < e))
            break;
        /* one batch: the next records of the current range, then of the next */
        u32 n1 = k < e ? (e - k < B ? e - k : B) : 0u;
        u32 n2 = (n1 < B && k2 < e2) ? (e2 - k2 < B - n1 ? e2 - k2 : B - n1) : 0u;
        u32 nt = n1 + n2;
        bool work = nt > 0;
        if (work) {
            /* the batch's records, all loaded before any is used (a lane
               without records loads nothing: its registers stay unset) */
            F4 pp[B], pd[B];
            u32 kl = n2 ? k2 + n2 - 1 : k + n1 - 1;
            u32 pdr[DECODED ? 1 : B];
#pragma unroll
            for (u32 q = 0; q < B; q++) {
                u32 kk = q < n1 ? k + q : (q < nt ? k2 + (q - n1) : kl);
                pp[q] = ld(a.ppos + kk);
                if (DECODED) {
                    pd[q] = ld(a.pdirpow + kk);
                } else {
                    pd[q].w = phx_as_float(a.ppow[kk]);
                    pdr[q] = a.pdir[kk];
                }
            }
#pragma unroll
            for (u32 q = 0; q < B; q++)
                keep_loaded(DECODED ? pd[q] : mk(pd[q].w, phx_as_float(pdr[q]), 0.0f, 0.0f));
            if (!DECODED) {
                /* records not decoded once per pass (phx_gather_flux): the
                   same decode_oct at every visit */
#pragma unroll
                for (u32 q = 0; q < B; q++) {
                    F4 dv = decode_oct(pdr[q]);
                    pd[q] = mk(dv.x, dv.y, dv.z, pd[q].w);
                }
            }
#pragma unroll
            for (u32 q = 0; q < B; q++) {
                if (q >= nt)
                    break;
                /* photon_hash.h:84-98: within the radius and facing the hitpoint;
                   one branch, so the sums have one merge point */
                F4 dd = pos - pp[q];
                f32 dsq = dot(dd, dd);
                /* `s + wz > 0` as `s > wz` (gather_cell) */
                bool acc = dsq < r_sq && !(((pd[q].x * sn.x + pd[q].y * sn.y) + pd[q].z * sn.z) > wz);
                if (acc) {
                    f32 wgt = gauss_weight(dsq, two_r_sq, inv_two_r_sq);
                    nf = nf + wgt * decode_rgbe_ldexp(phx_as_uint(pd[q].w));
                    found += wgt;
                }
            }
            k += n1;
            k2 += n2;
        }
    }
    if (px)
        gather_end(a, tid, alpha, fsf, wsf, r_sq, nf, found);
}

/* ---------------- workspace layouts (kernels_trace.hip) ---------------- */
WfState wf_layout(void *ws, u32 n_pixels, u32 direct_samples, bool with_scene);

struct EmitLayout {
    EmitStaging g;
    u32 *offs, *total;
    void *scan_ws;
    EmitWfState st;
};
EmitWfState emit_wf_layout(char *&p, u32 photons, u32 max_bounces);
EmitLayout emit_layout(void *ws, u32 photons, u32 max_bounces);
/* scan of the per-photon deposit counts, compaction to *n_photons + offset,
 * counter advance */
hipError_t launch_emit_compact(const EmitLayout &L, u32 photons, u32 max_photons, phx_float4 *photon_pos,
                               u32 *photon_power, u32 *photon_dir, u32 *n_photons, hipStream_t s,
                               phx_float4 *deposit_rec = nullptr);

} // namespace phx

#endif /* PHX_PASS_BODIES_H */
