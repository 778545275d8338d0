/*
 * kernels_batch.hip — one SPPM pass over many scenes at once.
 *
 * Every pixel, photon and deposit runs the same per-item code as in a
 * one-scene pass (pass_bodies.h) with its own scene's data, PCG stream and
 * grid, so each scene's results are bit-identical to rendering it alone.
 * What changes is how work is spread:
 *
 *  - camera paths: a persistent kernel pulls 128-pixel chunks (each inside one
 *    scene) from an atomic ticket, so large and small images share the GPU
 *    and blocks never relaunch;
 *  - shadow rays and photon bounces come from queues that mix scenes; a
 *    ballot-driven loop serves one scene per iteration with that scene's data
 *    wave-uniform (almost always one iteration: queues are filled wave by wave
 *    and scenes are contiguous);
 *  - the photon map is segmented: scene s owns grid buckets
 *    [bucket_base_s, bucket_base_s + g_s³ + 1), deposits are packed scene by
 *    scene in canonical order, and one scan and one stable sort over all
 *    buckets give every scene the cell ranges of a one-scene pass shifted by
 *    its first deposit (the bucket total may exceed 4096²);
 *  - gather: one wave per 8×8 tile of any scene, tiles assigned XCD by XCD.
 *
 * Pixels are traced and photons emitted in groups of whole scenes so the
 * wavefront workspaces stay bounded.  All sizes are static and all counts
 * live on the device, so the pass can be captured into a hipGraph.
 */
#include "pass_bodies.h"

namespace phx {

/* largest s in [lo, hi) with base[s] <= i (base non-decreasing) */
PHX_DEV u32 find_scene(const u32 *base, u32 lo, u32 hi, u32 i) {
    while (hi - lo > 1u) {
        u32 mid = (lo + hi) >> 1;
        if (base[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

/* Calls f(s0) once per distinct scene s0 among the active lanes, with s0
 * wave-uniform, on the lanes whose scene is s0.  The loop is driven by
 * ballots: the lowest pending lane leads, every lane of its scene is served
 * and leaves the pending set, so each iteration removes at least the leader
 * and the loop ends after as many iterations as there are distinct scenes. */
template <class F> PHX_DEV void for_each_scene(u32 s, F &&f) {
    unsigned long long pending = __ballot(1);
    while (pending) {
        int leader = __ffsll((long long)pending) - 1;
        u32 s0 = __builtin_amdgcn_readfirstlane(__shfl(s, leader, 64));
        bool mine = s == s0;
        if (mine)
            f(s0);
        pending &= ~__ballot(mine);
    }
}

/* ---------------- trace ---------------- */
struct BCamArgs {
    const BatchScene *scenes;
    const u32 *chunk_scene, *round_seed;
    u32 chunk0, chunks, px0;
    u32 *ticket;
    WfState st;
    CamOut o;
};

__global__ __launch_bounds__(WF_CAM_BLOCK) void k_bcam(BCamArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * WF_CAM_BLOCK];
    __shared__ f32 s_ior[IOR_STACK_SIZE * WF_CAM_BLOCK];
    __shared__ u32 s_chunk;
    for (;;) {
        if (threadIdx.x == 0)
            s_chunk = atomicAdd(a.ticket, 1u);
        __syncthreads();
        u32 c = __builtin_amdgcn_readfirstlane(s_chunk);
        __syncthreads(); /* every lane has read s_chunk before the next ticket */
        if (c >= a.chunks)
            break;
        c += a.chunk0;
        u32 s = __builtin_amdgcn_readfirstlane(a.chunk_scene[c]);
        const BatchScene &B = a.scenes[s];
        u32 local = (c - B.chunk_base) * WF_CAM_BLOCK + threadIdx.x;
        bool live = local < B.n_pixels;
        u32 out = B.px_base + local;
        camera_path(B.sc, B.cam, B.width, B.height, a.round_seed[s], B.direct_samples, live, local, out, out - a.px0,
                    s, a.st, a.o, s_stack + threadIdx.x, s_ior + threadIdx.x, WF_CAM_BLOCK, B.direct_samples);
    }
}

struct BShadowArgs {
    const BatchScene *scenes;
    WfState st;
};

/* six waves per SIMD like k_trace_shadow (80 VGPRs, some spills), and five
   for the emission kernels: the batch64 pass 302 -> 286 ms against the
   unconstrained 126-VGPR build */
__global__ __launch_bounds__(WF_SHADOW_BLOCK) __attribute__((amdgpu_waves_per_eu(6))) void k_bshadow(BShadowArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * WF_SHADOW_BLOCK];
    u32 *stk = s_stack + threadIdx.x;
    const WfState &S = a.st;
    u32 count = *S.counter;
    for (u32 slot = blockIdx.x * WF_SHADOW_BLOCK + threadIdx.x; slot < count; slot += gridDim.x * WF_SHADOW_BLOCK) {
        if (S.sh_ev[slot] == NO_EVENT) { /* no sample drawn for this slot */
            st(S.sh_tr + slot, BLACK);
            continue;
        }
        for_each_scene(S.sh_scene[slot], [&](u32 s0) {
            const BatchScene &B = a.scenes[s0];
            if (B.any_hit)
                shadow_one<true>(B.sc, S, slot, stk);
            else
                shadow_one<false>(B.sc, S, slot, stk);
        });
    }
}

struct BResolveArgs {
    const BatchScene *scenes;
    const u32 *chunk_scene;
    u32 chunk0, px0;
    WfState st;
    phx_float4 *hp_direct, *hp_emission;
};

__global__ __launch_bounds__(WF_CAM_BLOCK) void k_bresolve(BResolveArgs a) {
    u32 c = a.chunk0 + blockIdx.x;
    u32 s = __builtin_amdgcn_readfirstlane(a.chunk_scene[c]);
    const BatchScene &B = a.scenes[s];
    u32 local = (c - B.chunk_base) * WF_CAM_BLOCK + threadIdx.x;
    if (local >= B.n_pixels)
        return;
    u32 out = B.px_base + local;
    resolve_pixel(a.st, B.direct_samples, out - a.px0, out, a.hp_direct, a.hp_emission);
}

/* ---------------- emission ---------------- */
struct BEmitArgs {
    const BatchScene *scenes;
    const u32 *ph_base, *round_seed;
    u32 s0, s1, first, photons, depth;
    EmitWfState st;
    EmitStaging g;
};

__global__ __launch_bounds__(256) void k_bemit_start(BEmitArgs a) {
    u32 li = blockIdx.x * 256 + threadIdx.x;
    if (li >= a.photons)
        return;
    u32 gi = a.first + li;
    for_each_scene(find_scene(a.ph_base, a.s0, a.s1, gi), [&](u32 s0) {
        const BatchScene &B = a.scenes[s0];
        emit_start_one(B.sc, a.round_seed[s0], gi - B.ph_base, li, a.st, a.photons);
    });
}

/* bounce a.depth of every live photon of the group; survivors go to the next queue */
__global__ __launch_bounds__(EMIT_WF_BLOCK) __attribute__((amdgpu_waves_per_eu(5))) void k_bemit_bounce(BEmitArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * EMIT_WF_BLOCK];
    u32 *stk = s_stack + threadIdx.x;
    const EmitWfState &S = a.st;
    const u32 depth = a.depth;
    const u32 k = blockIdx.x % EMIT_NQ, step = (gridDim.x / EMIT_NQ) * EMIT_WF_BLOCK;
    const u32 n = emit_qlen(S, depth, k, a.photons);
    u32 *cnt_out = emit_qcounter(S.qcount, depth + 1, k);
    u32 *qout = S.queue[(depth + 1) & 1] + (size_t)k * S.slice;
    for (u32 base = (blockIdx.x / EMIT_NQ) * EMIT_WF_BLOCK; base < n; base += step) {
        u32 i = base + threadIdx.x;
        bool keep = false;
        u32 li = 0;
        if (i < n) {
            li = emit_qentry(S, depth, k, i);
            for_each_scene(find_scene(a.ph_base, a.s0, a.s1, a.first + li), [&](u32 s0) {
                const BatchScene &B = a.scenes[s0];
                if (depth < B.max_bounces)
                    keep = emit_advance(B.sc, S, a.g, li, depth, B.max_bounces, true, stk);
            });
        }
        wave_append(keep, li, cnt_out, qout);
    }
}

/* the rest of every path still live at bounce a.depth, in one launch */
__global__ __launch_bounds__(EMIT_WF_BLOCK) __attribute__((amdgpu_waves_per_eu(5))) void k_bemit_finish(BEmitArgs a) {
    __shared__ u32 s_stack[BVH_STACK_SIZE * EMIT_WF_BLOCK];
    u32 *stk = s_stack + threadIdx.x;
    const EmitWfState &S = a.st;
    const u32 k = blockIdx.x % EMIT_NQ, step = (gridDim.x / EMIT_NQ) * EMIT_WF_BLOCK;
    const u32 n = emit_qlen(S, a.depth, k, a.photons);
    for (u32 i = (blockIdx.x / EMIT_NQ) * EMIT_WF_BLOCK + threadIdx.x; i < n; i += step) {
        u32 li = emit_qentry(S, a.depth, k, i);
        for_each_scene(find_scene(a.ph_base, a.s0, a.s1, a.first + li), [&](u32 s0) {
            const BatchScene &B = a.scenes[s0];
            if (a.depth < B.max_bounces)
                emit_advance(B.sc, S, a.g, li, a.depth, B.max_bounces, false, stk);
        });
    }
}

/* first deposit of each scene of the group: counter + scanned offset of its
 * first photon (before the counter advances) */
__global__ void k_bdep_base(const u32 *__restrict__ ph_base, u32 s0, u32 s1, u32 first,
                            const u32 *__restrict__ offs, const u32 *__restrict__ counter, u32 *__restrict__ dep_base) {
    for (u32 s = s0 + threadIdx.x; s < s1; s += blockDim.x)
        dep_base[s] = *counter + offs[ph_base[s] - first];
}

/* group photon li's deposits to counter + offs[li] + k (no truncation: the
 * buffer holds max_bounces deposits per photon of every scene) */
__global__ __launch_bounds__(256) void k_bemit_compact(EmitStaging g, const u32 *__restrict__ count,
                                                       const u32 *__restrict__ offs, u32 photons,
                                                       const u32 *__restrict__ counter, phx_float4 *__restrict__ pos,
                                                       u32 *__restrict__ power, u32 *__restrict__ dirs) {
    u32 li = blockIdx.x * 256 + threadIdx.x;
    if (li >= photons)
        return;
    u32 c = count[li];
    size_t dst = (size_t)*counter + offs[li];
    for (u32 k = 0; k < c; k++, dst++) {
        size_t slot = (size_t)k * g.stride + li;
        pos[dst] = g.pos[slot];
        power[dst] = g.power[slot];
        dirs[dst] = g.dir[slot];
    }
}

__global__ void k_bcounter_add(u32 *counter, const u32 *total) { *counter += *total; }

/* end of emission: sentinel, per-scene counts and passes without deposits */
__global__ void k_bdep_end(u32 n_scenes, const u32 *__restrict__ counter, u32 *__restrict__ dep_base,
                           u32 *__restrict__ counts, u32 *__restrict__ zero_passes) {
    if (threadIdx.x == 0)
        dep_base[n_scenes] = *counter;
    __syncthreads();
    for (u32 s = threadIdx.x; s < n_scenes; s += blockDim.x) {
        u32 c = dep_base[s + 1] - dep_base[s];
        counts[s] = c;
        if (c == 0)
            zero_passes[s] += 1u;
    }
}

/* ---------------- photon map ---------------- */
/* hash_photons.cl per deposit with its scene's grid, key offset by the
 * scene's first bucket */
__global__ __launch_bounds__(256) void k_bhash(const BatchScene *__restrict__ scenes, u32 n_scenes,
                                               const u32 *__restrict__ dep_base, const phx_float4 *__restrict__ pos,
                                               const u32 *__restrict__ d_n, u32 max_n, u32 *__restrict__ hashes,
                                               u32 *__restrict__ counts) {
    u32 n = *d_n;
    n = n < max_n ? n : max_n;
    for (u32 i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        u32 s = find_scene(dep_base, 0u, n_scenes, i);
        const BatchScene &B = scenes[s];
        phx_float4 p = pos[i];
        /* octant key of the batch-wide bucket (8 keys per bucket) */
        u32 key = B.bucket_base * 8u + get_hash_octant(p.x, p.y, p.z, B.hash);
        hashes[i] = key;
        atomicAdd(&counts[key >> 3], 1u);
    }
}

/* ---------------- gather ---------------- */
struct BGatherArgs {
    const BatchScene *scenes;
    const u32 *tile_scene;
    u32 total_tiles;
    GatherArgs ga;
};

__global__ __launch_bounds__(256) void k_bgather(BGatherArgs a) {
    u32 t = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    if (t >= a.total_tiles)
        return;
    u32 lane = threadIdx.x & 63u;
    u32 s = __builtin_amdgcn_readfirstlane(a.tile_scene[t]);
    const BatchScene &B = a.scenes[s];
    u32 lt = t - B.tile_base;
    u32 x = (lt % B.tiles_x) * 8 + (lane & 7u), y = (lt / B.tiles_x) * 8 + (lane >> 3);
    gather_pixel_flat(a.ga, x < B.width && y < B.height, B.px_base + y * B.width + x, B.hash, B.alpha,
                      B.bucket_base);
}

/* ---------------- pass ---------------- */
size_t batch_trace_workspace_bytes(u32 max_pixels, u32 max_direct_samples) {
    return trace_wavefront_workspace_bytes(max_pixels, max_direct_samples, true);
}

size_t batch_emit_workspace_bytes(u32 max_photons, u32 max_bounces) {
    return emit_workspace_bytes(max_photons, max_bounces);
}

uint32_t batch_camera_blocks() {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_bcam, WF_CAM_BLOCK, 0) != hipSuccess || per_cu <= 0)
        per_cu = 4;
    return (uint32_t)(cus * per_cu);
}

#define BTRY(expr)                                                                                                     \
    do {                                                                                                               \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess)                                                                                          \
            return e_;                                                                                                 \
    } while (0)
#define BLAUNCH() BTRY(hipGetLastError())

hipError_t launch_batch_pass(const BatchPlan &P, const BatchStreams &B) {
    u32 S = P.n_scenes;
    hipStream_t s = B.s, side = B.side;
    hipEvent_t ev_fork = B.ev_fork, ev_join = B.ev_join;
    BTRY(launch_next_round_seed(P.rng, P.round_seed, S, s));
    BTRY(launch_fill_u32(P.tickets, 0u, P.trace_groups.size(), s));
    /* the emission and the photon map need only the round seeds: they run on
       the side stream while the camera trace runs on s */
    BTRY(hipEventRecord(ev_fork, s));
    BTRY(hipStreamWaitEvent(side, ev_fork, 0));
    {
        hipStream_t main = s;
        s = side;
        /* emit_photons, group by group, packed scene by scene */
        BTRY(launch_fill_u32(P.dep_counter, 0u, 1, s));
        for (const BatchGroup &G : P.emit_groups) {
            EmitLayout L = emit_layout(P.emit_ws, G.count, P.max_bounces);
            BTRY(launch_fill_u32(L.st.qcount, 0u, emit_counter_words(P.max_bounces), s));
            BEmitArgs a{P.scenes, P.ph_base, P.round_seed, G.s0, G.s1, G.first, G.count, 0u, L.st, L.g};
            hipLaunchKernelGGL(k_bemit_start, dim3((G.count + 255) / 256), dim3(256), 0, s, a);
            BLAUNCH();
            u32 blocks = emit_queue_blocks(EMIT_WF_BLOCK, L.st.slice);
            u32 wave_bounces = P.max_bounces < EMIT_WF_BOUNCES ? P.max_bounces : EMIT_WF_BOUNCES;
            for (u32 d = 0; d < wave_bounces; d++) {
                a.depth = d;
                hipLaunchKernelGGL(k_bemit_bounce, dim3(blocks), dim3(EMIT_WF_BLOCK), 0, s, a);
                BLAUNCH();
            }
            if (wave_bounces < P.max_bounces) {
                a.depth = wave_bounces;
                hipLaunchKernelGGL(k_bemit_finish, dim3(blocks), dim3(EMIT_WF_BLOCK), 0, s, a);
                BLAUNCH();
            }
            BTRY(launch_scan(L.st.count, G.count, L.offs, nullptr, L.total, L.scan_ws, s));
            hipLaunchKernelGGL(k_bdep_base, dim3(1), dim3(64), 0, s, P.ph_base, G.s0, G.s1, G.first, L.offs,
                               P.dep_counter, P.dep_base);
            BLAUNCH();
            hipLaunchKernelGGL(k_bemit_compact, dim3((G.count + 255) / 256), dim3(256), 0, s, L.g, L.st.count, L.offs,
                               G.count, P.dep_counter, P.dep_pos, P.dep_power, P.dep_dir);
            BLAUNCH();
            hipLaunchKernelGGL(k_bcounter_add, dim3(1), dim3(1), 0, s, P.dep_counter, L.total);
            BLAUNCH();
        }
        hipLaunchKernelGGL(k_bdep_end, dim3(1), dim3(256), 0, s, S, P.dep_counter, P.dep_base, P.counts, P.zero_passes);
        BLAUNCH();
        if (B.h_info) { /* the pass's counts, before the next emission overwrites them */
            BTRY(hipMemcpyAsync(B.h_info, P.counts, S * sizeof(u32), hipMemcpyDeviceToHost, s));
            BTRY(hipMemcpyAsync(B.h_info + S, P.zero_passes, S * sizeof(u32), hipMemcpyDeviceToHost, s));
        }
        if (B.ev_emit)
            BTRY(hipEventRecord(B.ev_emit, s));

        /* hash, scan, stable sort over every scene's buckets */
        BTRY(launch_fill_u32(P.cell_counts, 0u, P.total_buckets, s));
        {
            u32 blocks = (P.dep_cap + 255) / 256;
            if (blocks > 8192)
                blocks = 8192;
            if (blocks == 0)
                blocks = 1;
            hipLaunchKernelGGL(k_bhash, dim3(blocks), dim3(256), 0, s, P.scenes, S, P.dep_base, P.dep_pos,
                               P.dep_counter, P.dep_cap, P.hashes, P.cell_counts);
            BLAUNCH();
        }
        BTRY(launch_scan(P.cell_counts, P.total_buckets, P.cell_start, P.cell_end, nullptr, P.scan_ws, s));
        BTRY(launch_sort(P.dep_pos, P.dep_power, P.dep_dir, P.hashes, P.cell_start, P.cell_end, P.dep_cap,
                         P.total_buckets, P.map_pos, P.map_power, P.map_dir, P.sort_ws, s, P.dep_counter, nullptr,
                         nullptr, P.oct_start));
        BTRY(launch_decode_photons(P.map_power, P.map_dir, P.dep_counter, P.dep_cap, P.map_dirpow, s));
        BTRY(hipEventRecord(ev_join, side));
        s = main;
    }

    /* trace_rays, group by group */
    WfState W = wf_layout(P.trace_ws, P.trace_max_pixels, P.max_direct_samples, true);
    CamOut o{P.hp[0], P.hp[1], P.hp[2], P.hp[3]};
    for (size_t gi = 0; gi < P.trace_groups.size(); gi++) {
        const BatchGroup &G = P.trace_groups[gi];
        BTRY(launch_fill_u32(W.counter, 0u, 1, s));
        BCamArgs ca{P.scenes, P.chunk_scene, P.round_seed, G.chunk0, G.chunks, G.first, P.tickets + gi, W, o};
        u32 blocks = P.cam_blocks < G.chunks ? P.cam_blocks : G.chunks;
        hipLaunchKernelGGL(k_bcam, dim3(blocks), dim3(WF_CAM_BLOCK), 0, s, ca);
        BLAUNCH();
        size_t expect = (size_t)G.count * P.max_direct_samples;
        u32 sblocks = (u32)((expect + WF_SHADOW_BLOCK - 1) / WF_SHADOW_BLOCK);
        if (sblocks > 16384)
            sblocks = 16384;
        if (sblocks == 0)
            sblocks = 1;
        hipLaunchKernelGGL(k_bshadow, dim3(sblocks), dim3(WF_SHADOW_BLOCK), 0, s, BShadowArgs{P.scenes, W});
        BLAUNCH();
        BResolveArgs ra{P.scenes, P.chunk_scene, G.chunk0, G.first, W, P.hp[4], P.hp[5]};
        hipLaunchKernelGGL(k_bresolve, dim3(G.chunks), dim3(WF_CAM_BLOCK), 0, s, ra);
        BLAUNCH();
    }

    if (B.gather) {
        BTRY(hipEventRecord(B.ev_trace, s));
        BTRY(hipStreamWaitEvent(B.gather, B.ev_trace, 0));
        s = B.gather;
    }
    BTRY(hipStreamWaitEvent(s, ev_join, 0));

    /* gather_flux over every tile of every scene */
    GatherArgs ga{P.hp[0], P.hp[1], P.hp[2], P.hp[3], P.hp[4], P.hp[5], 0u, P.map_pos, P.map_power, P.map_dir,
                  P.cell_start, P.cell_end, PhxPhotonHashData{}, 0.0f, P.flux, P.photon_weight, P.radiance,
                  nullptr, nullptr, P.map_dirpow, P.oct_start};
    BGatherArgs gargs{P.scenes, P.tile_scene, P.total_tiles, ga};
    hipLaunchKernelGGL(k_bgather, dim3((P.total_tiles + 3) / 4), dim3(256), 0, s, gargs);
    return hipGetLastError();
}

} // namespace phx
