/*
 * renderer.cpp — the SPPM pass driver on HIP.
 *
 * Mirrors src/renderer/renderer.cpp (Renderer::set_scene/run/trace_rays/
 * emit_photons/hash_and_sort_photons/gather_flux) and the buffer owners of
 * include/buffers/*.hpp (GeometryBuffers, ShadingBuffers, PhotonBuffers,
 * ImageBuffers, RenderingBuffers), with cl::Buffer replaced by RAII device
 * allocations and the OpenCL enqueues replaced by the kernels of this
 * library on one HIP stream.
 *
 * Differences from the reference, all internal: the scan is one kernel, the
 * scatter is the stable sort (canonical photon order), and a pass never waits
 * for the host.  The round seed is drawn on the device from a PCG state kept
 * there, and the deposit count that the reference reads back between emit and
 * hash (renderer.cpp:113-118) stays on the device: hash, sort and cell-end
 * kernels read it.  The count is copied to pinned host memory at the end of
 * the pass; a pass without deposits ("no photon hit", renderer.cpp:117-118)
 * is reported by the next synchronisation.  With PHX_PASS_GRAPH the pass is
 * captured once into a hipGraph and replayed.
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

/* emission workspace allowed for one launch sequence over a whole round */
static const size_t EMIT_ONE_SHOT_BYTES = 4ull << 30;

#define PHX_TRY(expr, what)                                                                                            \
    do {                                                                                                               \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess)                                                                                          \
            return hip_status(e_, what);                                                                               \
    } while (0)

namespace phx {
int check_render_args(const PhxRenderArgs &a, const char *who) {
    std::string w(who);
    if (a.width == 0 || a.height == 0 || a.photons_per_round == 0)
        return fail(PHX_ERR_INVALID_ARG, w + ": zero image or photon count");
    if (a.grid_res == 0 || a.grid_res > 255)
        return fail(PHX_ERR_INVALID_ARG, w + ": grid_res outside 1..255 (cmd_args.hpp:51)");
    if (a.max_photon_bounces == 0 || a.max_photon_bounces > 32)
        return fail(PHX_ERR_INVALID_ARG, w + ": max_bounces outside 1..32 (cmd_args.hpp:49)");
    if (a.trace_mode > 1)
        return fail(PHX_ERR_INVALID_ARG, w + ": trace_mode must be 0 or 1");
    if (a.direct_samples == 0 || a.direct_samples > 128)
        return fail(PHX_ERR_INVALID_ARG, w + ": direct_samples outside 1..128 (cmd_args.hpp:59)");
    if (a.flags & ~(uint32_t)(PHX_PASS_GRAPH | PHX_REFERENCE_BVH | PHX_PASS_SERIAL | PHX_FUSED_TRACE))
        return fail(PHX_ERR_INVALID_ARG, w + ": unknown flags");
    if (a.emit_wave_bounces > 32)
        return fail(PHX_ERR_INVALID_ARG, w + ": emit_wave_bounces outside 0..32");
    for (uint32_t b : {a.camera_block, a.shadow_block, a.emit_block})
        if (b != 0 && b != 64 && b != 128 && b != 256)
            return fail(PHX_ERR_INVALID_ARG, w + ": camera, shadow and emission blocks must be 64, 128 or 256 threads");
    if (a.gather_tile_width > 64 || (a.gather_tile_width & (a.gather_tile_width - 1)) != 0)
        return fail(PHX_ERR_INVALID_ARG, w + ": gather_tile_width must be 0 or a power of two up to 64");
    /* pixel indices, shadow-queue slots (5 NEE events per pixel at most, each
       with direct_samples light samples) and deposit slots are u32 on the device */
    uint64_t px = (uint64_t)a.width * a.height;
    if (px > 0xFFFFFFFFull)
        return fail(PHX_ERR_TOO_LARGE, w + ": width * height exceeds 2^32 pixels");
    if (px * 5u * a.direct_samples > 0xFFFFFFFFull)
        return fail(PHX_ERR_TOO_LARGE, w + ": width * height * 5 * direct_samples exceeds 2^32 shadow slots");
    if ((uint64_t)a.photons_per_round * a.max_photon_bounces > 0xFFFFFFFFull)
        return fail(PHX_ERR_TOO_LARGE, w + ": photons_per_round * max_bounces exceeds 2^32 deposit slots");
    return PHX_OK;
}
} // namespace phx

/* renderer.cpp:22-57 + rendering.cpp:15-30 + geometry.cpp / shading.cpp / image.cpp / photon.cpp */
int PhxRenderer::set_scene(const PhxSceneDesc &sc) {
    hipStream_t s = stream;
    const PhxRenderArgs &a = args;
    int rc = upload_scene(sc, s, (a.flags & PHX_REFERENCE_BVH) != 0, a.bvh_leaf_size, buffers.scene);
    if (rc != PHX_OK)
        return rc;
    shadow_any_hit = buffers.scene.shadow_any_hit;
    const PhxBvhNode &root = buffers.scene.root;

    /* alloc_image_buffers (image.cpp:11-21, zero-filled) */
    ImageBuffers &im = buffers.image_buf;
    im.camera = *sc.camera;
    im.image_width = a.width;
    im.image_height = a.height;
    im.n_pixels = a.width * a.height;
    size_t np = im.n_pixels;
    DevBuf *f4bufs[] = {&im.d_hp_pos,   &im.d_hp_normal, &im.d_hp_throughput, &im.d_hp_base_color,
                        &im.d_hp_direct, &im.d_hp_emission, &im.d_flux,        &im.d_radiance};
    for (DevBuf *b : f4bufs) {
        PHX_TRY(b->alloc(np * sizeof(phx_float4)), "alloc image");
        PHX_TRY(hipMemsetAsync(b->p, 0, b->bytes, s), "clear image");
    }
    PHX_TRY(im.d_photon_weight.alloc(np * sizeof(float)), "alloc image");
    PHX_TRY(hipMemsetAsync(im.d_photon_weight.p, 0, im.d_photon_weight.bytes, s), "clear image");
    if (a.initial_radius > 0.0f) {
        /* explicit initial radius: flux.w starts at r0² instead of 0 */
        std::vector<phx_float4> f(np, phx_float4{0.0f, 0.0f, 0.0f, a.initial_radius * a.initial_radius});
        PHX_TRY(hipMemcpyAsync(im.d_flux.p, f.data(), np * sizeof(phx_float4), hipMemcpyHostToDevice, s), "init flux");
        PHX_TRY(hipStreamSynchronize(s), "init flux");
    }

    /* renderer.cpp:37-52 */
    phx_build_photon_hash_info(&root, a.grid_res, &photon_hash_info);
    bucket_count = a.grid_res * a.grid_res * a.grid_res + 1;
    uint64_t buf = (uint64_t)a.photons_per_round * a.max_photon_bounces;
    if (buf > 0xFFFFFFFFull)
        return fail(PHX_ERR_TOO_LARGE, "set_scene: photons_per_round * max_bounces overflows u32");
    photon_buf_size = (uint32_t)buf;
    PHX_TRY(d_deposit_rec.alloc((size_t)photon_buf_size * 2 * sizeof(phx_float4)), "alloc photons");
    PHX_TRY(d_batch_size.alloc(sizeof(uint32_t)), "alloc counter");
    PhotonBuffers &p = buffers.photon_buf;
    PHX_TRY(p.d_photon_pos.alloc((size_t)photon_buf_size * sizeof(phx_float4)), "alloc photon map");
    PHX_TRY(p.d_photon_power.alloc((size_t)photon_buf_size * sizeof(uint32_t)), "alloc photon map");
    PHX_TRY(p.d_photon_dir.alloc((size_t)photon_buf_size * sizeof(uint32_t)), "alloc photon map");
    PHX_TRY(p.d_photon_hashes.alloc((size_t)photon_buf_size * sizeof(uint32_t)), "alloc photon map");
    PHX_TRY(p.d_photon_dirpow.alloc((size_t)photon_buf_size * sizeof(phx_float4)), "alloc photon map");
    PHX_TRY(p.d_cell_start.alloc((size_t)bucket_count * sizeof(uint32_t)), "alloc cells");
    PHX_TRY(p.d_cell_end.alloc((size_t)bucket_count * sizeof(uint32_t)), "alloc cells");
    PHX_TRY(p.d_oct_start.alloc((size_t)bucket_count * 8u * sizeof(uint32_t)), "alloc cells");
    /* --batch-size (renderer.cpp:103-111) bounds the reference's device memory;
       deposits do not depend on it here (canonical order), so the whole round
       is emitted in one launch sequence whenever its workspace fits
       EMIT_ONE_SHOT_BYTES, and in batches of --batch-size otherwise */
    uint32_t per_batch = std::min(a.photons_per_round, a.max_photons_per_batch ? a.max_photons_per_batch : 1u);
    if (emit_workspace_bytes(a.photons_per_round, a.max_photon_bounces) <= EMIT_ONE_SHOT_BYTES)
        per_batch = a.photons_per_round;
    emit_chunk = per_batch;
    PHX_TRY(d_emit_ws.alloc(emit_workspace_bytes(per_batch, a.max_photon_bounces)), "alloc emit workspace");
    PHX_TRY(d_scan_ws.alloc(scan_workspace_bytes(bucket_count)), "alloc scan workspace");
    PHX_TRY(d_sort_ws.alloc(sort_workspace_bytes(photon_buf_size, bucket_count)), "alloc sort workspace");
    if (a.trace_mode == 0)
        PHX_TRY(d_trace_ws.alloc(trace_wavefront_workspace_bytes(im.n_pixels, a.direct_samples)),
                "alloc trace workspace");

    /* renderer.cpp:54: the round-seed generator, kept on the device */
    uint32_t rng0[3] = {host_pcg_seed(a.seed), 0u, 0u};
    PHX_TRY(d_rng.alloc(sizeof(rng0)), "alloc seeds");
    PHX_TRY(hipMemcpyAsync(d_rng.p, rng0, sizeof(rng0), hipMemcpyHostToDevice, s), "upload seeds");
    PHX_TRY(d_info.alloc(2 * sizeof(uint32_t)), "alloc pass info");
    PHX_TRY(hipMemsetAsync(d_info.p, 0, d_info.bytes, s), "clear pass info");
    PHX_TRY(hipStreamSynchronize(s), "set_scene");
    return PHX_OK;
}

/* the six hitpoint buffers of a set (0: in `buffers`, 1: alt_hp) */
void PhxRenderer::hp_bufs(int set, DevBuf *out[6]) {
    ImageBuffers &i = buffers.image_buf;
    DevBuf *b[6] = {&i.d_hp_pos, &i.d_hp_normal, &i.d_hp_throughput, &i.d_hp_base_color, &i.d_hp_direct,
                    &i.d_hp_emission};
    for (int k = 0; k < 6; k++)
        out[k] = set ? &alt_hp[k] : b[k];
}

/* the sorted photon map of a set: pos, power, dir, dirpow, cell_start, cell_end, oct_start */
void PhxRenderer::map_bufs(int set, DevBuf *out[7]) {
    PhotonBuffers &p = buffers.photon_buf;
    DevBuf *b[7] = {&p.d_photon_pos, &p.d_photon_power, &p.d_photon_dir, &p.d_photon_dirpow, &p.d_cell_start,
                    &p.d_cell_end,   &p.d_oct_start};
    for (int k = 0; k < 7; k++)
        out[k] = set ? &alt_map[k] : b[k];
}

/* renderer.cpp:90-95 */
int PhxRenderer::trace_rays(int set) {
    const ImageBuffers &i = buffers.image_buf;
    DevBuf *h[6];
    hp_bufs(set, h);
    if (args.trace_mode == 0)
        return hip_status(launch_trace_wavefront(i.camera, i.image_width, i.image_height, 0u, args.direct_samples,
                                                 buffers.scene_args(), shadow_any_hit, h[0]->as<phx_float4>(),
                                                 h[1]->as<phx_float4>(), h[2]->as<phx_float4>(),
                                                 h[3]->as<phx_float4>(), h[4]->as<phx_float4>(),
                                                 h[5]->as<phx_float4>(), d_trace_ws.p, stream, seed_ptr(),
                                                 &shape),
                          "trace_rays");
    return hip_status(launch_trace(i.camera, i.image_width, i.image_height, 0u, args.direct_samples,
                                   buffers.scene_args(), h[0]->as<phx_float4>(), h[1]->as<phx_float4>(),
                                   h[2]->as<phx_float4>(), h[3]->as<phx_float4>(), h[4]->as<phx_float4>(),
                                   h[5]->as<phx_float4>(), stream, seed_ptr()),
                      "trace_rays");
}

/* renderer.cpp:97-121; the count stays on the device (clamped to the buffer
 * by its readers, as renderer.cpp:115 clamps it) */
int PhxRenderer::emit_photons(hipStream_t s) {
    PHX_TRY(launch_fill_u32(d_batch_size.as<uint32_t>(), 0u, 1, s), "emit_photons: reset counter");
    uint32_t ppr = args.photons_per_round;
    uint32_t per_batch = emit_chunk;
    SceneArgs sc = buffers.scene_args();
    for (uint32_t off = 0; off < ppr; off += per_batch) {
        uint32_t to_emit = std::min(per_batch, ppr - off);
        PHX_TRY(launch_emit(sc, photon_buf_size, off, to_emit, args.max_photon_bounces, 0u, nullptr, nullptr, nullptr,
                            d_batch_size.as<uint32_t>(), d_emit_ws.p, s, seed_ptr(), args.emit_wave_bounces, &shape,
                            d_deposit_rec.as<phx_float4>()),
                "emit_photons");
    }
    PHX_TRY(launch_note_count(d_batch_size.as<uint32_t>(), d_info.as<uint32_t>(), s), "emit_photons: count");
    return PHX_OK;
}

/* renderer.cpp:123-153, sized by the device-side count */
int PhxRenderer::hash_and_sort_photons(hipStream_t s, int set) {
    PhotonBuffers &p = buffers.photon_buf;
    DevBuf *m[7];
    map_bufs(set, m);
    const uint32_t *d_n = d_batch_size.as<uint32_t>();
    uint32_t n_max = photon_buf_size;
    /* keys only: the stable sort orders the records, and the cell ranges
       (the scan of the cell counts, renderer.cpp:134-147) are read off the
       sorted keys */
    const phx_float4 *rec = d_deposit_rec.as<phx_float4>();
    PHX_TRY(launch_hash_dev(rec, d_n, n_max, photon_hash_info, p.d_photon_hashes.as<uint32_t>(), nullptr, s, true, 2u),
            "hash_photons");
    /* the last radix pass moves the records and writes each one's decoded
       direction (the gather's input) */
    PHX_TRY(launch_sort(nullptr, nullptr, nullptr, p.d_photon_hashes.as<uint32_t>(), nullptr, m[5]->as<uint32_t>(),
                        n_max, bucket_count, m[0]->as<phx_float4>(), m[1]->as<uint32_t>(), m[2]->as<uint32_t>(),
                        d_sort_ws.p, s, d_n, sort_perm, m[4]->as<uint32_t>(), m[6]->as<uint32_t>(), rec,
                        m[3]->as<phx_float4>()),
            "sort_photons");
    return PHX_OK;
}

/* renderer.cpp:155-161 */
int PhxRenderer::gather_flux(hipStream_t s, int set) {
    const ImageBuffers &i = buffers.image_buf;
    DevBuf *h[6], *m[7];
    hp_bufs(set, h);
    map_bufs(set, m);
    return hip_status(launch_gather(h[0]->as<phx_float4>(), h[1]->as<phx_float4>(), h[2]->as<phx_float4>(),
                                    h[3]->as<phx_float4>(), h[4]->as<phx_float4>(), h[5]->as<phx_float4>(),
                                    i.n_pixels, m[0]->as<phx_float4>(), m[1]->as<uint32_t>(), m[2]->as<uint32_t>(),
                                    m[4]->as<uint32_t>(), m[5]->as<uint32_t>(), photon_hash_info, args.sppm_alpha,
                                    i.d_flux.as<phx_float4>(), i.d_photon_weight.as<float>(),
                                    i.d_radiance.as<phx_float4>(), i.image_width, s, rec_ratio, rec_r2,
                                    m[3]->as<phx_float4>(), m[6]->as<uint32_t>(), shape.gather_tile_w),
                      "gather_flux");
}

/* the second set of hitpoint and photon-map buffers, on the first pipelined
 * pass (cleared, like the first set) */
int PhxRenderer::alloc_alt() {
    if (alt_ready)
        return PHX_OK;
    DevBuf *h[6], *m[7];
    hp_bufs(0, h);
    map_bufs(0, m);
    for (int k = 0; k < 6; k++) {
        PHX_TRY(alt_hp[k].alloc(h[k]->bytes), "alloc second hitpoint set");
        PHX_TRY(hipMemsetAsync(alt_hp[k].p, 0, alt_hp[k].bytes, stream), "clear second hitpoint set");
    }
    for (int k = 0; k < 7; k++) {
        PHX_TRY(alt_map[k].alloc(m[k]->bytes), "alloc second photon map");
        PHX_TRY(hipMemsetAsync(alt_map[k].p, 0, alt_map[k].bytes, stream), "clear second photon map");
    }
    alt_ready = true;
    return PHX_OK;
}

/* after a pipelined pass: its set becomes the one in `buffers` */
void PhxRenderer::swap_sets() {
    DevBuf *h[6], *m[7];
    hp_bufs(0, h);
    map_bufs(0, m);
    for (int k = 0; k < 6; k++)
        h[k]->swap(alt_hp[k]);
    for (int k = 0; k < 7; k++)
        m[k]->swap(alt_map[k]);
    std::swap(ev_read[0], ev_read[1]);
    std::swap(read_pending[0], read_pending[1]);
}

/* waits for every stream of the renderer */
hipError_t PhxRenderer::sync_all() {
    for (hipStream_t s : {stream, side, gstream}) {
        hipError_t e = s ? hipStreamSynchronize(s) : hipSuccess;
        if (e != hipSuccess)
            return e;
    }
    read_pending[0] = read_pending[1] = false;
    emit_pending[0] = emit_pending[1] = false;
    return hipSuccess;
}

/* renderer.cpp:64-78, one round, enqueued without host synchronisation;
 * stage events are recorded between the stages when `events` is set.
 *
 * Serial form: round seed, then the emission and the photon map on `side`
 * beside the camera trace on `stream`, then the gather on `stream` after both.
 * Pipelined form (`pipe`): the same, with the trace and the photon map
 * written into the second set and the gather on `gstream`, so that the next
 * pass's seed, trace and emission do not wait for it.  Ordering:
 *   - pipelined passes alternate between two round-seed slots: the round
 *     seed waits for the emission of the pass before last, the last reader
 *     of its slot (serial passes use slot 0 and wait for every emission);
 *   - the trace and the photon map overwrite the set the gather before last
 *     read: `stream` waits for that gather before the seed, and `side` follows
 *     `stream` from the seed on;
 *   - the gather waits for the trace and the photon map; gathers run in order
 *     on `gstream`, and a serial pass first waits for every pending gather. */
int PhxRenderer::enqueue_pass(bool events, bool pipe) {
    int rc;
    const int set = pipe ? 1 : 0;
    /* serial passes use seed slot 0 */
    seed_slot = pipe ? rounds & 1u : 0u;
    for (uint32_t k = 0; k < 2; k++)
        if (emit_pending[k] && (!pipe || k == seed_slot)) {
            PHX_TRY(hipStreamWaitEvent(stream, ev_emit_done[k], 0), "event");
            emit_pending[k] = false;
        }
    for (int k = pipe ? 1 : 0; k < 2; k++)
        if (read_pending[k]) {
            PHX_TRY(hipStreamWaitEvent(stream, ev_read[k], 0), "event");
            read_pending[k] = false;
        }
    PHX_TRY(launch_next_round_seed(d_rng.as<uint32_t>(), seed_ptr(), 1, stream), "round seed");
    PHX_TRY(hipEventRecord(ev_fork, stream), "event");
    PHX_TRY(hipStreamWaitEvent(side, ev_fork, 0), "event");
    if (events)
        PHX_TRY(hipEventRecord(ev_emit, side), "event");
    if ((rc = emit_photons(side)) != PHX_OK)
        return rc;
    if (events)
        PHX_TRY(hipEventRecord(ev[2], side), "event");
    if (pipe) {
        PHX_TRY(hipMemcpyAsync(h_info, d_info.p, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, side), "pass info");
        PHX_TRY(hipEventRecord(ev_info, side), "event");
        PHX_TRY(hipEventRecord(ev_emit_done[seed_slot], side), "event");
        emit_pending[seed_slot] = true;
    }
    if ((rc = hash_and_sort_photons(side, set)) != PHX_OK)
        return rc;
    if (events)
        PHX_TRY(hipEventRecord(ev[3], side), "event");
    PHX_TRY(hipEventRecord(ev_join, side), "event");
    if (events)
        PHX_TRY(hipEventRecord(ev_trace, stream), "event");
    if ((rc = trace_rays(set)) != PHX_OK)
        return rc;
    if (events)
        PHX_TRY(hipEventRecord(ev[1], stream), "event");
    hipEvent_t *ring = ev_gring[gring_count % GATHER_RING];
    if (!pipe) {
        PHX_TRY(hipStreamWaitEvent(stream, ev_join, 0), "event");
        if (events) {
            PHX_TRY(hipEventRecord(ev_gather, stream), "event");
            PHX_TRY(hipEventRecord(ring[0], stream), "event");
        }
        if ((rc = gather_flux(stream, 0)) != PHX_OK)
            return rc;
        if (events) {
            PHX_TRY(hipEventRecord(ring[1], stream), "event");
            gring_count++;
        }
        return PHX_OK;
    }
    PHX_TRY(hipEventRecord(ev_trace_done, stream), "event");
    PHX_TRY(hipStreamWaitEvent(gstream, ev_trace_done, 0), "event");
    PHX_TRY(hipStreamWaitEvent(gstream, ev_join, 0), "event");
    if (events) {
        PHX_TRY(hipEventRecord(ev_gather, gstream), "event");
        PHX_TRY(hipEventRecord(ring[0], gstream), "event");
    }
    if ((rc = gather_flux(gstream, 1)) != PHX_OK)
        return rc;
    if (events) {
        PHX_TRY(hipEventRecord(ring[1], gstream), "event");
        gring_count++;
        PHX_TRY(hipEventRecord(ev[4], gstream), "event");
    }
    PHX_TRY(hipEventRecord(ev_read[1], gstream), "event");
    read_pending[1] = true;
    swap_sets();
    return PHX_OK;
}

/* records one pass into a graph (kernel nodes only: counters and flags are
 * cleared by fill kernels, never by memset nodes) */
int PhxRenderer::capture_pass() {
    PHX_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "graph capture");
    int rc = enqueue_pass(false, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(stream, &g);
    if (rc != PHX_OK) {
        if (g)
            (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess)
        return hip_status(e, "graph capture");
    graph = g;
    return hip_status(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0), "graph instantiate");
}

int PhxRenderer::run_pass() {
    int rc;
    bool use_graph = (args.flags & PHX_PASS_GRAPH) != 0;
    /* the adjoint driver's recording passes stay serial */
    bool pipe = !use_graph && !(args.flags & PHX_PASS_SERIAL) && !rec_ratio && !sort_perm;
    if (use_graph && !graph_exec && (rc = capture_pass()) != PHX_OK)
        return rc;
    if (pipe && (rc = alloc_alt()) != PHX_OK)
        return rc;
    PHX_TRY(hipEventRecord(ev[0], stream), "event");
    if (use_graph) {
        PHX_TRY(hipGraphLaunch(graph_exec, stream), "graph launch");
    } else if ((rc = enqueue_pass(true, pipe)) != PHX_OK) {
        return rc;
    }
    if (!pipe) {
        PHX_TRY(hipEventRecord(ev[4], stream), "event");
        PHX_TRY(hipMemcpyAsync(h_info, d_info.p, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream),
                "pass info");
        PHX_TRY(hipEventRecord(ev_info, stream), "event");
    }
    info_pending = true;
    timed = !use_graph;
    rounds++;
    total_photons_emitted += args.photons_per_round;
    return PHX_OK;
}

/* waits for the last pass's deposit count */
int PhxRenderer::collect() {
    if (!info_pending)
        return PHX_OK;
    PHX_TRY(hipEventSynchronize(ev_info), "pass info");
    info_pending = false;
    buffers.photon_buf.n_photons = std::min(h_info[0], photon_buf_size);
    return PHX_OK;
}

/* reports, once, passes that deposited nothing (renderer.cpp:117-118) */
int PhxRenderer::check_passes() {
    int rc = collect();
    if (rc != PHX_OK)
        return rc;
    if (h_info[1] > zero_reported) {
        zero_reported = h_info[1];
        return fail(PHX_ERR_NO_PHOTONS, "no photon hit");
    }
    return PHX_OK;
}

extern "C" {

int phx_renderer_create(const PhxRenderArgs *args, const PhxSceneDesc *scene, PhxRenderer **out) {
    if (!args || !scene || !out)
        return fail(PHX_ERR_INVALID_ARG, "phx_renderer_create: null argument");
    *out = nullptr;
    if (check_scene_desc(*scene, "phx_renderer_create") != PHX_OK)
        return PHX_ERR_INVALID_ARG;
    int rc = check_render_args(*args, "phx_renderer_create");
    if (rc != PHX_OK)
        return rc;
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
        return fail(PHX_ERR_NO_DEVICE, "phx_renderer_create: no HIP device");
    std::unique_ptr<PhxRenderer> r(new PhxRenderer());
    r->args = *args;
    if (args->camera_block)
        r->shape.camera = args->camera_block;
    if (args->shadow_block)
        r->shape.shadow = args->shadow_block;
    if (args->emit_block)
        r->shape.emit = args->emit_block;
    if (args->gather_tile_width)
        r->shape.gather_tile_w = args->gather_tile_width;
    r->shape.split_trace = (args->flags & PHX_FUSED_TRACE) ? 0u : 1u;
    hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&r->side, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&r->gstream, hipStreamNonBlocking);
    if (e != hipSuccess)
        return hip_status(e, "phx_renderer_create: stream");
    for (hipEvent_t &ev : r->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess)
            return hip_status(e, "phx_renderer_create: event");
    for (hipEvent_t *ev : {&r->ev_emit, &r->ev_gather, &r->ev_trace})
        if ((e = hipEventCreate(ev)) != hipSuccess)
            return hip_status(e, "phx_renderer_create: event");
    for (auto &pair : r->ev_gring)
        for (hipEvent_t &ev : pair)
            if ((e = hipEventCreate(&ev)) != hipSuccess)
                return hip_status(e, "phx_renderer_create: event");
    for (hipEvent_t *ev : {&r->ev_info, &r->ev_fork, &r->ev_join, &r->ev_trace_done, &r->ev_emit_done[0], &r->ev_emit_done[1],
                           &r->ev_read[0], &r->ev_read[1]})
        if ((e = hipEventCreateWithFlags(ev, hipEventDisableTiming)) != hipSuccess)
            return hip_status(e, "phx_renderer_create: event");
    if ((e = hipHostMalloc((void **)&r->h_info, 2 * sizeof(uint32_t), hipHostMallocDefault)) != hipSuccess)
        return hip_status(e, "phx_renderer_create: pinned info");
    r->h_info[0] = r->h_info[1] = 0;
    rc = r->set_scene(*scene);
    if (rc != PHX_OK)
        return rc;
    *out = r.release();
    return PHX_OK;
}

void phx_renderer_destroy(PhxRenderer *r) { delete r; }

int phx_renderer_run_pass(PhxRenderer *r) {
    if (!r)
        return fail(PHX_ERR_INVALID_ARG, "phx_renderer_run_pass: null renderer");
    return r->run_pass();
}

int phx_renderer_synchronize(PhxRenderer *r) {
    if (!r)
        return fail(PHX_ERR_INVALID_ARG, "null renderer");
    PHX_TRY(r->sync_all(), "phx_renderer_synchronize");
    return r->check_passes();
}

int phx_renderer_last_pass_ms(PhxRenderer *r, float *out_ms) {
    if (!r || !out_ms)
        return fail(PHX_ERR_INVALID_ARG, "null argument");
    *out_ms = 0.0f;
    if (r->rounds == 0)
        return PHX_OK;
    PHX_TRY(hipEventSynchronize(r->ev[4]), "pass timing");
    return hip_status(hipEventElapsedTime(out_ms, r->ev[0], r->ev[4]), "pass timing");
}

int phx_renderer_last_stage_ms(PhxRenderer *r, float out_ms[4]) {
    if (!r || !out_ms)
        return fail(PHX_ERR_INVALID_ARG, "null argument");
    if (!r->timed) {
        for (int k = 0; k < 4; k++)
            out_ms[k] = 0.0f;
        return PHX_OK;
    }
    hipError_t e = hipEventSynchronize(r->ev[4]);
    if (e != hipSuccess)
        return hip_status(e, "stage timing");
    /* the stages overlap: each is timed on its own stream */
    const hipEvent_t from[4] = {r->ev_trace, r->ev_emit, r->ev[2], r->ev_gather};
    const hipEvent_t to[4] = {r->ev[1], r->ev[2], r->ev[3], r->ev[4]};
    for (int k = 0; k < 4; k++)
        if ((e = hipEventElapsedTime(&out_ms[k], from[k], to[k])) != hipSuccess)
            return hip_status(e, "stage timing");
    return PHX_OK;
}

int phx_renderer_gather_ms(PhxRenderer *r, float *out_ms, uint32_t max, uint32_t *n_out) {
    if (!r || !n_out || (max && !out_ms))
        return fail(PHX_ERR_INVALID_ARG, "null argument");
    uint32_t have = std::min(r->gring_count, PhxRenderer::GATHER_RING);
    uint32_t n = std::min(have, max);
    *n_out = n;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t slot = (r->gring_count - n + k) % PhxRenderer::GATHER_RING;
        hipError_t e = hipEventSynchronize(r->ev_gring[slot][1]);
        if (e == hipSuccess)
            e = hipEventElapsedTime(&out_ms[k], r->ev_gring[slot][0], r->ev_gring[slot][1]);
        if (e != hipSuccess)
            return hip_status(e, "gather timing");
    }
    return PHX_OK;
}

int phx_renderer_reset_gather_ms(PhxRenderer *r) {
    if (!r)
        return fail(PHX_ERR_INVALID_ARG, "null argument");
    r->gring_count = 0;
    return PHX_OK;
}

uint32_t phx_renderer_last_photon_count(PhxRenderer *r) {
    if (!r)
        return 0u;
    (void)r->collect();
    return r->rounds ? r->h_info[0] : 0u;
}
uint32_t phx_renderer_rounds_done(PhxRenderer *r) { return r ? r->rounds : 0u; }

int phx_renderer_read_state(PhxRenderer *r, phx_float4 *flux, float *photon_weight, phx_float4 *radiance) {
    if (!r)
        return fail(PHX_ERR_INVALID_ARG, "null renderer");
    PHX_TRY(r->sync_all(), "synchronize");
    const ImageBuffers &i = r->buffers.image_buf;
    size_t np = i.n_pixels;
    hipStream_t s = r->stream;
    if (flux)
        PHX_TRY(hipMemcpyAsync(flux, i.d_flux.p, np * sizeof(phx_float4), hipMemcpyDeviceToHost, s), "read state");
    if (photon_weight)
        PHX_TRY(hipMemcpyAsync(photon_weight, i.d_photon_weight.p, np * sizeof(float), hipMemcpyDeviceToHost, s),
                "read state");
    if (radiance)
        PHX_TRY(hipMemcpyAsync(radiance, i.d_radiance.p, np * sizeof(phx_float4), hipMemcpyDeviceToHost, s),
                "read state");
    return hip_status(hipStreamSynchronize(s), "read state");
}

int phx_renderer_read_hitpoints(PhxRenderer *r, phx_float4 *hp6) {
    if (!r || !hp6)
        return fail(PHX_ERR_INVALID_ARG, "null argument");
    PHX_TRY(r->sync_all(), "synchronize");
    const ImageBuffers &i = r->buffers.image_buf;
    size_t np = i.n_pixels;
    std::vector<phx_float4> tmp(np);
    const DevBuf *b[6] = {&i.d_hp_pos,        &i.d_hp_normal, &i.d_hp_throughput,
                          &i.d_hp_base_color, &i.d_hp_direct, &i.d_hp_emission};
    for (int k = 0; k < 6; k++) {
        PHX_TRY(hipMemcpyAsync(tmp.data(), b[k]->p, np * sizeof(phx_float4), hipMemcpyDeviceToHost, r->stream),
                "read hitpoints");
        PHX_TRY(hipStreamSynchronize(r->stream), "read hitpoints");
        for (size_t p = 0; p < np; p++)
            hp6[p * 6 + k] = tmp[p];
    }
    return PHX_OK;
}

int phx_renderer_read_photon_map(PhxRenderer *r, phx_float4 *pos, uint32_t *power, uint32_t *dir,
                                 uint32_t *cell_start, uint32_t *cell_end) {
    if (!r)
        return fail(PHX_ERR_INVALID_ARG, "null renderer");
    PHX_TRY(r->sync_all(), "synchronize");
    (void)r->collect();
    const PhotonBuffers &p = r->buffers.photon_buf;
    size_t n = p.n_photons, nb = r->bucket_count;
    hipStream_t s = r->stream;
    if (pos)
        PHX_TRY(hipMemcpyAsync(pos, p.d_photon_pos.p, n * sizeof(phx_float4), hipMemcpyDeviceToHost, s), "read map");
    if (power)
        PHX_TRY(hipMemcpyAsync(power, p.d_photon_power.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "read map");
    if (dir)
        PHX_TRY(hipMemcpyAsync(dir, p.d_photon_dir.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "read map");
    if (cell_start)
        PHX_TRY(hipMemcpyAsync(cell_start, p.d_cell_start.p, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                "read map");
    if (cell_end)
        PHX_TRY(hipMemcpyAsync(cell_end, p.d_cell_end.p, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                "read map");
    return hip_status(hipStreamSynchronize(s), "read map");
}

int phx_renderer_device_buffers(PhxRenderer *r, void **ptrs, uint32_t n_ptrs) {
    if (!r || !ptrs)
        return fail(PHX_ERR_INVALID_ARG, "null argument");
    PHX_TRY(r->sync_all(), "synchronize");
    const ImageBuffers &i = r->buffers.image_buf;
    const PhotonBuffers &p = r->buffers.photon_buf;
    void *all[] = {i.d_hp_pos.p,       i.d_hp_normal.p,     i.d_hp_throughput.p, i.d_hp_base_color.p,
                   i.d_hp_direct.p,    i.d_hp_emission.p,   i.d_flux.p,          i.d_photon_weight.p,
                   i.d_radiance.p,     p.d_photon_pos.p,    p.d_photon_power.p,  p.d_photon_dir.p,
                   p.d_cell_start.p,   p.d_cell_end.p,      r->d_deposit_rec.p,  nullptr,
                   nullptr,            p.d_photon_hashes.p, nullptr,             (void *)r->stream};
    uint32_t total = (uint32_t)(sizeof(all) / sizeof(all[0]));
    for (uint32_t k = 0; k < n_ptrs && k < total; k++)
        ptrs[k] = all[k];
    return PHX_OK;
}

/* Checkpoint of the progressive state: header, flux, photon_weight,
 * radiance.  Resuming from it continues the same sequence of round seeds. */
struct PhxCheckpointHeader {
    uint32_t magic, version, width, height;
    uint32_t rounds, rng_state, photons_per_round, reserved;
    uint64_t total_photons_emitted;
};
static const uint32_t PHX_CKPT_MAGIC = 0x43584850u; /* "PHXC" */

size_t phx_renderer_checkpoint_bytes(PhxRenderer *r) {
    if (!r)
        return 0;
    size_t np = r->buffers.image_buf.n_pixels;
    return sizeof(PhxCheckpointHeader) + np * (2 * sizeof(phx_float4) + sizeof(float));
}

int phx_renderer_save_checkpoint(PhxRenderer *r, void *buf, size_t bytes) {
    if (!r || !buf)
        return fail(PHX_ERR_INVALID_ARG, "null argument");
    if (bytes < phx_renderer_checkpoint_bytes(r))
        return fail(PHX_ERR_WORKSPACE, "phx_renderer_save_checkpoint: buffer too small");
    PHX_TRY(r->sync_all(), "checkpoint");
    const ImageBuffers &i = r->buffers.image_buf;
    size_t np = i.n_pixels;
    PhxCheckpointHeader h{};
    h.magic = PHX_CKPT_MAGIC;
    h.version = 1;
    h.width = i.image_width;
    h.height = i.image_height;
    h.rounds = r->rounds;
    h.photons_per_round = r->args.photons_per_round;
    h.total_photons_emitted = r->total_photons_emitted;
    PHX_TRY(hipMemcpy(&h.rng_state, r->d_rng.p, sizeof(uint32_t), hipMemcpyDeviceToHost), "checkpoint");
    char *p = (char *)buf;
    std::memcpy(p, &h, sizeof(h));
    p += sizeof(h);
    PHX_TRY(hipMemcpy(p, i.d_flux.p, np * sizeof(phx_float4), hipMemcpyDeviceToHost), "checkpoint");
    p += np * sizeof(phx_float4);
    PHX_TRY(hipMemcpy(p, i.d_photon_weight.p, np * sizeof(float), hipMemcpyDeviceToHost), "checkpoint");
    p += np * sizeof(float);
    PHX_TRY(hipMemcpy(p, i.d_radiance.p, np * sizeof(phx_float4), hipMemcpyDeviceToHost), "checkpoint");
    return PHX_OK;
}

int phx_renderer_load_checkpoint(PhxRenderer *r, const void *buf, size_t bytes) {
    if (!r || !buf)
        return fail(PHX_ERR_INVALID_ARG, "null argument");
    if (bytes < sizeof(PhxCheckpointHeader))
        return fail(PHX_ERR_INVALID_ARG, "phx_renderer_load_checkpoint: truncated checkpoint");
    PhxCheckpointHeader h;
    std::memcpy(&h, buf, sizeof(h));
    const ImageBuffers &i = r->buffers.image_buf;
    size_t np = i.n_pixels;
    if (h.magic != PHX_CKPT_MAGIC || h.version != 1)
        return fail(PHX_ERR_INVALID_ARG, "phx_renderer_load_checkpoint: not a checkpoint");
    if (h.width != i.image_width || h.height != i.image_height || bytes < phx_renderer_checkpoint_bytes(r))
        return fail(PHX_ERR_INVALID_ARG, "phx_renderer_load_checkpoint: image size differs");
    PHX_TRY(r->sync_all(), "checkpoint");
    const char *p = (const char *)buf + sizeof(h);
    PHX_TRY(hipMemcpy(i.d_flux.p, p, np * sizeof(phx_float4), hipMemcpyHostToDevice), "checkpoint");
    p += np * sizeof(phx_float4);
    PHX_TRY(hipMemcpy(i.d_photon_weight.p, p, np * sizeof(float), hipMemcpyHostToDevice), "checkpoint");
    p += np * sizeof(float);
    PHX_TRY(hipMemcpy(i.d_radiance.p, p, np * sizeof(phx_float4), hipMemcpyHostToDevice), "checkpoint");
    PHX_TRY(hipMemcpy(r->d_rng.p, &h.rng_state, sizeof(uint32_t), hipMemcpyHostToDevice), "checkpoint");
    r->rounds = h.rounds;
    r->total_photons_emitted = h.total_photons_emitted;
    return PHX_OK;
}

int phx_renderer_estimate(PhxRenderer *r, float *rgb) {
    if (!r || !rgb)
        return fail(PHX_ERR_INVALID_ARG, "null argument");
    size_t np = r->buffers.image_buf.n_pixels;
    std::vector<phx_float4> flux(np), rad(np);
    int rc = phx_renderer_read_state(r, flux.data(), nullptr, rad.data());
    if (rc != PHX_OK)
        return rc;
    /* output.cpp:52-83 */
    const float PI_F = 3.141592654f;
    float inv_ph = r->total_photons_emitted > 0 ? 1.0f / (float)r->total_photons_emitted : 0.0f;
    float inv_rd = r->rounds > 0 ? 1.0f / (float)r->rounds : 0.0f;
    for (size_t p = 0; p < np; p++) {
        float area = PI_F * flux[p].w;
        float ir = 0.0f, ig = 0.0f, ib = 0.0f;
        if (area > 0.0f) {
            float pa = inv_ph / area;
            ir = flux[p].x * pa;
            ig = flux[p].y * pa;
            ib = flux[p].z * pa;
        }
        rgb[3 * p + 0] = rad[p].x * inv_rd + ir;
        rgb[3 * p + 1] = rad[p].y * inv_rd + ig;
        rgb[3 * p + 2] = rad[p].z * inv_rd + ib;
    }
    return PHX_OK;
}

} /* extern "C" */
