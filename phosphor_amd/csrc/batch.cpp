/*
 * batch.cpp — SPPM passes over many scenes at once.
 *
 * Each scene keeps the arguments, PCG seed schedule (renderer.cpp:54,65) and
 * results of a one-scene Renderer; the batch concatenates their per-pixel
 * state, photons and grid buckets into batch-wide arrays so one pass is a
 * fixed sequence of kernel launches over all scenes (kernels_batch.hip).
 * Scenes are grouped for the trace and emission stages so the wavefront
 * workspaces stay bounded.  The pass never waits for the host; with
 * PHX_PASS_GRAPH it is captured once into a hipGraph and replayed.
 */
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "internal.h"
#include "scene_buffers.h"

using namespace phx;

#define PHX_TRY(expr, what)                                                                                            \
    do {                                                                                                               \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess)                                                                                          \
            return hip_status(e_, what);                                                                               \
    } while (0)

struct PhxBatch {
    std::vector<PhxRenderArgs> args;
    PhxBatchOptions opt{};
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr; /* emission and photon map, beside the camera trace */
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_info = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    /* Pipelined passes (without PHX_PASS_GRAPH or PHX_PASS_SERIAL), as for
       one scene: the gather on `gstream` beside the next pass's trace and
       emission; the hitpoints and the photon map alternate between the set
       in d_hp / d_map_* (the last pass's; `plan` points at it) and the alt
       set (`plan_alt`), swapped after each pass; ev_read[k] ends the last
       gather that read set k, ev_emit_done the last emission (which reads
       the round seeds). */
    hipStream_t gstream = nullptr;
    hipEvent_t ev_trace_done = nullptr, ev_emit_done = nullptr, ev_read[2] = {};
    bool read_pending[2] = {false, false};
    bool emit_pending = false;
    bool alt_ready = false;
    DevBuf alt_hp[6], alt_map[7]; /* map: pos, power, dir, dirpow, cell_start, cell_end, oct_start */
    BatchPlan plan_alt;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<std::unique_ptr<SceneBuffers>> scenes;
    std::vector<BatchScene> table;
    DevBuf d_table, d_chunk_scene, d_tile_scene, d_ph_base, d_dep_base, d_rng, d_round_seed, d_counts, d_zero;
    DevBuf d_tickets, d_counter;
    DevBuf d_hp[6], d_flux, d_weight, d_radiance;
    DevBuf d_dep_pos, d_dep_power, d_dep_dir, d_map_pos, d_map_power, d_map_dir, d_hashes, d_map_dirpow;
    DevBuf d_cell_counts, d_cell_start, d_cell_end, d_oct_start;
    DevBuf d_trace_ws, d_emit_ws, d_scan_ws, d_sort_ws;
    BatchPlan plan;
    uint32_t *h_info = nullptr; /* pinned: per-scene deposit counts, then per-scene passes without deposits */
    std::vector<uint32_t> zero_reported;
    uint32_t rounds = 0;
    bool info_pending = false;

    ~PhxBatch() {
        (void)sync_all();
        if (graph_exec)
            (void)hipGraphExecDestroy(graph_exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        for (hipEvent_t e : {ev0, ev1, ev_info, ev_fork, ev_join, ev_trace_done, ev_emit_done, ev_read[0], ev_read[1]})
            if (e)
                (void)hipEventDestroy(e);
        if (h_info)
            (void)hipHostFree(h_info);
        if (gstream)
            (void)hipStreamDestroy(gstream);
        if (side)
            (void)hipStreamDestroy(side);
        if (stream)
            (void)hipStreamDestroy(stream);
    }

    uint32_t n() const { return (uint32_t)args.size(); }
    int build(const PhxSceneDesc *descs);
    int collect();
    int run_pass();
    int alloc_alt();
    void map_bufs(DevBuf *out[7]) {
        DevBuf *m[7] = {&d_map_pos, &d_map_power, &d_map_dir, &d_map_dirpow, &d_cell_start, &d_cell_end, &d_oct_start};
        for (int k = 0; k < 7; k++)
            out[k] = m[k];
    }
    /* waits for every stream of the batch */
    hipError_t sync_all() {
        for (hipStream_t q : {stream, side, gstream}) {
            hipError_t e = q ? hipStreamSynchronize(q) : hipSuccess;
            if (e != hipSuccess)
                return e;
        }
        read_pending[0] = read_pending[1] = false;
        emit_pending = false;
        return hipSuccess;
    }
};

/* the second set of hitpoint and photon-map buffers (cleared), and the plan
 * that writes it */
int PhxBatch::alloc_alt() {
    if (alt_ready)
        return PHX_OK;
    DevBuf *m[7];
    map_bufs(m);
    for (int k = 0; k < 6; k++) {
        PHX_TRY(alt_hp[k].alloc(d_hp[k].bytes), "batch second hitpoint set");
        PHX_TRY(hipMemsetAsync(alt_hp[k].p, 0, alt_hp[k].bytes, stream), "batch second hitpoint set");
    }
    for (int k = 0; k < 7; k++) {
        PHX_TRY(alt_map[k].alloc(m[k]->bytes), "batch second photon map");
        PHX_TRY(hipMemsetAsync(alt_map[k].p, 0, alt_map[k].bytes, stream), "batch second photon map");
    }
    plan_alt = plan;
    for (int k = 0; k < 6; k++)
        plan_alt.hp[k] = alt_hp[k].as<phx_float4>();
    plan_alt.map_pos = alt_map[0].as<phx_float4>();
    plan_alt.map_power = alt_map[1].as<uint32_t>();
    plan_alt.map_dir = alt_map[2].as<uint32_t>();
    plan_alt.map_dirpow = alt_map[3].as<phx_float4>();
    plan_alt.cell_start = alt_map[4].as<uint32_t>();
    plan_alt.cell_end = alt_map[5].as<uint32_t>();
    plan_alt.oct_start = alt_map[6].as<uint32_t>();
    alt_ready = true;
    return PHX_OK;
}

/* one pass: serial (graph replay, PHX_PASS_SERIAL) or pipelined.  Ordering
 * of a pipelined pass: the round seeds wait for the previous emission; the
 * trace and the photon map, which overwrite the set the gather before last
 * read, wait for that gather (on `stream`, before the seeds; `side` follows
 * `stream` from the seeds on); the gather waits for the trace and the photon
 * map, and gathers run in order on `gstream`.  A serial pass first waits for
 * every pending gather. */
int PhxBatch::run_pass() {
    bool use_graph = (opt.flags & PHX_PASS_GRAPH) != 0;
    bool pipe = !use_graph && !(opt.flags & PHX_PASS_SERIAL);
    BatchStreams B;
    B.s = stream;
    B.side = side;
    B.ev_fork = ev_fork;
    B.ev_join = ev_join;
    if (use_graph && !graph_exec) {
        PHX_TRY(sync_all(), "batch graph capture");
        PHX_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "batch graph capture");
        hipError_t le = launch_batch_pass(plan, B);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(stream, &g);
        if (le != hipSuccess || ce != hipSuccess) {
            if (g)
                (void)hipGraphDestroy(g);
            return hip_status(le != hipSuccess ? le : ce, "batch graph capture");
        }
        graph = g;
        PHX_TRY(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0), "batch graph instantiate");
    }
    if (pipe) {
        int rc = alloc_alt();
        if (rc != PHX_OK)
            return rc;
    }
    if (emit_pending) {
        PHX_TRY(hipStreamWaitEvent(stream, ev_emit_done, 0), "event");
        emit_pending = false;
    }
    for (int k = pipe ? 1 : 0; k < 2; k++)
        if (read_pending[k]) {
            PHX_TRY(hipStreamWaitEvent(stream, ev_read[k], 0), "event");
            read_pending[k] = false;
        }
    PHX_TRY(hipEventRecord(ev0, stream), "event");
    size_t S = n();
    if (use_graph) {
        PHX_TRY(hipGraphLaunch(graph_exec, stream), "batch graph launch");
    } else if (!pipe) {
        PHX_TRY(launch_batch_pass(plan, B), "batch pass");
    } else {
        B.gather = gstream;
        B.ev_trace = ev_trace_done;
        B.ev_emit = ev_emit_done;
        B.h_info = h_info;
        PHX_TRY(launch_batch_pass(plan_alt, B), "batch pass");
        PHX_TRY(hipEventRecord(ev_info, side), "event");
        PHX_TRY(hipEventRecord(ev1, gstream), "event");
        PHX_TRY(hipEventRecord(ev_read[1], gstream), "event");
        read_pending[1] = true;
        emit_pending = true;
        /* the set just written becomes the one in d_hp / d_map_* */
        DevBuf *m[7];
        map_bufs(m);
        for (int k = 0; k < 6; k++)
            d_hp[k].swap(alt_hp[k]);
        for (int k = 0; k < 7; k++)
            m[k]->swap(alt_map[k]);
        std::swap(plan, plan_alt);
        std::swap(ev_read[0], ev_read[1]);
        std::swap(read_pending[0], read_pending[1]);
    }
    if (!pipe) {
        PHX_TRY(hipEventRecord(ev1, stream), "event");
        PHX_TRY(hipMemcpyAsync(h_info, d_counts.p, S * 4, hipMemcpyDeviceToHost, stream), "batch info");
        PHX_TRY(hipMemcpyAsync(h_info + S, d_zero.p, S * 4, hipMemcpyDeviceToHost, stream), "batch info");
        PHX_TRY(hipEventRecord(ev_info, stream), "event");
    }
    info_pending = true;
    rounds++;
    return PHX_OK;
}

/* groups of whole scenes with at most `budget` items each (a scene larger
 * than the budget forms its own group) */
static std::vector<BatchGroup> make_groups(const std::vector<uint32_t> &items, const std::vector<uint32_t> &first,
                                           const std::vector<uint32_t> &chunk_first,
                                           const std::vector<uint32_t> &chunks, uint64_t budget) {
    std::vector<BatchGroup> out;
    uint32_t S = (uint32_t)items.size();
    uint32_t s = 0;
    while (s < S) {
        BatchGroup g{s, s, first[s], 0u, chunk_first.empty() ? 0u : chunk_first[s], 0u};
        uint64_t sum = 0;
        while (g.s1 < S && (g.s1 == g.s0 || sum + items[g.s1] <= budget)) {
            sum += items[g.s1];
            if (!chunks.empty())
                g.chunks += chunks[g.s1];
            g.s1++;
        }
        g.count = (uint32_t)sum;
        out.push_back(g);
        s = g.s1;
    }
    return out;
}

int PhxBatch::build(const PhxSceneDesc *descs) {
    hipStream_t s = stream;
    uint32_t S = n();
    table.resize(S);
    std::vector<uint32_t> chunk_scene, tile_scene, ph_base(S + 1), rng(S);
    std::vector<uint32_t> px_items(S), px_first(S), chunk_first(S), chunk_count(S), ph_items(S), ph_first(S);
    uint64_t px = 0, chunks = 0, tiles = 0, photons = 0, buckets = 0, deps = 0;
    uint32_t max_ds = 1, max_b = 1;
    for (uint32_t i = 0; i < S; i++) {
        const PhxRenderArgs &a = args[i];
        scenes.emplace_back(new SceneBuffers());
        int rc = upload_scene(descs[i], s, ((opt.flags | a.flags) & PHX_REFERENCE_BVH) != 0, a.bvh_leaf_size,
                              *scenes[i]);
        if (rc != PHX_OK)
            return rc;
        BatchScene &B = table[i];
        std::memset(&B, 0, sizeof(B));
        B.sc = scenes[i]->args();
        B.cam = *descs[i].camera;
        phx_build_photon_hash_info(&scenes[i]->root, a.grid_res, &B.hash);
        B.width = a.width;
        B.height = a.height;
        B.n_pixels = a.width * a.height;
        B.px_base = (uint32_t)px;
        B.chunk_base = (uint32_t)chunks;
        B.tile_base = (uint32_t)tiles;
        B.tiles_x = (a.width + 7) / 8;
        B.tiles_y = (a.height + 7) / 8;
        B.photons = a.photons_per_round;
        B.ph_base = (uint32_t)photons;
        B.max_bounces = a.max_photon_bounces;
        B.direct_samples = a.direct_samples;
        B.bucket_base = (uint32_t)buckets;
        B.n_buckets = a.grid_res * a.grid_res * a.grid_res + 1;
        B.any_hit = scenes[i]->shadow_any_hit;
        B.alpha = a.sppm_alpha;
        uint32_t nchunks = (B.n_pixels + 127) / 128;
        px_items[i] = B.n_pixels;
        px_first[i] = B.px_base;
        chunk_first[i] = B.chunk_base;
        chunk_count[i] = nchunks;
        ph_items[i] = B.photons;
        ph_first[i] = B.ph_base;
        ph_base[i] = B.ph_base;
        chunk_scene.insert(chunk_scene.end(), nchunks, i);
        tile_scene.insert(tile_scene.end(), (size_t)B.tiles_x * B.tiles_y, i);
        px += B.n_pixels;
        chunks += nchunks;
        tiles += (uint64_t)B.tiles_x * B.tiles_y;
        photons += B.photons;
        buckets += B.n_buckets;
        deps += (uint64_t)B.photons * B.max_bounces;
        max_ds = std::max(max_ds, a.direct_samples);
        max_b = std::max(max_b, a.max_photon_bounces);
        rng[i] = host_pcg_seed(a.seed);
    }
    /* octant sort keys are 8 per bucket */
    if (px > 0xFFFFFFFFull || photons > 0xFFFFFFFFull || buckets * 8u > 0xFFFFFFFFull || deps > 0xFFFFFFFFull)
        return fail(PHX_ERR_TOO_LARGE, "phx_batch_create: batch totals exceed 32-bit indices");
    ph_base[S] = (uint32_t)photons;

    /* batch64 workload (64 scenes, 34.7 M pixels, 33.5 M photons per pass):
       1 Mi / 2 Mi 245 ms per pass, 2 Mi / 8 Mi 257 ms, 8 Mi / 16 Mi 270 ms */
    uint64_t trace_budget = opt.trace_group_pixels ? opt.trace_group_pixels : (1u << 20);
    uint64_t emit_budget = opt.emit_group_photons ? opt.emit_group_photons : (1u << 21);
    plan.trace_groups = make_groups(px_items, px_first, chunk_first, chunk_count, trace_budget);
    plan.emit_groups = make_groups(ph_items, ph_first, {}, {}, emit_budget);
    uint32_t max_tp = 0, max_ep = 0;
    for (const BatchGroup &g : plan.trace_groups)
        max_tp = std::max(max_tp, g.count);
    for (const BatchGroup &g : plan.emit_groups)
        max_ep = std::max(max_ep, g.count);
    /* a trace group's shadow queue holds 5 NEE events per pixel, each with up to
       max_ds light samples, indexed by a u32 slot counter */
    if ((uint64_t)max_tp * 5u * max_ds > 0xFFFFFFFFull)
        return fail(PHX_ERR_TOO_LARGE, "phx_batch_create: a trace group's shadow queue exceeds 2^32 slots "
                                       "(lower trace_group_pixels)");

    PHX_TRY(copy_to_dev_buf(d_table, table.data(), S, sizeof(BatchScene), s), "batch table");
    PHX_TRY(copy_to_dev_buf(d_chunk_scene, chunk_scene.data(), chunk_scene.size(), 4, s), "batch table");
    PHX_TRY(copy_to_dev_buf(d_tile_scene, tile_scene.data(), tile_scene.size(), 4, s), "batch table");
    PHX_TRY(copy_to_dev_buf(d_ph_base, ph_base.data(), S + 1, 4, s), "batch table");
    PHX_TRY(copy_to_dev_buf(d_dep_base, nullptr, S + 1, 4, s), "batch table");
    PHX_TRY(copy_to_dev_buf(d_rng, rng.data(), S, 4, s), "batch seeds");
    PHX_TRY(copy_to_dev_buf(d_round_seed, nullptr, S, 4, s), "batch seeds");
    PHX_TRY(copy_to_dev_buf(d_counts, nullptr, S, 4, s), "batch counts");
    PHX_TRY(copy_to_dev_buf(d_zero, nullptr, S, 4, s), "batch counts");
    PHX_TRY(copy_to_dev_buf(d_tickets, nullptr, plan.trace_groups.size(), 4, s), "batch tickets");
    PHX_TRY(copy_to_dev_buf(d_counter, nullptr, 1, 4, s), "batch counter");
    for (DevBuf &b : d_hp)
        PHX_TRY(copy_to_dev_buf(b, nullptr, px, sizeof(phx_float4), s), "batch hitpoints");
    PHX_TRY(copy_to_dev_buf(d_radiance, nullptr, px, sizeof(phx_float4), s), "batch state");
    PHX_TRY(copy_to_dev_buf(d_weight, nullptr, px, sizeof(float), s), "batch state");
    {
        /* flux.w starts at r0² where an initial radius is given, else 0 */
        std::vector<phx_float4> f(px);
        for (uint32_t i = 0; i < S; i++) {
            float r2 = args[i].initial_radius > 0.0f ? args[i].initial_radius * args[i].initial_radius : 0.0f;
            std::fill(f.begin() + table[i].px_base, f.begin() + table[i].px_base + table[i].n_pixels,
                      phx_float4{0.0f, 0.0f, 0.0f, r2});
        }
        PHX_TRY(copy_to_dev_buf(d_flux, f.data(), px, sizeof(phx_float4), s), "batch state");
        PHX_TRY(hipStreamSynchronize(s), "batch state");
    }
    PHX_TRY(d_dep_pos.alloc(deps * sizeof(phx_float4)), "batch photons");
    PHX_TRY(d_dep_power.alloc(deps * 4), "batch photons");
    PHX_TRY(d_dep_dir.alloc(deps * 4), "batch photons");
    PHX_TRY(d_map_pos.alloc(deps * sizeof(phx_float4)), "batch photon map");
    PHX_TRY(d_map_power.alloc(deps * 4), "batch photon map");
    PHX_TRY(d_map_dir.alloc(deps * 4), "batch photon map");
    PHX_TRY(d_hashes.alloc(deps * 4), "batch photon map");
    PHX_TRY(d_map_dirpow.alloc(deps * sizeof(phx_float4)), "batch photon map");
    PHX_TRY(d_cell_counts.alloc(buckets * 4), "batch cells");
    PHX_TRY(d_cell_start.alloc(buckets * 4), "batch cells");
    PHX_TRY(d_cell_end.alloc(buckets * 4), "batch cells");
    PHX_TRY(d_oct_start.alloc(buckets * 32), "batch cells");
    PHX_TRY(d_trace_ws.alloc(batch_trace_workspace_bytes(max_tp, max_ds)), "batch trace workspace");
    PHX_TRY(d_emit_ws.alloc(batch_emit_workspace_bytes(max_ep, max_b)), "batch emit workspace");
    PHX_TRY(d_scan_ws.alloc(scan_workspace_bytes((uint32_t)buckets)), "batch scan workspace");
    PHX_TRY(d_sort_ws.alloc(sort_workspace_bytes((uint32_t)deps, (uint32_t)buckets)), "batch sort workspace");

    BatchPlan &P = plan;
    P.n_scenes = S;
    P.scenes = d_table.as<BatchScene>();
    P.chunk_scene = d_chunk_scene.as<uint32_t>();
    P.tile_scene = d_tile_scene.as<uint32_t>();
    P.ph_base = d_ph_base.as<uint32_t>();
    P.dep_base = d_dep_base.as<uint32_t>();
    P.rng = d_rng.as<uint32_t>();
    P.round_seed = d_round_seed.as<uint32_t>();
    P.counts = d_counts.as<uint32_t>();
    P.zero_passes = d_zero.as<uint32_t>();
    P.tickets = d_tickets.as<uint32_t>();
    P.dep_counter = d_counter.as<uint32_t>();
    for (int k = 0; k < 6; k++)
        P.hp[k] = d_hp[k].as<phx_float4>();
    P.flux = d_flux.as<phx_float4>();
    P.radiance = d_radiance.as<phx_float4>();
    P.photon_weight = d_weight.as<float>();
    P.dep_pos = d_dep_pos.as<phx_float4>();
    P.dep_power = d_dep_power.as<uint32_t>();
    P.dep_dir = d_dep_dir.as<uint32_t>();
    P.map_pos = d_map_pos.as<phx_float4>();
    P.map_power = d_map_power.as<uint32_t>();
    P.map_dir = d_map_dir.as<uint32_t>();
    P.hashes = d_hashes.as<uint32_t>();
    P.map_dirpow = d_map_dirpow.as<phx_float4>();
    P.cell_counts = d_cell_counts.as<uint32_t>();
    P.cell_start = d_cell_start.as<uint32_t>();
    P.cell_end = d_cell_end.as<uint32_t>();
    P.oct_start = d_oct_start.as<uint32_t>();
    P.dep_cap = (uint32_t)deps;
    P.total_buckets = (uint32_t)buckets;
    P.total_tiles = (uint32_t)tiles;
    P.trace_ws = d_trace_ws.p;
    P.emit_ws = d_emit_ws.p;
    P.scan_ws = d_scan_ws.p;
    P.sort_ws = d_sort_ws.p;
    P.trace_max_pixels = max_tp;
    P.max_direct_samples = max_ds;
    P.emit_max_photons = max_ep;
    P.max_bounces = max_b;
    P.cam_blocks = batch_camera_blocks();
    zero_reported.assign(S, 0u);
    PHX_TRY(hipStreamSynchronize(s), "phx_batch_create");
    return PHX_OK;
}

int PhxBatch::collect() {
    if (!info_pending)
        return PHX_OK;
    PHX_TRY(hipEventSynchronize(ev_info), "batch info");
    info_pending = false;
    return PHX_OK;
}

extern "C" {

int phx_batch_create(const PhxRenderArgs *args, const PhxSceneDesc *scenes, uint32_t n_scenes,
                     const PhxBatchOptions *options, PhxBatch **out) {
    if (!args || !scenes || !out || n_scenes == 0)
        return fail(PHX_ERR_INVALID_ARG, "phx_batch_create: null argument or no scene");
    *out = nullptr;
    for (uint32_t i = 0; i < n_scenes; i++) {
        int rc = check_scene_desc(scenes[i], "phx_batch_create");
        if (rc == PHX_OK)
            rc = check_render_args(args[i], "phx_batch_create");
        if (rc != PHX_OK)
            return fail(rc, "phx_batch_create: scene " + std::to_string(i) + ": " + std::string(phx_last_error()));
    }
    if (options && (options->flags & ~(uint32_t)(PHX_PASS_GRAPH | PHX_REFERENCE_BVH | PHX_PASS_SERIAL)))
        return fail(PHX_ERR_INVALID_ARG, "phx_batch_create: unknown flags");
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
        return fail(PHX_ERR_NO_DEVICE, "phx_batch_create: no HIP device");
    std::unique_ptr<PhxBatch> b(new PhxBatch());
    b->args.assign(args, args + n_scenes);
    if (options)
        b->opt = *options;
    PHX_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "phx_batch_create: stream");
    PHX_TRY(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking), "phx_batch_create: stream");
    PHX_TRY(hipStreamCreateWithFlags(&b->gstream, hipStreamNonBlocking), "phx_batch_create: stream");
    for (hipEvent_t *e : {&b->ev_trace_done, &b->ev_emit_done, &b->ev_read[0], &b->ev_read[1]})
        PHX_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming), "phx_batch_create: event");
    PHX_TRY(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming), "phx_batch_create: event");
    PHX_TRY(hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming), "phx_batch_create: event");
    PHX_TRY(hipEventCreate(&b->ev0), "phx_batch_create: event");
    PHX_TRY(hipEventCreate(&b->ev1), "phx_batch_create: event");
    PHX_TRY(hipEventCreateWithFlags(&b->ev_info, hipEventDisableTiming), "phx_batch_create: event");
    PHX_TRY(hipHostMalloc((void **)&b->h_info, 2ull * n_scenes * sizeof(uint32_t), hipHostMallocDefault),
            "phx_batch_create: pinned info");
    std::memset(b->h_info, 0, 2ull * n_scenes * sizeof(uint32_t));
    int rc = b->build(scenes);
    if (rc != PHX_OK)
        return rc;
    *out = b.release();
    return PHX_OK;
}

void phx_batch_destroy(PhxBatch *b) { delete b; }

int phx_batch_run_pass(PhxBatch *b) {
    if (!b)
        return fail(PHX_ERR_INVALID_ARG, "phx_batch_run_pass: null batch");
    return b->run_pass();
}

int phx_batch_synchronize(PhxBatch *b) {
    if (!b)
        return fail(PHX_ERR_INVALID_ARG, "null batch");
    PHX_TRY(b->sync_all(), "phx_batch_synchronize");
    int rc = b->collect();
    if (rc != PHX_OK)
        return rc;
    uint32_t S = b->n();
    for (uint32_t i = 0; i < S; i++) {
        if (b->h_info[S + i] > b->zero_reported[i]) {
            b->zero_reported[i] = b->h_info[S + i];
            return fail(PHX_ERR_NO_PHOTONS, "no photon hit (scene " + std::to_string(i) + ")");
        }
    }
    return PHX_OK;
}

int phx_batch_last_pass_ms(PhxBatch *b, float *out_ms) {
    if (!b || !out_ms)
        return fail(PHX_ERR_INVALID_ARG, "null argument");
    *out_ms = 0.0f;
    if (b->rounds == 0)
        return PHX_OK;
    PHX_TRY(hipEventSynchronize(b->ev1), "batch timing");
    return hip_status(hipEventElapsedTime(out_ms, b->ev0, b->ev1), "batch timing");
}

uint32_t phx_batch_rounds_done(PhxBatch *b) { return b ? b->rounds : 0u; }

int phx_batch_photon_counts(PhxBatch *b, uint32_t *counts) {
    if (!b || !counts)
        return fail(PHX_ERR_INVALID_ARG, "null argument");
    int rc = b->collect();
    if (rc != PHX_OK)
        return rc;
    std::memcpy(counts, b->h_info, b->n() * sizeof(uint32_t));
    return PHX_OK;
}

static int batch_scene(PhxBatch *b, uint32_t scene, const char *who) {
    if (!b)
        return fail(PHX_ERR_INVALID_ARG, std::string(who) + ": null batch");
    if (scene >= b->n())
        return fail(PHX_ERR_INVALID_ARG, std::string(who) + ": scene index out of range");
    return PHX_OK;
}

int phx_batch_read_state(PhxBatch *b, uint32_t scene, phx_float4 *flux, float *photon_weight,
                         phx_float4 *radiance) {
    int rc = batch_scene(b, scene, "phx_batch_read_state");
    if (rc != PHX_OK)
        return rc;
    PHX_TRY(b->sync_all(), "read");
    const BatchScene &B = b->table[scene];
    hipStream_t s = b->stream;
    size_t np = B.n_pixels, off = B.px_base;
    if (flux)
        PHX_TRY(hipMemcpyAsync(flux, b->d_flux.as<phx_float4>() + off, np * 16, hipMemcpyDeviceToHost, s), "read");
    if (photon_weight)
        PHX_TRY(hipMemcpyAsync(photon_weight, b->d_weight.as<float>() + off, np * 4, hipMemcpyDeviceToHost, s),
                "read");
    if (radiance)
        PHX_TRY(hipMemcpyAsync(radiance, b->d_radiance.as<phx_float4>() + off, np * 16, hipMemcpyDeviceToHost, s),
                "read");
    return hip_status(hipStreamSynchronize(s), "read");
}

int phx_batch_read_hitpoints(PhxBatch *b, uint32_t scene, phx_float4 *hp6) {
    int rc = batch_scene(b, scene, "phx_batch_read_hitpoints");
    if (rc != PHX_OK)
        return rc;
    if (!hp6)
        return fail(PHX_ERR_INVALID_ARG, "null argument");
    PHX_TRY(b->sync_all(), "read");
    const BatchScene &B = b->table[scene];
    size_t np = B.n_pixels;
    std::vector<phx_float4> tmp(np);
    for (int k = 0; k < 6; k++) {
        PHX_TRY(hipMemcpyAsync(tmp.data(), b->d_hp[k].as<phx_float4>() + B.px_base, np * 16, hipMemcpyDeviceToHost,
                               b->stream),
                "read");
        PHX_TRY(hipStreamSynchronize(b->stream), "read");
        for (size_t p = 0; p < np; p++)
            hp6[p * 6 + k] = tmp[p];
    }
    return PHX_OK;
}

/* the scene's photon map as a one-scene pass holds it: records from its
 * first deposit on, cell ranges relative to that deposit */
int phx_batch_read_photon_map(PhxBatch *b, uint32_t scene, phx_float4 *pos, uint32_t *power, uint32_t *dir,
                              uint32_t *cell_start, uint32_t *cell_end) {
    int rc = batch_scene(b, scene, "phx_batch_read_photon_map");
    if (rc != PHX_OK)
        return rc;
    PHX_TRY(b->sync_all(), "read");
    const BatchScene &B = b->table[scene];
    hipStream_t s = b->stream;
    uint32_t base[2];
    PHX_TRY(hipMemcpyAsync(base, b->d_dep_base.as<uint32_t>() + scene, 8, hipMemcpyDeviceToHost, s), "read");
    PHX_TRY(hipStreamSynchronize(s), "read");
    size_t n = base[1] - base[0], nb = B.n_buckets;
    if (pos)
        PHX_TRY(hipMemcpyAsync(pos, b->d_map_pos.as<phx_float4>() + base[0], n * 16, hipMemcpyDeviceToHost, s),
                "read");
    if (power)
        PHX_TRY(hipMemcpyAsync(power, b->d_map_power.as<uint32_t>() + base[0], n * 4, hipMemcpyDeviceToHost, s),
                "read");
    if (dir)
        PHX_TRY(hipMemcpyAsync(dir, b->d_map_dir.as<uint32_t>() + base[0], n * 4, hipMemcpyDeviceToHost, s), "read");
    if (cell_start)
        PHX_TRY(hipMemcpyAsync(cell_start, b->d_cell_start.as<uint32_t>() + B.bucket_base, nb * 4,
                               hipMemcpyDeviceToHost, s),
                "read");
    if (cell_end)
        PHX_TRY(hipMemcpyAsync(cell_end, b->d_cell_end.as<uint32_t>() + B.bucket_base, nb * 4, hipMemcpyDeviceToHost,
                               s),
                "read");
    PHX_TRY(hipStreamSynchronize(s), "read");
    for (size_t h = 0; h < nb; h++) {
        if (cell_start)
            cell_start[h] -= base[0];
        if (cell_end)
            cell_end[h] -= base[0];
    }
    return PHX_OK;
}

int phx_batch_estimate(PhxBatch *b, uint32_t scene, float *rgb) {
    int rc = batch_scene(b, scene, "phx_batch_estimate");
    if (rc != PHX_OK)
        return rc;
    if (!rgb)
        return fail(PHX_ERR_INVALID_ARG, "null argument");
    size_t np = b->table[scene].n_pixels;
    std::vector<phx_float4> flux(np), rad(np);
    if ((rc = phx_batch_read_state(b, scene, flux.data(), nullptr, rad.data())) != PHX_OK)
        return rc;
    /* output.cpp:52-83 */
    const float PI_F = 3.141592654f;
    uint64_t emitted = (uint64_t)b->args[scene].photons_per_round * b->rounds;
    float inv_ph = emitted > 0 ? 1.0f / (float)emitted : 0.0f;
    float inv_rd = b->rounds > 0 ? 1.0f / (float)b->rounds : 0.0f;
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
