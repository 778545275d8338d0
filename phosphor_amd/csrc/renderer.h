/*
 * renderer.h — the one-scene pass driver's state (src/renderer/renderer.hpp,
 * include/buffers/photon.hpp, include/buffers/image.hpp,
 * include/buffers/rendering.hpp), shared by renderer.cpp and the adjoint
 * driver, which replays the driver's stages.
 */
#ifndef PHX_RENDERER_H
#define PHX_RENDERER_H

#include "internal.h"
#include "scene_buffers.h"

namespace phx {

struct PhotonBuffers { /* include/buffers/photon.hpp (sorted photon map) */
    DevBuf d_photon_pos, d_photon_power, d_photon_dir, d_photon_hashes;
    DevBuf d_photon_dirpow; /* each record's decoded direction, packed power in w */
    DevBuf d_cell_start, d_cell_end; /* cell counts are not kept: ranges come from the sorted keys */
    DevBuf d_oct_start;              /* 8 octant starts per cell (records sorted by cell, then octant) */
    uint32_t n_photons = 0;
};

struct ImageBuffers { /* include/buffers/image.hpp */
    DevBuf d_hp_pos, d_hp_normal, d_hp_throughput, d_hp_base_color, d_hp_direct, d_hp_emission;
    DevBuf d_flux, d_photon_weight, d_radiance;
    PhxCameraParams camera{};
    uint32_t image_width = 0, image_height = 0, n_pixels = 0;
};

struct RenderingBuffers { /* include/buffers/rendering.hpp */
    SceneBuffers scene; /* geometry_buf + shading_buf */
    PhotonBuffers photon_buf;
    ImageBuffers image_buf;

    SceneArgs scene_args() const { return scene.args(); }
};

} // namespace phx

/* the C ABI's opaque renderer (kept outside the namespace: include/phosphor_hip.h
   declares it at global scope) */
struct PhxRenderer {
    typedef phx::DevBuf DevBuf;
    PhxRenderArgs args{};
    phx::LaunchShape shape; /* from args: block sizes and gather tile */
    hipStream_t stream = nullptr;
    /* the emission and the photon map run on `side` while the camera trace
       runs on `stream` (the two are independent until the gather) */
    hipStream_t side = nullptr;
    /* pass timing: ev[0] start, ev[1] trace done, ev[2] emission done,
       ev[3] photon map done, ev[4] end; ev_trace / ev_emit / ev_gather the
       starts of the trace, the emission and the gather */
    hipEvent_t ev[5] = {};
    hipEvent_t ev_emit = nullptr, ev_gather = nullptr, ev_trace = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr; /* no timing */
    hipEvent_t ev_info = nullptr;
    /* Pipelined passes (without PHX_PASS_GRAPH or PHX_PASS_SERIAL): a
       pass's gather runs on `gstream` while the next pass's trace and
       emission run.  The hitpoints and the sorted photon map alternate
       between two sets: the one in `buffers` (the last pass's, which readers
       see) and `alt_hp` / `alt_map` (the one the next pass writes), each with
       the end of the last gather that read it (ev_read, [0] for the set in
       `buffers`).  Pipelined passes also alternate between two round-seed
       slots (seed_slot); a pass's round seed waits for the emission of the
       pass before last (ev_emit_done[slot]), the last one that read that
       slot. */
    hipStream_t gstream = nullptr;
    hipEvent_t ev_trace_done = nullptr, ev_emit_done[2] = {}, ev_read[2] = {};
    /* start and end of each pass's gather on the stream it runs on, for the
       last GATHER_RING eager passes (phx_renderer_gather_ms) */
    static constexpr uint32_t GATHER_RING = 64;
    hipEvent_t ev_gring[GATHER_RING][2] = {};
    uint32_t gring_count = 0;
    bool read_pending[2] = {false, false};
    bool emit_pending[2] = {false, false};
    uint32_t seed_slot = 0;
    bool alt_ready = false;
    DevBuf alt_hp[6];  /* pos, normal, throughput, base colour, direct, emission */
    DevBuf alt_map[7]; /* pos, power, dir, dirpow, cell_start, cell_end, oct_start */
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint32_t *h_info = nullptr; /* pinned: last deposit count, passes without deposits */
    phx::RenderingBuffers buffers;
    PhxPhotonHashData photon_hash_info{};
    uint32_t bucket_count = 0;
    uint32_t photon_buf_size = 0;
    /* Renderer's own unsorted photon buffers (renderer.hpp:48-51), kept as
       32-byte records (position; packed power, octahedral direction, 0, 0):
       the sort's last pass reads each deposit with one access */
    DevBuf d_deposit_rec, d_batch_size;
    DevBuf d_emit_ws, d_scan_ws, d_sort_ws, d_trace_ws;
    DevBuf d_rng;  /* [0] PCG state of the round seeds, [1] and [2] the round-seed slots */
    DevBuf d_info; /* [0] last deposit count, [1] passes without deposits */
    uint32_t shadow_any_hit = 0;
    uint32_t emit_chunk = 1; /* photons per emission launch sequence */
    uint32_t rounds = 0;
    uint64_t total_photons_emitted = 0;
    uint32_t zero_reported = 0;
    bool timed = false;
    bool info_pending = false;
    /* set by the adjoint driver: the gather records each pixel's radius
       ratio and r² of the pass, and the sort records its permutation */
    float *rec_ratio = nullptr, *rec_r2 = nullptr;
    uint32_t *sort_perm = nullptr;

    ~PhxRenderer() {
        (void)sync_all();
        if (graph_exec)
            (void)hipGraphExecDestroy(graph_exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        for (hipEvent_t &e : ev)
            if (e)
                (void)hipEventDestroy(e);
        for (auto &pair : ev_gring)
            for (hipEvent_t e : pair)
                if (e)
                    (void)hipEventDestroy(e);
        for (hipEvent_t e : {ev_emit, ev_gather, ev_trace, ev_fork, ev_join, ev_info, ev_trace_done, ev_emit_done[0],
                             ev_emit_done[1], ev_read[0],
                             ev_read[1]})
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

    uint32_t *seed_ptr() const { return d_rng.as<uint32_t>() + 1 + seed_slot; }
    int set_scene(const PhxSceneDesc &scene);
    int run_pass();
    int enqueue_pass(bool events, bool pipe);
    int alloc_alt();
    void hp_bufs(int set, DevBuf *out[6]);
    void map_bufs(int set, DevBuf *out[7]);
    void swap_sets();
    hipError_t sync_all();
    int capture_pass();
    int collect();
    int check_passes();
    /* set 0: the buffers in `buffers`; set 1: alt_hp / alt_map */
    int trace_rays(int set = 0);
    int emit_photons() { return emit_photons(stream); }
    int emit_photons(hipStream_t s);
    int hash_and_sort_photons() { return hash_and_sort_photons(stream, 0); }
    int hash_and_sort_photons(hipStream_t s, int set);
    int gather_flux() { return gather_flux(stream, 0); }
    int gather_flux(hipStream_t s, int set);
};


#endif /* PHX_RENDERER_H */
