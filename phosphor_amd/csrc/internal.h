/*
 * internal.h — launchers and workspace layouts shared by the C ABI, the pass
 * driver and the kernels.  Not part of the public interface.
 */
#ifndef PHX_INTERNAL_H
#define PHX_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/phosphor_hip.h"

namespace phx {

/* Pointers of one scene on the device (kernel argument block). */
struct SceneArgs {
    const PhxTriangleVertices *tv;
    const PhxTriangleNormals *tn;
    const PhxTriangleTangents *tt;
    const PhxTriangleUvs *tuv;
    const uint32_t *tmat;
    const PhxBvhNode *nodes;
    uint32_t n_tris;
    const PhxTriangleVertices *etv;
    const PhxTriangleNormals *etn;
    const PhxTriangleUvs *etuv;
    const uint32_t *etmat;
    const PhxMaterial *materials;
    const PhxUvTransformBlock *uvt;
    const PhxTextureMeta *tex_meta;
    const uint8_t *atlas;
    const PhxLight *lights;
    uint32_t n_lights;
    const float *lum_pref_sum;
    float total_lum;
    phx_float4 scene_center;
    float scene_radius;
    const float *envmap;
    uint32_t env_w, env_h, has_env;
    /* 1: search `pairs`, the library's SAH tree (root `root_ref`) with both
     * children's boxes in each node, nearer child first; its leaves name
     * positions in the tree's triangle order, whose vertices are `tvo` and
     * whose triangle indices are `tri_order`; equal-distance ties are settled
     * by tri_key (by position: the triangle's leaf index in the reference
     * tree).  0: search `nodes`, the reference tree, depth-first. */
    uint32_t near_first;
    uint32_t root_ref;
    const PhxBvhPair *pairs;
    const uint32_t *tri_key;
    const uint32_t *tri_order;
    const PhxTriangleVertices *tvo;
};

/* Launch shapes of the one-scene pass (threads per block of the camera,
 * shadow and emission kernels: 64, 128 or 256; the gather's 64-pixel wave
 * tile: gather_tile_w × 64/gather_tile_w pixels, or 64 consecutive pixels
 * when gather_tile_w is 64).  They change how work is grouped, never a
 * result. */
struct LaunchShape {
    uint32_t camera = 128, shadow = 128, emit = 128, gather_tile_w = 8;
    /* primary camera rays and photon bounces traced by persistent kernels
       with ray replacement ahead of the kernels that shade them (pair tree
       only); 0: traced inside the camera kernel and one fused kernel per
       bounce */
    uint32_t split_trace = 1;
};

/* ---- the SAH pair tree (accel.cpp) ---- */
/* Depth of the deepest interior node of a reference-layout tree (root 0);
 * an error code when a child index is out of range. */
int tree_depth(const PhxBvhNode *nodes, size_t n_nodes, uint32_t &max_depth);
/* Whether rays may be traced on the pair tree instead of this reference
 * tree with the same hits: one triangle per leaf, nodes in preorder, leaf
 * boxes holding their triangle's box grown by 1e-3, interior boxes holding
 * their children's, at most 22 interior levels, coordinates within ±1024.
 * `why` names the first property that fails. */
bool pair_tree_eligible(const PhxBvhNode *nodes, size_t n_nodes, const PhxTriangleVertices *tv, uint32_t n,
                        std::string &why);
struct PairTreeHost {
    std::vector<PhxBvhPair> pairs;
    std::vector<uint32_t> key, order;
    std::vector<PhxTriangleVertices> tvo;
    uint32_t root_ref = 0;
};
/* The SAH tree over the triangles (depth <= 22, leaves of up to leaf_size)
 * as pairs, the triangles in its order and their reference leaf indices. */
bool build_pair_tree(const PhxBvhNode *nodes, size_t n_nodes, const PhxTriangleVertices *tv, uint32_t n,
                     uint32_t leaf_size, PairTreeHost &out);
/* Fills the pair-tree fields of s when phx_accel_attach registered a pair
 * tree for s.nodes and s.tv on the current device. */
void apply_attached_tree(SceneArgs &s);

/* ---- scan (single pass, decoupled look-back) ---- */
constexpr uint32_t SCAN_TILE = 4096; /* 256 threads × 16 elements */
size_t scan_workspace_bytes(uint32_t n);
/* out = exclusive scan of in; out2 (optional) receives a copy; total
 * (optional, device) receives the sum.  `ws` holds scan_workspace_bytes(n).
 * With d_n (device) the length is min(*d_n, n). */
hipError_t launch_scan(const uint32_t *in, uint32_t n, uint32_t *out, uint32_t *out2, uint32_t *total, void *ws,
                       hipStream_t s, const uint32_t *d_n = nullptr);
/* Size of the file-scope device array the allocation-free scan uses as its
 * workspace (ticket + one status word per tile), and its device address on
 * the current device (looked up once per device, then cached). */
constexpr size_t NOALLOC_SCAN_SCRATCH_BYTES = 64 * 1024;
hipError_t noalloc_scan_scratch(void **out);

hipError_t launch_fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t s);
/* dst[i] = src[i] for n 16-byte items */
hipError_t launch_copy16(phx_float4 *dst, const phx_float4 *src, size_t n, hipStream_t s);
/* seed[i] = random_u32(state[i]) for i < n (PCG, random.h:29-34) */
hipError_t launch_next_round_seed(uint32_t *state, uint32_t *seed, uint32_t n, hipStream_t s);
/* info[0] = *count; info[1] counts passes with a zero count */
hipError_t launch_note_count(const uint32_t *count, uint32_t *info, hipStream_t s);

hipError_t launch_trace(const PhxCameraParams &cam, uint32_t w, uint32_t h, uint32_t seed, uint32_t direct_samples,
                        const SceneArgs &sc, phx_float4 *hp_pos, phx_float4 *hp_normal, phx_float4 *hp_throughput,
                        phx_float4 *hp_base_color, phx_float4 *hp_direct, phx_float4 *hp_emission, hipStream_t s,
                        const uint32_t *d_seed = nullptr);

/* with_scene: room for a scene index per shadow slot (multi-scene queues) */
size_t trace_wavefront_workspace_bytes(uint32_t n_pixels, uint32_t direct_samples, bool with_scene = false);
hipError_t launch_trace_wavefront(const PhxCameraParams &cam, uint32_t w, uint32_t h, uint32_t seed,
                                  uint32_t direct_samples, const SceneArgs &sc, uint32_t shadow_any_hit,
                                  phx_float4 *hp_pos, phx_float4 *hp_normal, phx_float4 *hp_throughput,
                                  phx_float4 *hp_base_color, phx_float4 *hp_direct, phx_float4 *hp_emission,
                                  void *ws, hipStream_t s, const uint32_t *d_seed = nullptr,
                                  const struct LaunchShape *shape = nullptr);

size_t emit_workspace_bytes(uint32_t photons_to_emit, uint32_t max_bounces);
hipError_t launch_emit(const SceneArgs &sc, uint32_t max_photons, uint32_t offset, uint32_t photons_to_emit,
                       uint32_t max_bounces, uint32_t seed, phx_float4 *photon_pos, uint32_t *photon_power,
                       uint32_t *photon_dir, uint32_t *n_photons, void *ws, hipStream_t s,
                       const uint32_t *d_seed = nullptr, uint32_t wave_bounces = 0,
                       const struct LaunchShape *shape = nullptr, phx_float4 *deposit_rec = nullptr);
/* deposit_rec: the deposits go to 32-byte records (position; packed power,
 * octahedral direction, 0, 0) instead of photon_pos / photon_power /
 * photon_dir */

hipError_t launch_hash(const phx_float4 *pos, uint32_t n, const PhxPhotonHashData &d, uint32_t *hashes,
                       uint32_t *counts, hipStream_t s, bool octants = false);
/* Same with the photon count read on the device (clamped to max_n); counts
 * may be null (keys only). */
hipError_t launch_hash_dev(const phx_float4 *pos, const uint32_t *d_n, uint32_t max_n, const PhxPhotonHashData &d,
                           uint32_t *hashes, uint32_t *counts, hipStream_t s, bool octants = false,
                           uint32_t pos_stride = 1);
/* pos_stride: float4s from one position to the next (2 for deposit records) */

hipError_t launch_scatter_atomic(const phx_float4 *pos_in, const uint32_t *pow_in, const uint32_t *dir_in,
                                 const uint32_t *hashes, uint32_t *cell_end, uint32_t n, phx_float4 *pos_out,
                                 uint32_t *pow_out, uint32_t *dir_out, hipStream_t s);

size_t sort_workspace_bytes(uint32_t n, uint32_t n_cells);
hipError_t launch_sort(const phx_float4 *pos_in, const uint32_t *pow_in, const uint32_t *dir_in,
                       const uint32_t *hashes, const uint32_t *cell_start, uint32_t *cell_end, uint32_t n,
                       uint32_t n_cells, phx_float4 *pos_out, uint32_t *pow_out, uint32_t *dir_out, void *ws,
                       hipStream_t s, const uint32_t *d_n = nullptr, uint32_t *perm_out = nullptr,
                       uint32_t *cell_start_out = nullptr, uint32_t *oct_start_out = nullptr,
                       const phx_float4 *rec_in = nullptr, phx_float4 *dirpow_out = nullptr);
/* rec_in: the records come as deposit records (launch_emit's deposit_rec)
 * instead of pos_in / pow_in / dir_in, so the last pass reads each with one
 * access; dirpow_out: the last pass also writes each record's decoded
 * direction with its packed power in w (launch_decode_photons' output). */
/* With cell_start_out the cell starts come from the sorted keys (binary
 * search: the same numbers as the exclusive scan of the cell counts) and
 * cell_start is ignored, so the caller needs neither counts nor scan.  With
 * oct_start_out (8·n_cells entries) the keys are
 * octant keys (get_hash_octant, 8h + octant): records are sorted by cell,
 * then octant, then input order, and oct_start_out receives every non-empty
 * cell's octant starts (k_octant_starts). */

hipError_t launch_gather(const phx_float4 *hp_pos, const phx_float4 *hp_normal, const phx_float4 *hp_throughput,
                         const phx_float4 *hp_base_color, const phx_float4 *hp_direct, const phx_float4 *hp_emission,
                         uint32_t n_pixels, const phx_float4 *ppos, const uint32_t *ppow, const uint32_t *pdir,
                         const uint32_t *cell_start, const uint32_t *cell_end, const PhxPhotonHashData &d,
                         float alpha, phx_float4 *flux, float *photon_weight, phx_float4 *radiance, uint32_t width,
                         hipStream_t s, float *rec_ratio = nullptr, float *rec_r2 = nullptr,
                         const phx_float4 *pdirpow = nullptr, const uint32_t *oct_start = nullptr,
                         uint32_t gather_tile_w = 8);
/* every stored photon record's decoded direction with its packed power in w
 * (count on the device, clamped to max_n) */
hipError_t launch_decode_photons(const uint32_t *ppow, const uint32_t *pdir, const uint32_t *d_n, uint32_t max_n,
                                 phx_float4 *dirpow, hipStream_t s);

/* phx_kernel_info's entries for the trace/emission and the photon-map
 * kernels: appended to out[n] while n < max, every kernel counted in n;
 * false when the runtime cannot describe one */
bool kernel_info_trace(PhxKernelInfo *out, uint32_t max, uint32_t &n);
bool kernel_info_map(PhxKernelInfo *out, uint32_t max, uint32_t &n);

/* ---- multi-scene passes ---- */
/* One scene of a batch: its device scene, camera, grid and its ranges in the
 * batch-wide arrays (dense pixels, 128-pixel chunks, 8×8 gather tiles,
 * photons, grid buckets). */
struct BatchScene {
    SceneArgs sc;
    PhxCameraParams cam;
    PhxPhotonHashData hash;
    uint32_t width, height, n_pixels, px_base;
    uint32_t chunk_base, tile_base, tiles_x, tiles_y;
    uint32_t photons, ph_base, max_bounces, direct_samples;
    uint32_t bucket_base, n_buckets, any_hit, pad0;
    float alpha;
    uint32_t pad1[3];
};

/* scenes [s0, s1) of a trace or emission group and their dense range */
struct BatchGroup {
    uint32_t s0, s1, first, count, chunk0, chunks;
};

/* Everything one batched pass launches over (device pointers, static sizes). */
struct BatchPlan {
    uint32_t n_scenes = 0;
    const BatchScene *scenes = nullptr; /* device table [n_scenes] */
    const uint32_t *chunk_scene = nullptr, *tile_scene = nullptr;
    const uint32_t *ph_base = nullptr; /* [n_scenes + 1] */
    uint32_t *dep_base = nullptr;      /* [n_scenes + 1], per pass */
    uint32_t *rng = nullptr, *round_seed = nullptr, *counts = nullptr, *zero_passes = nullptr;
    uint32_t *tickets = nullptr; /* one per trace group */
    uint32_t *dep_counter = nullptr;
    phx_float4 *hp[6] = {};
    phx_float4 *flux = nullptr, *radiance = nullptr;
    float *photon_weight = nullptr;
    phx_float4 *dep_pos = nullptr, *map_pos = nullptr;
    uint32_t *dep_power = nullptr, *dep_dir = nullptr, *map_power = nullptr, *map_dir = nullptr;
    uint32_t *hashes = nullptr, *cell_counts = nullptr, *cell_start = nullptr, *cell_end = nullptr;
    uint32_t *oct_start = nullptr; /* 8 octant starts per bucket (records sorted by cell, then octant) */
    phx_float4 *map_dirpow = nullptr; /* decoded direction, packed power in w */
    uint32_t dep_cap = 0, total_buckets = 0, total_tiles = 0;
    void *trace_ws = nullptr, *emit_ws = nullptr, *scan_ws = nullptr, *sort_ws = nullptr;
    uint32_t trace_max_pixels = 0, max_direct_samples = 0, emit_max_photons = 0, max_bounces = 0;
    uint32_t cam_blocks = 0;
    std::vector<BatchGroup> trace_groups, emit_groups;
};

size_t batch_trace_workspace_bytes(uint32_t max_pixels, uint32_t max_direct_samples);
size_t batch_emit_workspace_bytes(uint32_t max_photons, uint32_t max_bounces);
/* persistent camera-kernel blocks that fit the device at once */
uint32_t batch_camera_blocks();
/* the streams of one batched pass: the camera trace on `s`, the emission and
 * the photon map on `side` (forked after the round seeds with ev_fork, joined
 * before the gather with ev_join); the gather on `s`, or, when `gather` is
 * set, on that stream after the trace (ev_trace).  ev_emit, when set, is
 * recorded on `side` after the emission, and h_info, when set, receives the
 * per-scene deposit counts and passes without deposits there. */
struct BatchStreams {
    hipStream_t s = nullptr, side = nullptr, gather = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_trace = nullptr, ev_emit = nullptr;
    uint32_t *h_info = nullptr;
};
/* one batched pass over every scene of the plan (launches and event
 * records only) */
hipError_t launch_batch_pass(const BatchPlan &P, const BatchStreams &B);

/* ---- adjoint ---- */
/* what the gradient needs per material (host-filled, device table) */
struct AdjMaterial {
    float inv_bc[3];  /* 1 / base_color */
    float rough_gate; /* dα/dθ = 2·roughness where the roughness is not clamped, else 0 */
    float inv_emis;   /* 1 / emission scale (0 for materials that do not emit) */
    float pad[3];
};

/* samples over n·v ∈ [0, 1] of a material's VNDF rejection terms */
constexpr int ACC_MU = 65;
/* [∂log A/∂α, 8·∂log P/∂α] per sample (A: acceptance within 8 tries, P: the
 * rejection probability of one try) for a roughness, into out[2·ACC_MU] */
void rejection_table(float roughness, float *out);

/* everything one pass of the backward sweep reads and writes */
struct AdjPass {
    SceneArgs sc;
    PhxCameraParams cam;
    PhxPhotonHashData hash;
    uint32_t w, h, n_direct, photons, max_bounces, M, max_deposits;
    const uint32_t *d_seed;
    const AdjMaterial *am;
    const float *rej; /* [material][ACC_MU][2] rejection_table */
    const uint32_t *light_mat;
    const phx_float4 *hp_pos, *hp_normal, *hp_throughput, *hp_base_color, *hp_emission;
    const phx_float4 *ppos;
    const uint32_t *ppow, *pdir, *cell_start, *cell_end, *perm, *d_count;
    void *emit_ws;
    const float *ratio, *r2;
    float *R;
    const phx_float4 *lam_rad, *lam_tau;
    phx_float4 *nf, *adj_sorted, *adj_dep;
    float *grad;
    phx_float4 *chk_emission, *chk_direct;
    uint32_t *chk_count;
};
hipError_t launch_adjoint_pass(const AdjPass &p, hipStream_t s);
/* per-photon deposit counts of the last emission in an emission workspace */
const uint32_t *adjoint_emit_counts(void *emit_ws, uint32_t photons, uint32_t max_bounces);

/* error bookkeeping of the C ABI */
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_status(hipError_t e, const char *what);

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

/* pcg_seed of random.h:16-24 on the host (the round-seed generator) */
inline uint32_t host_pcg_seed(uint32_t seed) {
    uint32_t s = 0u;
    s = s * 747796405u + 2891336453u;
    s += seed;
    s = s * 747796405u + 2891336453u;
    return s;
}

/* PhxRenderArgs checks of cmd_args.hpp shared by the drivers */
int check_render_args(const PhxRenderArgs &a, const char *who);

} // namespace phx

#endif /* PHX_INTERNAL_H */
