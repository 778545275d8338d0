/*
 * phosphor_hip.h — C ABI of the MI355X (gfx950) SPPM pass kernels.
 *
 * Every entry point replaces one OpenCL enqueue of alfazet/phosphor's pass loop
 * (src/renderer/renderer.cpp:90-161).  Arguments keep the order and meaning of
 * the `.cl` kernel signature they replace: a `cl::Buffer` becomes a device
 * pointer, a by-value struct (CameraParams, PhotonHashData, float4) becomes a
 * `const *` to host memory holding the same bytes, scalars stay u32/f32, and a
 * trailing `stream` (a hipStream_t; NULL = the null stream) orders the launch.
 *
 * All launches are asynchronous on `stream`.  Every function returns 0 on
 * success or a PHX_ERR_* / hipError_t code and never throws; the message of
 * the last failure on the calling thread is returned by phx_last_error().
 *
 * The structs below restate the reference's host/device POD layouts byte for
 * byte (the structs of include/common/); their sizes and field offsets are checked by
 * static assertions in phosphor_amd/csrc/abi.cpp and by the CPU tests.
 */
#ifndef PHOSPHOR_HIP_H
#define PHOSPHOR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHX_ALIGN16 __attribute__((aligned(16)))

/* ---- scalar/vector types (include/common/typedefs.h:4-27) ---- */
typedef struct PHX_ALIGN16 phx_float4 { float x, y, z, w; } phx_float4;
typedef struct phx_float2 { float x, y; } phx_float2;

/* ---- boundary structs (byte-exact with the reference) ---- */

/* include/common/camera.h:7-22 — 112 bytes */
typedef struct PHX_ALIGN16 PhxCameraParams {
    phx_float4 position;
    phx_float4 lower_left_corner;
    phx_float4 horizontal;
    phx_float4 vertical;
    phx_float4 defocus_disk_u;
    phx_float4 defocus_disk_v;
    float defocus_angle;
    float focus_distance;
    uint8_t _padding[8];
} PhxCameraParams;

/* include/common/triangle.h:7-28 — 48/48/48/32 bytes */
typedef struct PHX_ALIGN16 PhxTriangleVertices { phx_float4 v0, v1, v2; } PhxTriangleVertices;
typedef struct PHX_ALIGN16 PhxTriangleNormals { phx_float4 n0, n1, n2; } PhxTriangleNormals;
typedef struct PHX_ALIGN16 PhxTriangleTangents { phx_float4 t0, t1, t2; } PhxTriangleTangents;
typedef struct PHX_ALIGN16 PhxTriangleUvs {
    phx_float2 uv0, uv1, uv2;
    uint8_t _padding[8];
} PhxTriangleUvs;

/* include/common/bvh_node.h:6-8 + bounding_box.h:7-10 — 32 bytes.
 * min.w = as_float(triangle index), 0xFFFFFFFF for an interior node;
 * max.w = as_float(index of the right child) (src/renderer/bvh.cpp:59-72). */
typedef struct PHX_ALIGN16 PhxBvhNode { phx_float4 min, max; } PhxBvhNode;

/* No counterpart in the reference: one interior node of a binary tree with
 * both children's boxes, 64 bytes (one cache line), so a search step tests
 * both children with one load.  lmin.w / rmin.w = as_float(child reference):
 * for a leaf child PHX_PAIR_LEAF | first << 2 | (count − 1), the leaf's
 * triangles being positions first .. first + count − 1 of the tree's
 * triangle order; else the child's pair index.  lmax.w = rmax.w = 0. */
typedef struct PHX_ALIGN16 PhxBvhPair { phx_float4 lmin, lmax, rmin, rmax; } PhxBvhPair;
#define PHX_PAIR_LEAF 0x80000000u
#define PHX_PAIR_MAX_LEAF 4u            /* triangles per leaf */
#define PHX_PAIR_MAX_TRIS 0x1FFFFFF0u   /* triangles a tree of pairs can order */

/* include/common/material.h:7-27 — 32 / 192 bytes */
typedef struct PHX_ALIGN16 PhxUvTransform {
    phx_float2 uv_offset;
    phx_float2 uv_scale;
    float uv_rotation;
    uint8_t _padding[12];
} PhxUvTransform;
typedef struct PHX_ALIGN16 PhxUvTransformBlock {
    PhxUvTransform diff, emis, norm, occlusion, metal_rough, transmission;
} PhxUvTransformBlock;

/* include/common/material.h:36-59 — 96 bytes */
typedef struct PHX_ALIGN16 PhxMaterial {
    phx_float4 base_color;
    phx_float4 emissive;
    phx_float4 att_color;
    float metallic, roughness, transmission, ior, att_dist, thickness;
    uint32_t diff_index, emis_index, norm_index, occlusion_index, metal_rough_index, transmission_index;
} PhxMaterial;

/* include/common/texture_meta.h:7-15 — 16 bytes */
typedef struct PHX_ALIGN16 PhxTextureMeta {
    uint32_t atlas_offset, width, height, channels;
} PhxTextureMeta;

/* include/common/light.h:9-31 — 128 bytes */
#define PHX_LIGHT_POINT 0
#define PHX_LIGHT_SPOT 1
#define PHX_LIGHT_DIRECTIONAL 2
#define PHX_LIGHT_TEXTURED 3
typedef struct PHX_ALIGN16 PhxLight {
    phx_float4 position, power, direction, tangent, bitangent, origin, aux;
    uint8_t kind;
    uint8_t _padding[15];
} PhxLight;

/* include/common/photon_hash.h:11-19 — 48 bytes */
typedef struct PHX_ALIGN16 PhxPhotonHashData {
    phx_float4 origin;
    phx_float4 cell_sizes;
    uint32_t grid_res;
    uint8_t _padding[12];
} PhxPhotonHashData;

/* ---- status codes ---- */
#define PHX_OK 0
#define PHX_ERR_INVALID_ARG 10001
#define PHX_ERR_NO_PHOTONS 10002     /* renderer.cpp:117-118 "no photon hit" */
#define PHX_ERR_WORKSPACE 10003      /* workspace smaller than the *_workspace_size query */
#define PHX_ERR_NO_DEVICE 10004
#define PHX_ERR_TOO_LARGE 10005

/* Message of the last failing call on this thread ("" if none). */
const char *phx_last_error(void);
/* Library version string. */
const char *phx_version(void);

/* ======================================================================
 * Kernel-level entry points (one per reference kernel)
 * ====================================================================== */

/* Replaces kernels/trace_rays.cl:14-40 (enqueued at renderer.cpp:90-95,
 * arguments bound at src/buffers/rendering.cpp:32-42).  One work item per
 * pixel: camera ray → first diffuse hitpoint, NEE direct light, specular
 * emission, envmap on miss.  Writes hp_pos/hp_normal/hp_throughput/
 * hp_base_color/hp_direct/hp_emission (float4 per pixel). */
int phx_trace_rays(const PhxCameraParams *camera, uint32_t image_width, uint32_t image_height, uint32_t seed,
                   uint32_t direct_samples,
                   const PhxTriangleVertices *tri_verts, const PhxTriangleNormals *tri_normals,
                   const PhxTriangleTangents *tri_tangents, const PhxTriangleUvs *tri_uvs,
                   const PhxBvhNode *bvh_nodes, const uint32_t *tri_mat_index, uint32_t n_triangles,
                   const PhxMaterial *materials, const PhxUvTransformBlock *uv_transforms,
                   const PhxTextureMeta *tex_meta, const uint8_t *tex_atlas,
                   const PhxLight *lights, uint32_t n_lights, const float *lum_pref_sum, float total_lum,
                   const phx_float4 *scene_center, float scene_radius,
                   const PhxTriangleVertices *etri_verts, const PhxTriangleNormals *etri_normals,
                   const PhxTriangleUvs *etri_uvs, const uint32_t *etri_mat_index,
                   const float *envmap_data, uint32_t envmap_width, uint32_t envmap_height, uint32_t has_envmap,
                   phx_float4 *hp_pos, phx_float4 *hp_normal, phx_float4 *hp_throughput,
                   phx_float4 *hp_base_color, phx_float4 *hp_direct, phx_float4 *hp_emission,
                   void *stream);

/* Device scratch needed by phx_trace_rays_wavefront. */
size_t phx_trace_rays_wavefront_workspace_size(uint32_t image_width, uint32_t image_height, uint32_t direct_samples);

/* Same arguments and the same outputs, bit for bit, as phx_trace_rays (the
 * trace_rays.cl:14-40 list), computed in three launches: camera paths that
 * draw every NEE light fragment and queue its shadow ray, a shadow-ray kernel
 * (the direct_samples rays of one event are adjacent in the queue), and a
 * resolve pass that sums the NEE terms in the reference's order.  Extra
 * arguments: shadow_any_hit = 1 lets shadow rays stop at their first hit,
 * which is exact only when no material has transmission >= 1e-5 and no BVH
 * interior node is deeper than 22 (so the 24-entry stack never overflows);
 * the caller vouches for both.  `workspace` holds
 * phx_trace_rays_wavefront_workspace_size(width, height, direct_samples) bytes. */
int phx_trace_rays_wavefront(const PhxCameraParams *camera, uint32_t image_width, uint32_t image_height,
                             uint32_t seed, uint32_t direct_samples,
                             const PhxTriangleVertices *tri_verts, const PhxTriangleNormals *tri_normals,
                             const PhxTriangleTangents *tri_tangents, const PhxTriangleUvs *tri_uvs,
                             const PhxBvhNode *bvh_nodes, const uint32_t *tri_mat_index, uint32_t n_triangles,
                             const PhxMaterial *materials, const PhxUvTransformBlock *uv_transforms,
                             const PhxTextureMeta *tex_meta, const uint8_t *tex_atlas,
                             const PhxLight *lights, uint32_t n_lights, const float *lum_pref_sum, float total_lum,
                             const phx_float4 *scene_center, float scene_radius,
                             const PhxTriangleVertices *etri_verts, const PhxTriangleNormals *etri_normals,
                             const PhxTriangleUvs *etri_uvs, const uint32_t *etri_mat_index,
                             const float *envmap_data, uint32_t envmap_width, uint32_t envmap_height,
                             uint32_t has_envmap, phx_float4 *hp_pos, phx_float4 *hp_normal,
                             phx_float4 *hp_throughput, phx_float4 *hp_base_color, phx_float4 *hp_direct,
                             phx_float4 *hp_emission, uint32_t shadow_any_hit, void *workspace,
                             size_t workspace_bytes, void *stream);

/* Device scratch needed by phx_emit_photons for one batch. */
size_t phx_emit_photons_workspace_size(uint32_t photons_to_emit, uint32_t max_photon_bounces);

/* Replaces kernels/emit_photons.cl:12-32 (renderer.cpp:97-121, arguments bound
 * at rendering.cpp:44-56).  Photon `offset + i` (i < photons_to_emit) is traced
 * with PCG stream `seed + offset + i`; every diffuse deposit is appended at
 * index *n_photons (a device counter the caller resets per pass, as
 * renderer.cpp:99 does).  Unlike the reference's atomic_inc order, deposits
 * land in the canonical order (photon index, bounce), so repeated runs are
 * bit-identical; the counter counts every deposit, and deposits at or beyond
 * max_photons are dropped (emit_photons.cl:98-105).  `workspace` must hold
 * phx_emit_photons_workspace_size(photons_to_emit, max_photon_bounces) bytes. */
int phx_emit_photons(const PhxLight *lights, uint32_t n_lights, uint32_t max_photons, uint32_t offset,
                     uint32_t photons_to_emit, uint32_t max_photon_bounces, uint32_t seed,
                     const PhxBvhNode *tree, const PhxTriangleVertices *tri_verts,
                     const PhxTriangleNormals *tri_normals, const PhxTriangleTangents *tri_tangents,
                     const PhxTriangleUvs *tri_uvs, const uint32_t *tri_mat_index, uint32_t n_tris,
                     const PhxTriangleVertices *etri_verts, const PhxTriangleNormals *etri_normals,
                     const PhxTriangleUvs *etri_uvs, const uint32_t *etri_mat_index,
                     const PhxMaterial *materials, const PhxUvTransformBlock *uv_transforms,
                     const PhxTextureMeta *tex_meta, const uint8_t *tex_atlas, const float *lum_pref_sum,
                     float total_lum, const phx_float4 *scene_center, float scene_radius,
                     phx_float4 *photon_pos, uint32_t *photon_power, uint32_t *photon_dir, uint32_t *n_photons,
                     void *workspace, size_t workspace_bytes, void *stream);

/* Replaces kernels/hash_photons.cl:4-5 (renderer.cpp:128-132).  photon_hashes[i]
 * = get_hash(photon_pos[i]) (photon_hash.h:21-37, 0 = outside the grid) and
 * cell_counts[h] += 1.  Like the reference, cell_counts is accumulated into:
 * the caller zeroes it first (renderer.cpp:128). */
int phx_hash_photons(const phx_float4 *photon_pos, uint32_t n_photons, const PhxPhotonHashData *data,
                     uint32_t *photon_hashes, uint32_t *cell_counts, void *stream);

/* Device scratch needed by phx_scan_cells for n_cells elements. */
size_t phx_scan_cells_workspace_size(uint32_t n_cells);

/* Replaces kernels/scan_small_chunks.cl + scan_big_chunks.cl +
 * scan_add_offsets.cl and the cell_start→cell_end copy (renderer.cpp:134-147)
 * with one single-pass scan.  Postcondition: cell_start = exclusive prefix sum
 * of cell_counts and cell_end = cell_start.  No 4096² bucket limit. */
int phx_scan_cells(const uint32_t *cell_counts, uint32_t n_cells, uint32_t *cell_start, uint32_t *cell_end,
                   void *workspace, size_t workspace_bytes, void *stream);

/* Largest n_cells phx_scan_cells_noalloc accepts (8160 tiles of 4096 cells:
 * above the reference's 4096² scan limit and the 255³ + 1 buckets of its
 * largest grid). */
#define PHX_SCAN_NOALLOC_MAX_CELLS 33423360u

/* No counterpart in the reference.  phx_scan_cells for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same scan and postcondition, with its
 * workspace in a 64 KiB file-scope device array instead of a caller buffer;
 * the call allocates nothing.  n_cells above PHX_SCAN_NOALLOC_MAX_CELLS is
 * refused with PHX_ERR_TOO_LARGE.  All calls on one GPU share that GPU's
 * array: calls that overlap on one GPU (concurrent streams or threads) race
 * and must be serialised by the caller; calls in order on one stream are
 * safe (each clears the words it uses). */
int phx_scan_cells_noalloc(const uint32_t *cell_counts, uint32_t n_cells, uint32_t *cell_start, uint32_t *cell_end,
                           void *stream);

/* Replaces kernels/scatter_photons.cl:4-7 with its exact semantics:
 * dst = atomic_inc(&cell_end[h]); records copied to dst.  Order inside a cell
 * follows atomic arrival (nondeterministic, as in the reference). */
int phx_scatter_photons(const phx_float4 *photon_pos_in, const uint32_t *photon_power_in,
                        const uint32_t *photon_dir_in, const uint32_t *photon_hashes, uint32_t *cell_end,
                        uint32_t n_photons, phx_float4 *out_photon_pos, uint32_t *out_photon_power,
                        uint32_t *out_photon_dir, void *stream);

/* Device scratch needed by phx_sort_photons. */
size_t phx_sort_photons_workspace_size(uint32_t n_photons, uint32_t n_cells);

/* Deterministic counterpart of scatter_photons.cl: a stable counting sort by
 * cell.  Same inputs and the same postcondition (records grouped by cell,
 * cell_end[h] = cell_start[h] + count(h)), but inside a cell records keep
 * their input order, so the sorted photon map is canonical.  cell_start must
 * hold the exclusive scan of the counts and cell_end must equal cell_start on
 * entry (the phx_scan_cells postcondition). */
int phx_sort_photons(const phx_float4 *photon_pos_in, const uint32_t *photon_power_in,
                     const uint32_t *photon_dir_in, const uint32_t *photon_hashes, const uint32_t *cell_start,
                     uint32_t *cell_end, uint32_t n_photons, uint32_t n_cells, phx_float4 *out_photon_pos,
                     uint32_t *out_photon_power, uint32_t *out_photon_dir, void *workspace,
                     size_t workspace_bytes, void *stream);

/* Replaces kernels/gather_flux.cl:5-22 (renderer.cpp:155-161, arguments bound
 * at rendering.cpp:71-77): per-pixel 27-cell Gaussian photon gather and the
 * SPPM radius/flux update (photon_hash.h:56-103, gather_flux.cl:28-77). */
int phx_gather_flux(const phx_float4 *hp_pos, const phx_float4 *hp_normal, const phx_float4 *hp_throughput,
                    const phx_float4 *hp_base_color, const phx_float4 *hp_direct, const phx_float4 *hp_emission,
                    uint32_t n_pixels, const phx_float4 *photon_pos, const uint32_t *photon_power,
                    const uint32_t *photon_dir, uint32_t n_photons, const uint32_t *cell_start,
                    const uint32_t *cell_end, const PhxPhotonHashData *data, float sppm_alpha,
                    phx_float4 *flux, float *photon_weight, phx_float4 *radiance, void *stream);

/* ---- Faster forms of the stages above, same results ----
 *
 * The SAH child-pair tree for the trace entry points.  No counterpart in the
 * reference: its set_scene builds the median tree (bvh.cpp:74-84) and
 * uploads it with the triangles (renderer.cpp:32-34, geometry.cpp); a
 * drop-in calls phx_accel_attach once after that upload with the device
 * pointers and the host arrays it uploaded.  Afterwards phx_trace_rays,
 * phx_trace_rays_wavefront and phx_emit_photons, called on the current
 * device with that bvh_nodes pointer and that tri_verts pointer, trace the
 * library's SAH tree of the same triangles (leaves of up to leaf_size, 0 = 3)
 * stored as 64-byte child pairs, and return the same hit records bit for bit.
 * That holds where the reference's search cannot drop a subtree and is
 * insensitive to visiting order: one triangle per leaf, nodes in preorder,
 * leaf boxes holding their triangle's box grown by 1e-3 (bounding_box.h),
 * interior boxes holding their children's, no interior node deeper than 22,
 * coordinates within ±1024.  phx_accel_eligible checks this on the host (1 or
 * 0; when 0, phx_last_error says which property fails); phx_accel_attach
 * sets *accelerated to 0 and keeps the reference tree when it fails.  The
 * device arrays must hold what the host arrays do (attach compares them) and
 * must not change while attached; phx_accel_detach (which synchronises the
 * device) before they are freed or rewritten. */
int phx_accel_eligible(const PhxBvhNode *host_tree, uint32_t n_nodes, const PhxTriangleVertices *host_tri_verts,
                       uint32_t n_triangles);
int phx_accel_attach(const PhxBvhNode *device_tree, const PhxTriangleVertices *device_tri_verts,
                     const PhxBvhNode *host_tree, uint32_t n_nodes, const PhxTriangleVertices *host_tri_verts,
                     uint32_t n_triangles, uint32_t leaf_size, uint32_t *accelerated);
int phx_accel_detach(const PhxBvhNode *device_tree);

/* Device scratch needed by phx_sort_photons_by_octant. */
size_t phx_sort_photons_by_octant_workspace_size(uint32_t n_photons, uint32_t n_cells);

/* hash_photons.cl + the scan + scatter_photons.cl (renderer.cpp:123-153) in
 * one call, for the gather below.  The reference leaves the order inside a
 * cell to atomic arrival (scatter_photons.cl:13); here records are sorted by
 * cell, then by the half-cell octant that holds them (per axis, whether the
 * position lies in the upper half of its cell, from get_hash's quotients,
 * photon_hash.h:21-37), then by input order.  Writes the sorted records,
 * cell_start/cell_end for all n_cells (>= grid_res^3 + 1) buckets, and for
 * every non-empty cell its eight octant starts (oct_start[8h + o]; entries of
 * empty cells are left as they were).  out_dir_power (nullable) receives each
 * sorted record's decoded direction (encoding.h decode_oct) with the packed
 * power in w, the form phx_gather_flux_octants reads. */
int phx_sort_photons_by_octant(const phx_float4 *photon_pos_in, const uint32_t *photon_power_in,
                               const uint32_t *photon_dir_in, uint32_t n_photons, const PhxPhotonHashData *data,
                               uint32_t n_cells, phx_float4 *out_photon_pos, uint32_t *out_photon_power,
                               uint32_t *out_photon_dir, phx_float4 *out_dir_power, uint32_t *cell_start,
                               uint32_t *cell_end, uint32_t *oct_start, void *workspace, size_t workspace_bytes,
                               void *stream);

/* gather_flux.cl on a map from phx_sort_photons_by_octant: the same
 * per-pixel sums as phx_gather_flux over that map (the order of the records
 * is the map's), with the photon directions read decoded (photon_dir_power),
 * whole half-cell octants beyond the radius skipped (oct_start; null: cells
 * only) and one wave per 8×8 tile of the image. */
int phx_gather_flux_octants(const phx_float4 *hp_pos, const phx_float4 *hp_normal, const phx_float4 *hp_throughput,
                            const phx_float4 *hp_base_color, const phx_float4 *hp_direct,
                            const phx_float4 *hp_emission, uint32_t image_width, uint32_t image_height,
                            const phx_float4 *photon_pos, const phx_float4 *photon_dir_power, uint32_t n_photons,
                            const uint32_t *cell_start, const uint32_t *cell_end, const uint32_t *oct_start,
                            const PhxPhotonHashData *data, float sppm_alpha, phx_float4 *flux, float *photon_weight,
                            phx_float4 *radiance, void *stream);

/* ======================================================================
 * Device-memory plumbing (no reference counterpart: lets hosts without their
 * own HIP allocator, such as the Python tooling, own device buffers through
 * the same HIP runtime as the kernels)
 * ====================================================================== */
int phx_device_count(void);
/* Register use and occupancy of the pass's main kernels on the current
 * device, as the HIP runtime reports them: VGPRs per lane, static LDS bytes
 * per block, scratch bytes per lane, block size, and the waves per SIMD that
 * can be resident at once (blocks per CU from the occupancy calculator ×
 * waves per block / 4 SIMDs). */
typedef struct PhxKernelInfo {
    char name[32];
    uint32_t vgprs, lds_bytes, scratch_bytes, block, waves_per_simd;
} PhxKernelInfo;
/* Writes up to `max` entries and returns the number of kernels described
 * (call with max = 0 to count them), or a negative PHX_ERR_* (no device). */
int phx_kernel_info(PhxKernelInfo *out, uint32_t max);
int phx_device_malloc(void **out, size_t bytes);
int phx_device_free(void *p);
int phx_memcpy_htod(void *dst, const void *src, size_t bytes);
int phx_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int phx_memset_device(void *dst, int value, size_t bytes);
/* Device-to-device copy with a plain streaming kernel (16-byte loads and
 * stores, bytes a multiple of 16): the known-traffic reference used to scale
 * hardware byte counters. */
int phx_device_copy(void *dst, const void *src, size_t bytes, void *stream);
int phx_device_synchronize(void);

/* ======================================================================
 * Host helpers the pass driver needs (src/renderer/bvh.cpp, photon_hash.h)
 * ====================================================================== */

/* src/renderer/bvh.cpp:18-84: median split on the longest axis of the node
 * bbox, one triangle per leaf, DFS-flattened (left child = index + 1).
 * `tri_verts` is host memory with n_tris entries; `out_nodes` must hold
 * 2*n_tris-1 nodes.  Returns the node count (0 on error). */
uint32_t phx_build_bvh_median(const PhxTriangleVertices *tri_verts, uint32_t n_tris, PhxBvhNode *out_nodes);

/* No counterpart in the reference: a binned surface-area-heuristic tree over
 * the same expanded triangle boxes, at most `max_leaf` (1 .. PHX_PAIR_MAX_LEAF)
 * triangles per leaf, no interior node deeper than `max_depth`, in the node
 * format above (preorder, left child at index + 1).  out_order receives the
 * tree's triangle order (n_tris entries, a permutation of the triangles);
 * leaf min.w = as_float(PHX_PAIR_LEAF | first << 2 | (count − 1)) names the
 * leaf's positions in that order, leaf max.w = 0; interior min.w =
 * as_float(0xFFFFFFF8 | hint), hint = split axis | 4 when the left child lies
 * above the right one.  `out_nodes` must hold 2*n_tris-1 nodes.  Returns the
 * node count, 0 on error (n_tris beyond PHX_PAIR_MAX_TRIS, max_leaf out of
 * range) or when no tree within max_depth exists. */
uint32_t phx_build_bvh_sah(const PhxTriangleVertices *tri_verts, uint32_t n_tris, uint32_t max_depth,
                           uint32_t max_leaf, PhxBvhNode *out_nodes, uint32_t *out_order);

/* photon_hash.h:41-50 build_photon_hash_info(bbox, grid_res), written to *out. */
/* No counterpart in the reference: the interior nodes of a tree built by
 * phx_build_bvh_sah (preorder, left child at index + 1, interior min.w >=
 * 0xFFFFFFF8, leaf min.w a leaf reference) as pairs, in the same preorder;
 * out_pairs holds n_nodes / 2 entries (none for a single leaf).  *root_ref is
 * the reference of node 0.  Returns the number of pairs, or -1 on a malformed
 * tree. */
int phx_build_bvh_pairs(const PhxBvhNode *tree, uint32_t n_nodes, PhxBvhPair *out_pairs, uint32_t *root_ref);

int phx_build_photon_hash_info(const PhxBvhNode *root, uint32_t grid_res, PhxPhotonHashData *out);

/* ======================================================================
 * Pass driver (src/renderer/renderer.cpp Renderer::set_scene / run)
 * ====================================================================== */

/* Host-side scene description (SceneData, include/host/scene_data.hpp:21-42,
 * already split to SoA as GeometryBuffers::upload does).  All pointers are
 * host memory; the renderer copies them to the device. */
typedef struct PhxSceneDesc {
    const PhxTriangleVertices *tri_verts;
    const PhxTriangleNormals *tri_normals;
    const PhxTriangleTangents *tri_tangents;
    const PhxTriangleUvs *tri_uvs;
    const uint32_t *tri_mat_index;
    uint32_t n_triangles;
    const PhxTriangleVertices *etri_verts;
    const PhxTriangleNormals *etri_normals;
    const PhxTriangleUvs *etri_uvs;
    const uint32_t *etri_mat_index;
    uint32_t n_e_triangles;
    const PhxMaterial *materials;
    const PhxUvTransformBlock *uv_transforms;
    uint32_t n_materials;
    const PhxLight *lights;
    uint32_t n_lights;
    const PhxTextureMeta *tex_meta;
    uint32_t n_textures;
    const uint8_t *tex_atlas;
    uint64_t tex_atlas_bytes;
    const float *envmap; /* 3 floats per texel, NULL if none */
    uint32_t envmap_width, envmap_height;
    const PhxCameraParams *camera;
    /* optional prebuilt BVH (NULL → median-split build as bvh.cpp does) */
    const PhxBvhNode *bvh_nodes;
    uint32_t n_bvh_nodes;
} PhxSceneDesc;

/* The subset of ArgsList (include/host/cmd_args.hpp:36-65) the pass reads. */
typedef struct PhxRenderArgs {
    uint32_t width, height;
    uint32_t photons_per_round;     /* -p */
    uint32_t max_photons_per_batch; /* --batch-size */
    uint32_t max_photon_bounces;    /* --max-bounces */
    uint32_t grid_res;              /* --grid-res */
    uint32_t direct_samples;        /* --direct-samples */
    uint32_t seed;                  /* --seed */
    float sppm_alpha;               /* --sppm-alpha */
    float initial_radius;           /* 0 → reference rule (cell half-diagonal, gather_flux.cl:51-56) */
    uint32_t trace_mode;            /* 0 = wavefront trace (default), 1 = one-kernel trace; same results */
    uint32_t flags;                 /* PHX_PASS_* */
    uint32_t emit_wave_bounces;     /* bounce-synchronous emission launches before the finishing one
                                       (0 = default: 4 split trace + shade bounces, 5 with
                                       PHX_FUSED_TRACE or the reference tree); same results for
                                       every value */
    uint32_t bvh_leaf_size;         /* triangles per leaf of the SAH tree rays are traced on, 1 ..
                                       PHX_PAIR_MAX_LEAF (0 = default 3); same results for every value */
    /* Launch shapes (the reference's --workgroup, cmd_args.hpp:44, per
       kernel); same results for every value.  Threads per block of the
       camera, shadow and emission kernels: 64, 128 or 256 (0 = 128); width of
       the gather's 64-pixel wave tile: 1, 2, 4, 8, 16, 32 or 64 pixels (height
       64 / width; 0 = 8). */
    uint32_t camera_block, shadow_block, emit_block, gather_tile_width;
} PhxRenderArgs;

/* Capture the pass into a hipGraph on the first run_pass and replay it. */
#define PHX_PASS_GRAPH 1u
/* Trace on the reference's median-split BVH in its depth-first node order
 * (bvh.cpp:18-84, hit.h:136-172) even where the library would use its SAH
 * tree with near-first order (reference tree no deeper than 22 interior
 * levels, built by the library); results are identical either way. */
#define PHX_REFERENCE_BVH 2u
/* Run each pass's gather before the next pass starts.  Without this flag
 * (and without PHX_PASS_GRAPH) the gather of pass N runs on a stream of its
 * own beside the camera trace and the emission of pass N+1; the hitpoints and
 * the sorted photon map then alternate between two sets of buffers (twice
 * their device memory).  Results are identical either way. */
#define PHX_PASS_SERIAL 4u
/* Trace the primary camera rays inside the camera kernel, and trace and
 * shade each photon bounce in one kernel, one lane per photon, instead of
 * persistent trace kernels whose lanes take a new ray as soon as their search
 * ends, ahead of the kernels that shade the hits.  Results are identical
 * either way. */
#define PHX_FUSED_TRACE 8u

typedef struct PhxRenderer PhxRenderer;

/* Renderer::Renderer + set_scene (renderer.cpp:9-57). */
int phx_renderer_create(const PhxRenderArgs *args, const PhxSceneDesc *scene, PhxRenderer **out);
void phx_renderer_destroy(PhxRenderer *r);
/* One SPPM round of Renderer::run (renderer.cpp:64-78): round_seed from the
 * PCG seeded with args.seed (drawn on the device), then trace → emit →
 * hash/scan/sort → gather.  Asynchronous: the pass never waits for the host
 * (the deposit count of renderer.cpp:113-118 stays on the device). */
int phx_renderer_run_pass(PhxRenderer *r);
/* Block until all enqueued passes have finished; PHX_ERR_NO_PHOTONS if a
 * pass since the last call deposited nothing (renderer.cpp:117-118). */
int phx_renderer_synchronize(PhxRenderer *r);
/* Stage timings of the last pass in milliseconds: trace, emit, hash+scan+sort,
 * gather (zeros for graph replays, which have no stage events). */
int phx_renderer_last_stage_ms(PhxRenderer *r, float out_ms[4]);
/* No counterpart in the reference (measurement): the gather kernel's time in
 * each of the last min(passes since the reset, 64, max) eager passes, oldest
 * first, from HIP events recorded around it on the stream it runs on (so a
 * pipelined pass's figure includes what runs beside it); *n_out receives the
 * count.  Waits for those gathers. */
int phx_renderer_gather_ms(PhxRenderer *r, float *out_ms, uint32_t max, uint32_t *n_out);
int phx_renderer_reset_gather_ms(PhxRenderer *r);
/* Duration of the last pass in milliseconds (both modes). */
int phx_renderer_last_pass_ms(PhxRenderer *r, float *out_ms);
/* Checkpoint / resume of the progressive state (flux, photon_weight,
 * radiance, round count, seed generator).  No counterpart in the reference. */
size_t phx_renderer_checkpoint_bytes(PhxRenderer *r);
int phx_renderer_save_checkpoint(PhxRenderer *r, void *buf, size_t bytes);
int phx_renderer_load_checkpoint(PhxRenderer *r, const void *buf, size_t bytes);
/* Number of deposited photons in the last pass (before the clamp). */
uint32_t phx_renderer_last_photon_count(PhxRenderer *r);
uint32_t phx_renderer_rounds_done(PhxRenderer *r);
/* Copy per-pixel state to host: flux (τ, r²), photon_weight N, radiance sum. */
int phx_renderer_read_state(PhxRenderer *r, phx_float4 *flux, float *photon_weight, phx_float4 *radiance);
/* Copy the last pass's hitpoint buffers (6 × float4 per pixel, hp_pos … hp_emission order). */
int phx_renderer_read_hitpoints(PhxRenderer *r, phx_float4 *hp6);
/* Copy the last pass's sorted photon map (n = last clamped count) and cell ranges. */
int phx_renderer_read_photon_map(PhxRenderer *r, phx_float4 *pos, uint32_t *power, uint32_t *dir,
                                 uint32_t *cell_start, uint32_t *cell_end);
/* Device pointers of the renderer's buffers (for the Python tooling), after
 * waiting for every enqueued pass; the hitpoint and photon-map pointers are
 * the last pass's set (pipelined passes alternate between two): hp_pos …
 * hp_emission, flux, photon_weight, radiance, sorted pos/power/dir, cell
 * start/end, the unsorted deposits (32-byte records: position, then packed
 * power, octahedral direction, 0, 0; they replace the reference's unsorted
 * pos/power/dir arrays), NULL, NULL, hash keys, NULL (the reference's cell
 * counts: the pass reads cell ranges off the sorted keys instead), stream. */
int phx_renderer_device_buffers(PhxRenderer *r, void **ptrs, uint32_t n_ptrs);
/* Final linear estimate (output.cpp:52-83 before exposure/gamma):
 * L = radiance/rounds + flux.xyz/(π r² photons_emitted). */
int phx_renderer_estimate(PhxRenderer *r, float *rgb_out);

/* ---- Multi-scene passes ----
 * No counterpart in the reference (one Renderer renders one scene): a batch
 * renders n scenes, each with its own PhxRenderArgs (max_photons_per_batch and
 * trace_mode are not used; emission and trace run in groups of whole scenes),
 * in one sequence of launches per pass.  Each scene's hitpoints, photon map
 * and SPPM state are bit-identical to a phx_renderer_* render of that scene
 * with the same arguments. */
typedef struct PhxBatch PhxBatch;
typedef struct PhxBatchOptions {
    uint32_t trace_group_pixels; /* pixels traced per group (0 → 1 Mi) */
    uint32_t emit_group_photons; /* photons emitted per group (0 → 2 Mi) */
    uint32_t flags;              /* PHX_PASS_GRAPH | PHX_PASS_SERIAL | PHX_REFERENCE_BVH */
    uint32_t reserved;
} PhxBatchOptions;

int phx_batch_create(const PhxRenderArgs *args, const PhxSceneDesc *scenes, uint32_t n_scenes,
                     const PhxBatchOptions *options, PhxBatch **out);
void phx_batch_destroy(PhxBatch *b);
/* One pass over every scene (asynchronous, no host synchronisation). */
int phx_batch_run_pass(PhxBatch *b);
/* Block until enqueued passes finish; PHX_ERR_NO_PHOTONS names a scene whose
 * pass deposited nothing. */
int phx_batch_synchronize(PhxBatch *b);
int phx_batch_last_pass_ms(PhxBatch *b, float *out_ms);
uint32_t phx_batch_rounds_done(PhxBatch *b);
/* Deposits of each scene in the last pass. */
int phx_batch_photon_counts(PhxBatch *b, uint32_t *counts);
int phx_batch_read_state(PhxBatch *b, uint32_t scene, phx_float4 *flux, float *photon_weight,
                         phx_float4 *radiance);
int phx_batch_read_hitpoints(PhxBatch *b, uint32_t scene, phx_float4 *hp6);
/* The scene's sorted photon map with cell ranges relative to its first record. */
int phx_batch_read_photon_map(PhxBatch *b, uint32_t scene, phx_float4 *pos, uint32_t *power, uint32_t *dir,
                              uint32_t *cell_start, uint32_t *cell_end);
int phx_batch_estimate(PhxBatch *b, uint32_t scene, float *rgb_out);

/* ---- Gradients of a render ----
 * No counterpart in the reference.  θ holds five floats per material m
 * (n_materials ≤ 16, no textures): albedo r, g, b (base_color), GGX roughness,
 * emission scale (multiplies the material's emissive colour and the power of
 * the lights built from its triangles; 1 = as given).  phx_adjoint_forward
 * renders `rounds` passes from a fresh state with θ and returns the linear
 * estimate (H·W·3 floats); phx_adjoint_backward maps ∂loss/∂estimate (same
 * layout) to ∂loss/∂θ by replaying the passes (detached sampling: every random
 * decision is held, the weights are differentiated; the roughness gradient
 * adds the score of each GGX half vector and of the visible-normal sampler's
 * rejection loop, whose acceptance depends on α). */
typedef struct PhxAdjoint PhxAdjoint;
int phx_adjoint_create(const PhxRenderArgs *args, const PhxSceneDesc *scene, uint32_t max_rounds,
                       PhxAdjoint **out);
void phx_adjoint_destroy(PhxAdjoint *a);
uint32_t phx_adjoint_num_params(PhxAdjoint *a);
int phx_adjoint_forward(PhxAdjoint *a, const float *theta, uint32_t rounds, float *image_out);
int phx_adjoint_backward(PhxAdjoint *a, const float *dloss_dimage, float *grad_out);
/* Replay check: the backward's replay kernels also write what they recompute
 * for the first pass (per pixel emission and direct, per photon deposits),
 * which must equal the forward kernels' results for that pass. */
int phx_adjoint_set_replay_check(PhxAdjoint *a, uint32_t on);
int phx_adjoint_read_replay(PhxAdjoint *a, phx_float4 *emission, phx_float4 *direct, uint32_t *counts,
                            uint32_t *forward_counts);
/* The renderer the adjoint drives (its hitpoints hold the last replayed pass). */
PhxRenderer *phx_adjoint_renderer(PhxAdjoint *a);

/* GGX VNDF rejection (bsdfs.h:27-72) as the roughness gradient uses it: for
 * α ∈ (0, 1] and n·v = mu ∈ [0, 1], out[0] = probability that one draw of
 * the visible-normal sampler is rejected (mirror direction below the
 * surface), out[1] = its derivative with respect to α (host computation, no
 * device needed). */
int phx_adjoint_ggx_rejection(float alpha, float mu, double *out);

#ifdef __cplusplus
}
#endif

#endif /* PHOSPHOR_HIP_H */
