/*
 * abi.cpp — the extern "C" kernel-level entry points of include/phosphor_hip.h.
 *
 * Each function validates its arguments, packs the device pointers into the
 * kernel argument block and launches on the caller's stream.  Errors are
 * returned as codes (never thrown); the message is kept per thread.
 */
#include <cstdio>
#include <cstring>
#include <string>

#include "internal.h"

namespace phx {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_status(hipError_t e, const char *what) {
    if (e == hipSuccess)
        return PHX_OK;
    return fail((int)e, std::string(what) + ": " + hipGetErrorString(e));
}

/* Layout checks: the boundary structs must match the reference byte for byte. */
static_assert(sizeof(phx_float4) == 16, "float4");
static_assert(256 + 8 * ((PHX_SCAN_NOALLOC_MAX_CELLS + phx::SCAN_TILE - 1) / phx::SCAN_TILE) <=
                  phx::NOALLOC_SCAN_SCRATCH_BYTES,
              "the allocation-free scan's largest input fits its scratch array");
static_assert(sizeof(PhxCameraParams) == 112, "CameraParams 112 B (camera.h:7-22)");
static_assert(offsetof(PhxCameraParams, defocus_angle) == 96, "CameraParams.defocus_angle");
static_assert(sizeof(PhxTriangleVertices) == 48 && sizeof(PhxTriangleNormals) == 48 &&
                  sizeof(PhxTriangleTangents) == 48 && sizeof(PhxTriangleUvs) == 32,
              "triangle.h:7-28");
static_assert(sizeof(PhxBvhNode) == 32, "BvhNode 32 B");
static_assert(sizeof(PhxUvTransform) == 32 && sizeof(PhxUvTransformBlock) == 192, "material.h:7-27");
static_assert(sizeof(PhxMaterial) == 96, "Material 96 B (material.h:36-59)");
static_assert(offsetof(PhxMaterial, metallic) == 48 && offsetof(PhxMaterial, diff_index) == 72, "Material fields");
static_assert(sizeof(PhxTextureMeta) == 16, "TextureMeta");
static_assert(sizeof(PhxLight) == 128 && offsetof(PhxLight, kind) == 112, "Light 128 B (light.h:14-31)");
static_assert(sizeof(PhxPhotonHashData) == 48 && offsetof(PhxPhotonHashData, grid_res) == 32, "PhotonHashData");
static_assert(sizeof(PhxRenderArgs) == 72 && offsetof(PhxRenderArgs, bvh_leaf_size) == 52 &&
                  offsetof(PhxRenderArgs, gather_tile_width) == 68,
              "PhxRenderArgs");
static_assert(sizeof(PhxKernelInfo) == 52, "PhxKernelInfo");

static SceneArgs scene_args(const PhxTriangleVertices *tv, const PhxTriangleNormals *tn,
                            const PhxTriangleTangents *tt, const PhxTriangleUvs *tuv, const PhxBvhNode *nodes,
                            const uint32_t *tmat, uint32_t n_tris, const PhxMaterial *materials,
                            const PhxUvTransformBlock *uvt, const PhxTextureMeta *tex_meta, const uint8_t *atlas,
                            const PhxLight *lights, uint32_t n_lights, const float *lum, float total_lum,
                            const phx_float4 *center, float radius, const PhxTriangleVertices *etv,
                            const PhxTriangleNormals *etn, const PhxTriangleUvs *etuv, const uint32_t *etmat) {
    SceneArgs s;
    std::memset(&s, 0, sizeof(s));
    s.tv = tv;
    s.tn = tn;
    s.tt = tt;
    s.tuv = tuv;
    s.tmat = tmat;
    s.nodes = nodes;
    s.n_tris = n_tris;
    s.etv = etv;
    s.etn = etn;
    s.etuv = etuv;
    s.etmat = etmat;
    s.materials = materials;
    s.uvt = uvt;
    s.tex_meta = tex_meta;
    s.atlas = atlas;
    s.lights = lights;
    s.n_lights = n_lights;
    s.lum_pref_sum = lum;
    s.total_lum = total_lum;
    if (center)
        s.scene_center = *center;
    s.scene_radius = radius;
    apply_attached_tree(s); /* the pair tree, where phx_accel_attach registered one for this tree */
    return s;
}

} // namespace phx

using namespace phx;

extern "C" {

const char *phx_last_error(void) { return g_last_error.c_str(); }

const char *phx_version(void) { return "phosphor-mi355x 0.1 (gfx950)"; }

int phx_trace_rays(const PhxCameraParams *camera, uint32_t image_width, uint32_t image_height, uint32_t seed,
                   uint32_t direct_samples, const PhxTriangleVertices *tri_verts, const PhxTriangleNormals *tri_normals,
                   const PhxTriangleTangents *tri_tangents, const PhxTriangleUvs *tri_uvs, const PhxBvhNode *bvh_nodes,
                   const uint32_t *tri_mat_index, uint32_t n_triangles, const PhxMaterial *materials,
                   const PhxUvTransformBlock *uv_transforms, const PhxTextureMeta *tex_meta, const uint8_t *tex_atlas,
                   const PhxLight *lights, uint32_t n_lights, const float *lum_pref_sum, float total_lum,
                   const phx_float4 *scene_center, float scene_radius, const PhxTriangleVertices *etri_verts,
                   const PhxTriangleNormals *etri_normals, const PhxTriangleUvs *etri_uvs,
                   const uint32_t *etri_mat_index, const float *envmap_data, uint32_t envmap_width,
                   uint32_t envmap_height, uint32_t has_envmap, phx_float4 *hp_pos, phx_float4 *hp_normal,
                   phx_float4 *hp_throughput, phx_float4 *hp_base_color, phx_float4 *hp_direct,
                   phx_float4 *hp_emission, void *stream) {
    if (!camera || !tri_verts || !tri_normals || !tri_uvs || !bvh_nodes || !tri_mat_index || !materials ||
        !hp_pos || !hp_normal || !hp_throughput || !hp_base_color || !hp_direct || !hp_emission)
        return fail(PHX_ERR_INVALID_ARG, "phx_trace_rays: null pointer argument");
    if (n_triangles == 0)
        return fail(PHX_ERR_INVALID_ARG, "phx_trace_rays: empty scene");
    if (n_lights > 0 && (!lights || !lum_pref_sum))
        return fail(PHX_ERR_INVALID_ARG, "phx_trace_rays: lights without a luminance prefix sum");
    if (has_envmap && (!envmap_data || envmap_width == 0 || envmap_height == 0))
        return fail(PHX_ERR_INVALID_ARG, "phx_trace_rays: has_envmap without envmap data");
    if (direct_samples == 0)
        return fail(PHX_ERR_INVALID_ARG, "phx_trace_rays: direct_samples must be >= 1 (cmd_args.hpp:59)");
    if ((uint64_t)image_width * image_height > 0xFFFFFFFFull)
        return fail(PHX_ERR_TOO_LARGE, "phx_trace_rays: image too large");
    SceneArgs sc = scene_args(tri_verts, tri_normals, tri_tangents, tri_uvs, bvh_nodes, tri_mat_index, n_triangles,
                              materials, uv_transforms, tex_meta, tex_atlas, lights, n_lights, lum_pref_sum, total_lum,
                              scene_center, scene_radius, etri_verts, etri_normals, etri_uvs, etri_mat_index);
    sc.envmap = envmap_data;
    sc.env_w = envmap_width;
    sc.env_h = envmap_height;
    sc.has_env = has_envmap ? 1u : 0u;
    return hip_status(launch_trace(*camera, image_width, image_height, seed, direct_samples, sc, hp_pos, hp_normal,
                                   hp_throughput, hp_base_color, hp_direct, hp_emission, (hipStream_t)stream),
                      "phx_trace_rays");
}

size_t phx_trace_rays_wavefront_workspace_size(uint32_t image_width, uint32_t image_height, uint32_t direct_samples) {
    return trace_wavefront_workspace_bytes(image_width * image_height, direct_samples);
}

int phx_trace_rays_wavefront(const PhxCameraParams *camera, uint32_t image_width, uint32_t image_height,
                             uint32_t seed, uint32_t direct_samples, const PhxTriangleVertices *tri_verts,
                             const PhxTriangleNormals *tri_normals, const PhxTriangleTangents *tri_tangents,
                             const PhxTriangleUvs *tri_uvs, const PhxBvhNode *bvh_nodes,
                             const uint32_t *tri_mat_index, uint32_t n_triangles, const PhxMaterial *materials,
                             const PhxUvTransformBlock *uv_transforms, const PhxTextureMeta *tex_meta,
                             const uint8_t *tex_atlas, const PhxLight *lights, uint32_t n_lights,
                             const float *lum_pref_sum, float total_lum, const phx_float4 *scene_center,
                             float scene_radius, const PhxTriangleVertices *etri_verts,
                             const PhxTriangleNormals *etri_normals, const PhxTriangleUvs *etri_uvs,
                             const uint32_t *etri_mat_index, const float *envmap_data, uint32_t envmap_width,
                             uint32_t envmap_height, uint32_t has_envmap, phx_float4 *hp_pos, phx_float4 *hp_normal,
                             phx_float4 *hp_throughput, phx_float4 *hp_base_color, phx_float4 *hp_direct,
                             phx_float4 *hp_emission, uint32_t shadow_any_hit, void *workspace,
                             size_t workspace_bytes, void *stream) {
    if (!camera || !tri_verts || !tri_normals || !tri_uvs || !bvh_nodes || !tri_mat_index || !materials ||
        !hp_pos || !hp_normal || !hp_throughput || !hp_base_color || !hp_direct || !hp_emission)
        return fail(PHX_ERR_INVALID_ARG, "phx_trace_rays_wavefront: null pointer argument");
    if (n_triangles == 0)
        return fail(PHX_ERR_INVALID_ARG, "phx_trace_rays_wavefront: empty scene");
    if (n_lights > 0 && (!lights || !lum_pref_sum))
        return fail(PHX_ERR_INVALID_ARG, "phx_trace_rays_wavefront: lights without a luminance prefix sum");
    if (has_envmap && (!envmap_data || envmap_width == 0 || envmap_height == 0))
        return fail(PHX_ERR_INVALID_ARG, "phx_trace_rays_wavefront: has_envmap without envmap data");
    if (direct_samples == 0 || direct_samples > 128)
        return fail(PHX_ERR_INVALID_ARG, "phx_trace_rays_wavefront: direct_samples outside 1..128");
    uint64_t slots = (uint64_t)image_width * image_height * 5u * direct_samples;
    if (slots > 0xFFFFFFFFull)
        return fail(PHX_ERR_TOO_LARGE, "phx_trace_rays_wavefront: shadow queue exceeds 2^32 slots");
    if (!workspace || workspace_bytes < trace_wavefront_workspace_bytes(image_width * image_height, direct_samples))
        return fail(PHX_ERR_WORKSPACE, "phx_trace_rays_wavefront: workspace smaller than the size query");
    SceneArgs sc = scene_args(tri_verts, tri_normals, tri_tangents, tri_uvs, bvh_nodes, tri_mat_index, n_triangles,
                              materials, uv_transforms, tex_meta, tex_atlas, lights, n_lights, lum_pref_sum, total_lum,
                              scene_center, scene_radius, etri_verts, etri_normals, etri_uvs, etri_mat_index);
    sc.envmap = envmap_data;
    sc.env_w = envmap_width;
    sc.env_h = envmap_height;
    sc.has_env = has_envmap ? 1u : 0u;
    return hip_status(launch_trace_wavefront(*camera, image_width, image_height, seed, direct_samples, sc,
                                             shadow_any_hit, hp_pos, hp_normal, hp_throughput, hp_base_color,
                                             hp_direct, hp_emission, workspace, (hipStream_t)stream),
                      "phx_trace_rays_wavefront");
}

size_t phx_emit_photons_workspace_size(uint32_t photons_to_emit, uint32_t max_photon_bounces) {
    return emit_workspace_bytes(photons_to_emit, max_photon_bounces);
}

int phx_emit_photons(const PhxLight *lights, uint32_t n_lights, uint32_t max_photons, uint32_t offset,
                     uint32_t photons_to_emit, uint32_t max_photon_bounces, uint32_t seed, const PhxBvhNode *tree,
                     const PhxTriangleVertices *tri_verts, const PhxTriangleNormals *tri_normals,
                     const PhxTriangleTangents *tri_tangents, const PhxTriangleUvs *tri_uvs,
                     const uint32_t *tri_mat_index, uint32_t n_tris, const PhxTriangleVertices *etri_verts,
                     const PhxTriangleNormals *etri_normals, const PhxTriangleUvs *etri_uvs,
                     const uint32_t *etri_mat_index, const PhxMaterial *materials,
                     const PhxUvTransformBlock *uv_transforms, const PhxTextureMeta *tex_meta,
                     const uint8_t *tex_atlas, const float *lum_pref_sum, float total_lum,
                     const phx_float4 *scene_center, float scene_radius, phx_float4 *photon_pos,
                     uint32_t *photon_power, uint32_t *photon_dir, uint32_t *n_photons, void *workspace,
                     size_t workspace_bytes, void *stream) {
    if (!tree || !tri_verts || !tri_normals || !tri_uvs || !tri_mat_index || !materials || !photon_pos ||
        !photon_power || !photon_dir || !n_photons)
        return fail(PHX_ERR_INVALID_ARG, "phx_emit_photons: null pointer argument");
    if (n_tris == 0)
        return fail(PHX_ERR_INVALID_ARG, "phx_emit_photons: empty scene");
    if (max_photon_bounces == 0 || max_photon_bounces > 32)
        return fail(PHX_ERR_INVALID_ARG, "phx_emit_photons: max_photon_bounces outside 1..32 (cmd_args.hpp:49)");
    if (n_lights > 0 && (!lights || !lum_pref_sum))
        return fail(PHX_ERR_INVALID_ARG, "phx_emit_photons: lights without a luminance prefix sum");
    if ((uint64_t)offset + photons_to_emit > 0xFFFFFFFFull)
        return fail(PHX_ERR_TOO_LARGE, "phx_emit_photons: photon index overflows u32");
    size_t need = emit_workspace_bytes(photons_to_emit, max_photon_bounces);
    if (photons_to_emit > 0 && (!workspace || workspace_bytes < need))
        return fail(PHX_ERR_WORKSPACE, "phx_emit_photons: workspace smaller than phx_emit_photons_workspace_size");
    SceneArgs sc = scene_args(tri_verts, tri_normals, tri_tangents, tri_uvs, tree, tri_mat_index, n_tris, materials,
                              uv_transforms, tex_meta, tex_atlas, lights, n_lights, lum_pref_sum, total_lum,
                              scene_center, scene_radius, etri_verts, etri_normals, etri_uvs, etri_mat_index);
    return hip_status(launch_emit(sc, max_photons, offset, photons_to_emit, max_photon_bounces, seed, photon_pos,
                                  photon_power, photon_dir, n_photons, workspace, (hipStream_t)stream),
                      "phx_emit_photons");
}

int phx_hash_photons(const phx_float4 *photon_pos, uint32_t n_photons, const PhxPhotonHashData *data,
                     uint32_t *photon_hashes, uint32_t *cell_counts, void *stream) {
    if (!data || (n_photons > 0 && (!photon_pos || !photon_hashes || !cell_counts)))
        return fail(PHX_ERR_INVALID_ARG, "phx_hash_photons: null pointer argument");
    if (data->grid_res == 0 || data->grid_res > 255)
        return fail(PHX_ERR_INVALID_ARG, "phx_hash_photons: grid_res outside 1..255 (cmd_args.hpp:51)");
    return hip_status(launch_hash(photon_pos, n_photons, *data, photon_hashes, cell_counts, (hipStream_t)stream),
                      "phx_hash_photons");
}

size_t phx_scan_cells_workspace_size(uint32_t n_cells) { return scan_workspace_bytes(n_cells); }

int phx_scan_cells(const uint32_t *cell_counts, uint32_t n_cells, uint32_t *cell_start, uint32_t *cell_end,
                   void *workspace, size_t workspace_bytes, void *stream) {
    if (n_cells > 0 && (!cell_counts || !cell_start))
        return fail(PHX_ERR_INVALID_ARG, "phx_scan_cells: null pointer argument");
    if (!workspace || workspace_bytes < scan_workspace_bytes(n_cells))
        return fail(PHX_ERR_WORKSPACE, "phx_scan_cells: workspace smaller than phx_scan_cells_workspace_size");
    return hip_status(launch_scan(cell_counts, n_cells, cell_start, cell_end, nullptr, workspace, (hipStream_t)stream),
                      "phx_scan_cells");
}

int phx_scan_cells_noalloc(const uint32_t *cell_counts, uint32_t n_cells, uint32_t *cell_start, uint32_t *cell_end,
                           void *stream) {
    if (n_cells > 0 && (!cell_counts || !cell_start))
        return fail(PHX_ERR_INVALID_ARG, "phx_scan_cells_noalloc: null pointer argument");
    if (n_cells > PHX_SCAN_NOALLOC_MAX_CELLS || scan_workspace_bytes(n_cells) > NOALLOC_SCAN_SCRATCH_BYTES)
        return fail(PHX_ERR_TOO_LARGE, "phx_scan_cells_noalloc: n_cells above PHX_SCAN_NOALLOC_MAX_CELLS (" +
                                           std::to_string(PHX_SCAN_NOALLOC_MAX_CELLS) + ")");
    void *ws = nullptr;
    hipError_t e = noalloc_scan_scratch(&ws);
    if (e != hipSuccess)
        return hip_status(e, "phx_scan_cells_noalloc: scratch address");
    return hip_status(launch_scan(cell_counts, n_cells, cell_start, cell_end, nullptr, ws, (hipStream_t)stream),
                      "phx_scan_cells_noalloc");
}

int phx_scatter_photons(const phx_float4 *photon_pos_in, const uint32_t *photon_power_in,
                        const uint32_t *photon_dir_in, const uint32_t *photon_hashes, uint32_t *cell_end,
                        uint32_t n_photons, phx_float4 *out_photon_pos, uint32_t *out_photon_power,
                        uint32_t *out_photon_dir, void *stream) {
    if (n_photons > 0 && (!photon_pos_in || !photon_power_in || !photon_dir_in || !photon_hashes || !cell_end ||
                          !out_photon_pos || !out_photon_power || !out_photon_dir))
        return fail(PHX_ERR_INVALID_ARG, "phx_scatter_photons: null pointer argument");
    return hip_status(launch_scatter_atomic(photon_pos_in, photon_power_in, photon_dir_in, photon_hashes, cell_end,
                                            n_photons, out_photon_pos, out_photon_power, out_photon_dir,
                                            (hipStream_t)stream),
                      "phx_scatter_photons");
}

size_t phx_sort_photons_workspace_size(uint32_t n_photons, uint32_t n_cells) {
    return sort_workspace_bytes(n_photons, n_cells);
}

int phx_sort_photons(const phx_float4 *photon_pos_in, const uint32_t *photon_power_in, const uint32_t *photon_dir_in,
                     const uint32_t *photon_hashes, const uint32_t *cell_start, uint32_t *cell_end,
                     uint32_t n_photons, uint32_t n_cells, phx_float4 *out_photon_pos, uint32_t *out_photon_power,
                     uint32_t *out_photon_dir, void *workspace, size_t workspace_bytes, void *stream) {
    if ((n_photons > 0 && (!photon_pos_in || !photon_power_in || !photon_dir_in || !photon_hashes ||
                           !out_photon_pos || !out_photon_power || !out_photon_dir)) ||
        (n_cells > 0 && (!cell_start || !cell_end)))
        return fail(PHX_ERR_INVALID_ARG, "phx_sort_photons: null pointer argument");
    if (!workspace || workspace_bytes < sort_workspace_bytes(n_photons, n_cells))
        return fail(PHX_ERR_WORKSPACE, "phx_sort_photons: workspace smaller than phx_sort_photons_workspace_size");
    return hip_status(launch_sort(photon_pos_in, photon_power_in, photon_dir_in, photon_hashes, cell_start, cell_end,
                                  n_photons, n_cells, out_photon_pos, out_photon_power, out_photon_dir, workspace,
                                  (hipStream_t)stream),
                      "phx_sort_photons");
}

int phx_gather_flux(const phx_float4 *hp_pos, const phx_float4 *hp_normal, const phx_float4 *hp_throughput,
                    const phx_float4 *hp_base_color, const phx_float4 *hp_direct, const phx_float4 *hp_emission,
                    uint32_t n_pixels, const phx_float4 *photon_pos, const uint32_t *photon_power,
                    const uint32_t *photon_dir, uint32_t n_photons, const uint32_t *cell_start,
                    const uint32_t *cell_end, const PhxPhotonHashData *data, float sppm_alpha, phx_float4 *flux,
                    float *photon_weight, phx_float4 *radiance, void *stream) {
    (void)n_photons; /* unused by the reference kernel as well (gather_flux.cl:13) */
    if (!data || (n_pixels > 0 && (!hp_pos || !hp_normal || !hp_throughput || !hp_base_color || !hp_direct ||
                                   !hp_emission || !cell_start || !cell_end || !flux || !photon_weight || !radiance)))
        return fail(PHX_ERR_INVALID_ARG, "phx_gather_flux: null pointer argument");
    if (data->grid_res == 0 || data->grid_res > 255)
        return fail(PHX_ERR_INVALID_ARG, "phx_gather_flux: grid_res outside 1..255");
    return hip_status(launch_gather(hp_pos, hp_normal, hp_throughput, hp_base_color, hp_direct, hp_emission, n_pixels,
                                    photon_pos, photon_power, photon_dir, cell_start, cell_end, *data, sppm_alpha,
                                    flux, photon_weight, radiance, 0u, (hipStream_t)stream),
                      "phx_gather_flux");
}

static size_t octant_keys_bytes(uint32_t n) { return ((size_t)n * sizeof(uint32_t) + 255) & ~(size_t)255; }

size_t phx_sort_photons_by_octant_workspace_size(uint32_t n_photons, uint32_t n_cells) {
    return octant_keys_bytes(n_photons) + sort_workspace_bytes(n_photons, n_cells);
}

int phx_sort_photons_by_octant(const phx_float4 *photon_pos_in, const uint32_t *photon_power_in,
                               const uint32_t *photon_dir_in, uint32_t n_photons, const PhxPhotonHashData *data,
                               uint32_t n_cells, phx_float4 *out_photon_pos, uint32_t *out_photon_power,
                               uint32_t *out_photon_dir, phx_float4 *out_dir_power, uint32_t *cell_start,
                               uint32_t *cell_end, uint32_t *oct_start, void *workspace, size_t workspace_bytes,
                               void *stream) {
    if (!data || !cell_start || !cell_end || !oct_start ||
        (n_photons > 0 && (!photon_pos_in || !photon_power_in || !photon_dir_in || !out_photon_pos ||
                           !out_photon_power || !out_photon_dir)))
        return fail(PHX_ERR_INVALID_ARG, "phx_sort_photons_by_octant: null pointer argument");
    if (data->grid_res == 0 || data->grid_res > 255)
        return fail(PHX_ERR_INVALID_ARG, "phx_sort_photons_by_octant: grid_res outside 1..255 (cmd_args.hpp:51)");
    uint64_t g = data->grid_res;
    if ((uint64_t)n_cells < g * g * g + 1)
        return fail(PHX_ERR_INVALID_ARG, "phx_sort_photons_by_octant: n_cells below grid_res^3 + 1");
    if ((uint64_t)n_cells * 8u > 0xFFFFFFFFull)
        return fail(PHX_ERR_TOO_LARGE, "phx_sort_photons_by_octant: 8 * n_cells overflows u32");
    if (!workspace || workspace_bytes < phx_sort_photons_by_octant_workspace_size(n_photons, n_cells))
        return fail(PHX_ERR_WORKSPACE,
                    "phx_sort_photons_by_octant: workspace smaller than phx_sort_photons_by_octant_workspace_size");
    hipStream_t s = (hipStream_t)stream;
    uint32_t *keys = (uint32_t *)workspace;
    void *sort_ws = (char *)workspace + octant_keys_bytes(n_photons);
    hipError_t e = launch_hash(photon_pos_in, n_photons, *data, keys, nullptr, s, true);
    if (e == hipSuccess)
        e = launch_sort(photon_pos_in, photon_power_in, photon_dir_in, keys, nullptr, cell_end, n_photons, n_cells,
                        out_photon_pos, out_photon_power, out_photon_dir, sort_ws, s, nullptr, nullptr, cell_start,
                        oct_start);
    if (e == hipSuccess && out_dir_power && n_photons > 0)
        e = launch_decode_photons(out_photon_power, out_photon_dir, nullptr, n_photons, out_dir_power, s);
    return hip_status(e, "phx_sort_photons_by_octant");
}

int phx_gather_flux_octants(const phx_float4 *hp_pos, const phx_float4 *hp_normal, const phx_float4 *hp_throughput,
                            const phx_float4 *hp_base_color, const phx_float4 *hp_direct,
                            const phx_float4 *hp_emission, uint32_t image_width, uint32_t image_height,
                            const phx_float4 *photon_pos, const phx_float4 *photon_dir_power, uint32_t n_photons,
                            const uint32_t *cell_start, const uint32_t *cell_end, const uint32_t *oct_start,
                            const PhxPhotonHashData *data, float sppm_alpha, phx_float4 *flux, float *photon_weight,
                            phx_float4 *radiance, void *stream) {
    (void)n_photons;
    uint64_t n_pixels = (uint64_t)image_width * image_height;
    if (!data || (n_pixels > 0 && (!hp_pos || !hp_normal || !hp_throughput || !hp_base_color || !hp_direct ||
                                   !hp_emission || !cell_start || !cell_end || !photon_pos || !photon_dir_power ||
                                   !flux || !photon_weight || !radiance)))
        return fail(PHX_ERR_INVALID_ARG, "phx_gather_flux_octants: null pointer argument");
    if (data->grid_res == 0 || data->grid_res > 255)
        return fail(PHX_ERR_INVALID_ARG, "phx_gather_flux_octants: grid_res outside 1..255");
    if (n_pixels > 0xFFFFFFFFull)
        return fail(PHX_ERR_TOO_LARGE, "phx_gather_flux_octants: width * height overflows u32");
    if (n_pixels == 0)
        return PHX_OK;
    return hip_status(launch_gather(hp_pos, hp_normal, hp_throughput, hp_base_color, hp_direct, hp_emission,
                                    (uint32_t)n_pixels, photon_pos, nullptr, nullptr, cell_start, cell_end, *data,
                                    sppm_alpha, flux, photon_weight, radiance, image_width, (hipStream_t)stream,
                                    nullptr, nullptr, photon_dir_power, oct_start),
                      "phx_gather_flux_octants");
}

int phx_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return 0;
    return n;
}

int phx_kernel_info(PhxKernelInfo *out, uint32_t max) {
    if (max && !out)
        return -fail(PHX_ERR_INVALID_ARG, "phx_kernel_info: null output");
    if (phx_device_count() == 0)
        return -fail(PHX_ERR_NO_DEVICE, "phx_kernel_info: no device");
    uint32_t n = 0;
    bool ok = kernel_info_trace(out, max, n);
    ok = kernel_info_map(out, max, n) && ok;
    if (!ok)
        return -fail(PHX_ERR_NO_DEVICE, "phx_kernel_info: the runtime could not describe a kernel");
    return (int)n;
}

int phx_device_malloc(void **out, size_t bytes) {
    if (!out)
        return fail(PHX_ERR_INVALID_ARG, "phx_device_malloc: null output");
    *out = nullptr;
    return hip_status(hipMalloc(out, bytes ? bytes : 16), "phx_device_malloc");
}

int phx_device_free(void *p) { return hip_status(hipFree(p), "phx_device_free"); }

int phx_memcpy_htod(void *dst, const void *src, size_t bytes) {
    if (bytes == 0)
        return PHX_OK;
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "phx_memcpy_htod");
}

int phx_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    if (bytes == 0)
        return PHX_OK;
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "phx_memcpy_dtoh");
}

int phx_memset_device(void *dst, int value, size_t bytes) {
    if (bytes == 0)
        return PHX_OK;
    return hip_status(hipMemset(dst, value, bytes), "phx_memset_device");
}

int phx_device_copy(void *dst, const void *src, size_t bytes, void *stream) {
    if ((!dst || !src) && bytes)
        return fail(PHX_ERR_INVALID_ARG, "phx_device_copy: null pointer");
    if (bytes % 16 != 0 || ((uintptr_t)dst | (uintptr_t)src) % 16 != 0)
        return fail(PHX_ERR_INVALID_ARG, "phx_device_copy: bytes and pointers must be multiples of 16");
    return hip_status(launch_copy16((phx_float4 *)dst, (const phx_float4 *)src, bytes / 16, (hipStream_t)stream),
                      "phx_device_copy");
}

int phx_device_synchronize(void) { return hip_status(hipDeviceSynchronize(), "phx_device_synchronize"); }

} /* extern "C" */
