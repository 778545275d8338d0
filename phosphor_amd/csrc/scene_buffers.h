/*
 * scene_buffers.h — device buffers of one scene (include/buffers/geometry.hpp,
 * include/buffers/shading.hpp) and their upload, shared by the one-scene pass
 * driver (renderer.cpp) and the multi-scene driver (batch.cpp).
 */
#ifndef PHX_SCENE_BUFFERS_H
#define PHX_SCENE_BUFFERS_H

#include <cstring>

#include "internal.h"

namespace phx {

/* cl::Buffer stand-in: one device allocation, freed on destruction */
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p)
            (void)hipFree(p);
    }
    hipError_t alloc(size_t n) {
        if (p) {
            (void)hipFree(p);
            p = nullptr;
        }
        bytes = n < 16 ? 16 : n;
        return hipMalloc(&p, bytes);
    }
    template <class T> T *as() const { return (T *)p; }
    void swap(DevBuf &o) {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
    }
};

/* rendering.cpp:3-13: zero items still allocate one item */
hipError_t copy_to_dev_buf(DevBuf &b, const void *src, size_t count, size_t item, hipStream_t s);

struct GeometryBuffers { /* include/buffers/geometry.hpp */
    DevBuf d_tri_verts, d_tri_normals, d_tri_tangents, d_tri_uvs, d_tri_mat_index;
    uint32_t n_triangles = 0;
    DevBuf d_etri_verts, d_etri_normals, d_etri_uvs, d_etri_mat_index;
    uint32_t n_e_triangles = 0;
    DevBuf d_bvh_nodes;
    /* the SAH tree's pairs; its triangle order, the triangles' vertices and
       reference leaf indices in that order (empty when the reference tree is
       searched) */
    DevBuf d_bvh_pairs, d_tri_key, d_tri_order, d_tri_verts_ord;
    uint32_t root_ref = 0;
};

/* uploads a pair tree (accel.cpp build_pair_tree) into g's pair-tree
 * buffers and sets g.root_ref */
hipError_t upload_pair_tree(const PairTreeHost &h, GeometryBuffers &g, hipStream_t s);

struct ShadingBuffers { /* include/buffers/shading.hpp */
    DevBuf d_materials, d_uv_transforms;
    uint32_t n_materials = 0;
    DevBuf d_lights, d_lum_pref_sum;
    uint32_t n_lights = 0;
    float total_lum = 0.0f;
    DevBuf d_tex_atlas, d_tex_meta;
    uint32_t n_textures = 0;
    phx_float4 scene_center{};
    float scene_radius = 0.0f;
    DevBuf d_envmap_data;
    uint32_t envmap_width = 0, envmap_height = 0, has_envmap = 0;
};

/* One uploaded scene and what the host derives from it. */
struct SceneBuffers {
    GeometryBuffers geometry;
    ShadingBuffers shading;
    PhxBvhNode root{};
    /* any-hit shadow rays are exact for this scene: every material is opaque
       to shadow rays (light_sampling.h:316) and no interior BVH node is deeper
       than 22, so the traversal stack never overflows (hit.h:163) */
    uint32_t shadow_any_hit = 0;
    /* the uploaded tree is the library's SAH tree (the reference tree is no
       deeper than 22 interior levels): closest-hit searches visit nodes
       near-first with leaf-key tie breaks (same hits, SceneArgs::near_first) */
    uint32_t shallow_bvh = 0;

    SceneArgs args() const;
};

/* Validates and uploads a scene: BVH (given or median-split, bvh.cpp:18-84),
 * SoA triangle arrays, materials, lights with their luminance prefix sum
 * (scene_data.cpp:14-23), texture atlas, bounding sphere (shading.cpp:45-58)
 * and envmap (shading.cpp:60-71).  Asynchronous on `s` except for the BVH
 * build; the caller synchronises before freeing host memory. */
int upload_scene(const PhxSceneDesc &sc, hipStream_t s, bool reference_bvh, uint32_t leaf_size, SceneBuffers &out);

/* Scene validity checks of main.cpp:40-47 shared by the drivers. */
int check_scene_desc(const PhxSceneDesc &sc, const char *who);

} // namespace phx

#endif /* PHX_SCENE_BUFFERS_H */
