/*
 * scene_upload.cpp — upload of one scene to the device (renderer.cpp:22-36,
 * rendering.cpp:15-30, geometry.cpp, shading.cpp) for both pass drivers.
 */
#include <algorithm>
#include <cmath>
#include <string>
#include <utility>
#include <vector>

#include "phx_math.h"
#include "scene_buffers.h"

namespace phx {

hipError_t copy_to_dev_buf(DevBuf &b, const void *src, size_t count, size_t item, hipStream_t s) {
    hipError_t e = b.alloc(count ? count * item : item);
    if (e != hipSuccess)
        return e;
    if (count && src)
        return hipMemcpyAsync(b.p, src, count * item, hipMemcpyHostToDevice, s);
    return hipMemsetAsync(b.p, 0, b.bytes, s);
}

SceneArgs SceneBuffers::args() const {
    const GeometryBuffers &g = geometry;
    const ShadingBuffers &s = shading;
    SceneArgs a;
    std::memset(&a, 0, sizeof(a));
    a.tv = g.d_tri_verts.as<PhxTriangleVertices>();
    a.tn = g.d_tri_normals.as<PhxTriangleNormals>();
    a.tt = g.d_tri_tangents.as<PhxTriangleTangents>();
    a.tuv = g.d_tri_uvs.as<PhxTriangleUvs>();
    a.tmat = g.d_tri_mat_index.as<uint32_t>();
    a.nodes = g.d_bvh_nodes.as<PhxBvhNode>();
    a.n_tris = g.n_triangles;
    a.etv = g.d_etri_verts.as<PhxTriangleVertices>();
    a.etn = g.d_etri_normals.as<PhxTriangleNormals>();
    a.etuv = g.d_etri_uvs.as<PhxTriangleUvs>();
    a.etmat = g.d_etri_mat_index.as<uint32_t>();
    a.materials = s.d_materials.as<PhxMaterial>();
    a.uvt = s.d_uv_transforms.as<PhxUvTransformBlock>();
    a.tex_meta = s.d_tex_meta.as<PhxTextureMeta>();
    a.atlas = s.d_tex_atlas.as<uint8_t>();
    a.lights = s.d_lights.as<PhxLight>();
    a.n_lights = s.n_lights;
    a.lum_pref_sum = s.d_lum_pref_sum.as<float>();
    a.total_lum = s.total_lum;
    a.scene_center = s.scene_center;
    a.scene_radius = s.scene_radius;
    a.envmap = s.d_envmap_data.as<float>();
    a.env_w = s.envmap_width;
    a.env_h = s.envmap_height;
    a.has_env = s.has_envmap;
    a.near_first = shallow_bvh;
    a.root_ref = g.root_ref;
    a.pairs = g.d_bvh_pairs.as<PhxBvhPair>();
    a.tri_key = g.d_tri_key.as<uint32_t>();
    a.tri_order = g.d_tri_order.as<uint32_t>();
    a.tvo = g.d_tri_verts_ord.as<PhxTriangleVertices>();
    return a;
}

int check_scene_desc(const PhxSceneDesc &sc, const char *who) {
    if (sc.n_triangles == 0 || !sc.tri_verts || !sc.tri_normals || !sc.tri_tangents || !sc.tri_uvs ||
        !sc.tri_mat_index || !sc.materials || sc.n_materials == 0 || !sc.camera)
        return fail(PHX_ERR_INVALID_ARG, std::string(who) + ": empty or incomplete scene (main.cpp:40-47)");
    if (sc.n_triangles >= 0x80000000u)
        return fail(PHX_ERR_INVALID_ARG, std::string(who) + ": at most 2^31 - 1 triangles");
    return PHX_OK;
}

#define PHX_TRY(expr, what)                                                                                            \
    do {                                                                                                               \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess)                                                                                          \
            return hip_status(e_, what);                                                                               \
    } while (0)

int upload_scene(const PhxSceneDesc &sc, hipStream_t s, bool reference_bvh, uint32_t leaf_size, SceneBuffers &out) {
    uint32_t n = sc.n_triangles, ne = sc.n_e_triangles;

    /* Bvh bvh(scene.triangles) */
    std::vector<PhxBvhNode> nodes;
    if (sc.bvh_nodes && sc.n_bvh_nodes) {
        nodes.assign(sc.bvh_nodes, sc.bvh_nodes + sc.n_bvh_nodes);
    } else {
        nodes.resize(2 * (size_t)n - 1);
        uint32_t got = phx_build_bvh_median(sc.tri_verts, n, nodes.data());
        if (got != nodes.size())
            return fail(PHX_ERR_INVALID_ARG, "set_scene: BVH build failed");
    }
    out.root = nodes[0];
    if (leaf_size == 0)
        leaf_size = 3; /* fastest on the caustic and atrium workloads */
    if (leaf_size > PHX_PAIR_MAX_LEAF)
        return fail(PHX_ERR_INVALID_ARG, "set_scene: bvh_leaf_size beyond PHX_PAIR_MAX_LEAF");
    bool transmissive = false;
    for (uint32_t i = 0; i < sc.n_materials; i++)
        transmissive = transmissive || !(sc.materials[i].transmission < 1e-5f);
    uint32_t max_depth = 0;
    if (tree_depth(nodes.data(), nodes.size(), max_depth) != PHX_OK)
        return fail(PHX_ERR_INVALID_ARG, "set_scene: BVH child index out of range");
    out.shadow_any_hit = (!transmissive && max_depth <= 22) ? 1u : 0u;
    out.shallow_bvh = 0;
    /* Where the reference tree never overflows its search stack, its closest
       hit is the least t over all triangles, equal t going to the larger leaf
       index: a property of the triangles, not of the tree.  Trace then on an
       SAH tree of the same leaves (far fewer nodes per ray) carrying those
       leaf indices as tie-break keys -- within |coordinate| <= 1024, where
       box-test rounding stays far inside the 1e-3 leaf-box margin that makes
       the search order-free.  A tree passed in is searched as given. */
    std::string why;
    if (!sc.bvh_nodes && !reference_bvh && pair_tree_eligible(nodes.data(), nodes.size(), sc.tri_verts, n, why)) {
        PairTreeHost h; /* host copies: live until the stream has read them */
        if (build_pair_tree(nodes.data(), nodes.size(), sc.tri_verts, n, leaf_size, h)) {
            PHX_TRY(upload_pair_tree(h, out.geometry, s), "upload");
            PHX_TRY(hipStreamSynchronize(s), "upload");
            out.shallow_bvh = 1;
        }
    }

    GeometryBuffers &g = out.geometry;
    g.n_triangles = n;
    g.n_e_triangles = ne;
    PHX_TRY(copy_to_dev_buf(g.d_tri_verts, sc.tri_verts, n, sizeof(PhxTriangleVertices), s), "upload");
    PHX_TRY(copy_to_dev_buf(g.d_tri_normals, sc.tri_normals, n, sizeof(PhxTriangleNormals), s), "upload");
    PHX_TRY(copy_to_dev_buf(g.d_tri_tangents, sc.tri_tangents, n, sizeof(PhxTriangleTangents), s), "upload");
    PHX_TRY(copy_to_dev_buf(g.d_tri_uvs, sc.tri_uvs, n, sizeof(PhxTriangleUvs), s), "upload");
    PHX_TRY(copy_to_dev_buf(g.d_tri_mat_index, sc.tri_mat_index, n, sizeof(uint32_t), s), "upload");
    PHX_TRY(copy_to_dev_buf(g.d_etri_verts, sc.etri_verts, ne, sizeof(PhxTriangleVertices), s), "upload");
    PHX_TRY(copy_to_dev_buf(g.d_etri_normals, sc.etri_normals, ne, sizeof(PhxTriangleNormals), s), "upload");
    PHX_TRY(copy_to_dev_buf(g.d_etri_uvs, sc.etri_uvs, ne, sizeof(PhxTriangleUvs), s), "upload");
    PHX_TRY(copy_to_dev_buf(g.d_etri_mat_index, sc.etri_mat_index, ne, sizeof(uint32_t), s), "upload");
    PHX_TRY(copy_to_dev_buf(g.d_bvh_nodes, nodes.data(), nodes.size(), sizeof(PhxBvhNode), s), "upload");

    ShadingBuffers &sh = out.shading;
    sh.n_materials = sc.n_materials;
    sh.n_lights = sc.n_lights;
    sh.n_textures = sc.n_textures;
    PHX_TRY(copy_to_dev_buf(sh.d_materials, sc.materials, sc.n_materials, sizeof(PhxMaterial), s), "upload");
    PHX_TRY(copy_to_dev_buf(sh.d_uv_transforms, sc.uv_transforms, sc.n_materials, sizeof(PhxUvTransformBlock), s),
            "upload");
    PHX_TRY(copy_to_dev_buf(sh.d_lights, sc.lights, sc.n_lights, sizeof(PhxLight), s), "upload");
    /* scene_data.cpp:14-23 luminance prefix sum; shading.cpp:21-23 total */
    std::vector<float> lum;
    float running = 0.0f;
    for (uint32_t i = 0; i < sc.n_lights; i++) {
        const PhxLight &l = sc.lights[i];
        float v = l.power.x * 0.2126f + l.power.y * 0.7152f + l.power.z * 0.0722f;
        running += std::max(v, 0.0f);
        lum.push_back(running);
    }
    sh.total_lum = lum.empty() ? 0.0f : lum.back();
    PHX_TRY(copy_to_dev_buf(sh.d_lum_pref_sum, lum.data(), lum.size(), sizeof(float), s), "upload");
    /* shading.cpp:25-43 texture atlas (never empty) */
    if (sc.n_textures) {
        PHX_TRY(copy_to_dev_buf(sh.d_tex_meta, sc.tex_meta, sc.n_textures, sizeof(PhxTextureMeta), s), "upload");
        PHX_TRY(copy_to_dev_buf(sh.d_tex_atlas, sc.tex_atlas, (size_t)sc.tex_atlas_bytes, 1, s), "upload");
    } else {
        PhxTextureMeta m{};
        uint8_t zero = 0;
        PHX_TRY(copy_to_dev_buf(sh.d_tex_meta, &m, 1, sizeof(PhxTextureMeta), s), "upload");
        PHX_TRY(copy_to_dev_buf(sh.d_tex_atlas, &zero, 1, 1, s), "upload");
        PHX_TRY(hipStreamSynchronize(s), "upload");
    }
    /* shading.cpp:45-58 bounding sphere of the root bbox */
    const PhxBvhNode &root = out.root;
    float cx = 0.5f * (root.min.x + root.max.x), cy = 0.5f * (root.min.y + root.max.y),
          cz = 0.5f * (root.min.z + root.max.z);
    sh.scene_center = phx_float4{cx, cy, cz, 0.0f};
    float radius = 0.0f;
    for (int c = 0; c < 8; c++) {
        float dx = ((c & 1) ? root.max.x : root.min.x) - cx;
        float dy = ((c & 2) ? root.max.y : root.min.y) - cy;
        float dz = ((c & 4) ? root.max.z : root.min.z) - cz;
        radius = std::max(radius, phx_sqrtf(dx * dx + dy * dy + dz * dz));
    }
    sh.scene_radius = radius;
    /* shading.cpp:60-71 envmap */
    if (sc.envmap && sc.envmap_width && sc.envmap_height) {
        sh.has_envmap = 1;
        sh.envmap_width = sc.envmap_width;
        sh.envmap_height = sc.envmap_height;
        PHX_TRY(copy_to_dev_buf(sh.d_envmap_data, sc.envmap, 3ull * sc.envmap_width * sc.envmap_height, sizeof(float),
                                s),
                "upload");
    } else {
        sh.has_envmap = 0;
        sh.envmap_width = sh.envmap_height = 0;
        PHX_TRY(copy_to_dev_buf(sh.d_envmap_data, nullptr, 0, sizeof(float), s), "upload");
    }
    /* `nodes`, `key`, `order`, `pairs`, `tv_ord` and `lum` are local: their uploads finish before they go */
    PHX_TRY(hipStreamSynchronize(s), "upload");
    return PHX_OK;
}

} // namespace phx
