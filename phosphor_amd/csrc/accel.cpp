/*
 * accel.cpp — the library's SAH child-pair tree for callers that hold the
 * reference's median tree (bvh.cpp:18-84, uploaded by set_scene,
 * renderer.cpp:22-36) and trace through the kernel-level entry points.
 *
 * The reference's closest hit is a property of the triangles alone when its
 * depth-first search can never overflow its stack and every box encloses
 * what lies below it with the 1e-3 leaf margin (the argument is written out
 * at traverse_pairs in phx_device.h).  pair_tree_eligible checks exactly
 * those properties on the caller's host copy of its tree; build_pair_tree
 * builds the pair tree with the reference leaf indices as tie-break keys.
 * Both drivers (scene_upload.cpp) and phx_accel_attach use them.
 *
 * phx_accel_attach registers a pair tree under the caller's device tree
 * pointer.  phx_trace_rays, phx_trace_rays_wavefront and phx_emit_photons,
 * given that tree pointer and the triangle array it was attached with, then
 * trace the pair tree: same hit records, bit for bit.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "internal.h"
#include "phx_math.h"
#include "scene_buffers.h"

namespace phx {

namespace {
constexpr uint32_t INTERIOR = 0xFFFFFFFFu;
constexpr float LEAF_MARGIN = 1e-3f; /* bounding_box.h: leaf boxes grow by 1e-3 */
constexpr uint32_t MAX_SAFE_DEPTH = 22; /* deepest interior level the 24-entry stack never drops below */
constexpr float MAX_EXTENT = 1024.0f;
} // namespace

int tree_depth(const PhxBvhNode *nodes, size_t n_nodes, uint32_t &max_depth) {
    max_depth = 0;
    if (n_nodes == 0)
        return fail(PHX_ERR_INVALID_ARG, "BVH: empty tree");
    std::vector<std::pair<uint32_t, uint32_t>> todo{{0u, 0u}};
    while (!todo.empty()) {
        auto [i, dpt] = todo.back();
        todo.pop_back();
        if (phx_as_uint(nodes[i].min.w) != INTERIOR)
            continue;
        max_depth = std::max(max_depth, dpt);
        uint32_t l = i + 1, r = phx_as_uint(nodes[i].max.w);
        if (l >= n_nodes || r >= n_nodes || r <= i || dpt >= n_nodes)
            return fail(PHX_ERR_INVALID_ARG, "BVH: child index out of range");
        todo.push_back({l, dpt + 1});
        todo.push_back({r, dpt + 1});
    }
    return PHX_OK;
}

static bool box_holds(const PhxBvhNode &outer, const phx_float4 &mn, const phx_float4 &mx) {
    return outer.min.x <= mn.x && outer.min.y <= mn.y && outer.min.z <= mn.z && outer.max.x >= mx.x &&
           outer.max.y >= mx.y && outer.max.z >= mx.z;
}

bool pair_tree_eligible(const PhxBvhNode *nodes, size_t n_nodes, const PhxTriangleVertices *tv, uint32_t n,
                        std::string &why) {
    if (n == 0 || n > PHX_PAIR_MAX_TRIS) {
        why = "triangle count outside 1..PHX_PAIR_MAX_TRIS";
        return false;
    }
    if (n_nodes != 2 * (size_t)n - 1) {
        why = "not a tree of 2n-1 nodes with one triangle per leaf";
        return false;
    }
    uint32_t depth = 0;
    if (tree_depth(nodes, n_nodes, depth) != PHX_OK) {
        why = "child index out of range";
        return false;
    }
    if (depth > MAX_SAFE_DEPTH) {
        why = "deeper than 22 interior levels: the reference's stack-overflow rule can drop subtrees";
        return false;
    }
    const PhxBvhNode &rt = nodes[0];
    float extent = std::max({std::fabs(rt.min.x), std::fabs(rt.min.y), std::fabs(rt.min.z), std::fabs(rt.max.x),
                             std::fabs(rt.max.y), std::fabs(rt.max.z)});
    if (!(extent <= MAX_EXTENT)) {
        why = "coordinates beyond 1024";
        return false;
    }
    /* preorder (left child at i + 1, then the right child) visits the nodes
       in index order, every node once: node indices are the search's visiting
       order, which the tie-break keys stand for */
    std::vector<uint32_t> stack{0u};
    std::vector<uint8_t> seen_tri(n, 0);
    uint32_t next = 0;
    while (!stack.empty()) {
        uint32_t i = stack.back();
        stack.pop_back();
        if (i != next) {
            why = "nodes not stored in preorder";
            return false;
        }
        next++;
        const PhxBvhNode &nd = nodes[i];
        uint32_t t = phx_as_uint(nd.min.w);
        if (t == INTERIOR) {
            uint32_t l = i + 1, r = phx_as_uint(nd.max.w);
            if (!box_holds(nd, nodes[l].min, nodes[l].max) || !box_holds(nd, nodes[r].min, nodes[r].max)) {
                why = "an interior box does not hold its children's boxes";
                return false;
            }
            stack.push_back(r);
            stack.push_back(l);
            continue;
        }
        if (t >= n || seen_tri[t]) {
            why = "a triangle index out of range or in two leaves";
            return false;
        }
        seen_tri[t] = 1;
        const phx_float4 *v[3] = {&tv[t].v0, &tv[t].v1, &tv[t].v2};
        phx_float4 mn = *v[0], mx = *v[0];
        for (int k = 1; k < 3; k++) {
            mn.x = std::min(mn.x, v[k]->x), mn.y = std::min(mn.y, v[k]->y), mn.z = std::min(mn.z, v[k]->z);
            mx.x = std::max(mx.x, v[k]->x), mx.y = std::max(mx.y, v[k]->y), mx.z = std::max(mx.z, v[k]->z);
        }
        mn.x -= LEAF_MARGIN, mn.y -= LEAF_MARGIN, mn.z -= LEAF_MARGIN;
        mx.x += LEAF_MARGIN, mx.y += LEAF_MARGIN, mx.z += LEAF_MARGIN;
        if (!box_holds(nd, mn, mx)) {
            why = "a leaf box smaller than its triangle's box grown by 1e-3";
            return false;
        }
    }
    if (next != n_nodes) {
        why = "nodes unreachable from the root";
        return false;
    }
    return true;
}

bool build_pair_tree(const PhxBvhNode *nodes, size_t n_nodes, const PhxTriangleVertices *tv, uint32_t n,
                     uint32_t leaf_size, PairTreeHost &out) {
    std::vector<uint32_t> leaf_of(n, 0u);
    for (uint32_t i = 0; i < n_nodes; i++) {
        uint32_t t = phx_as_uint(nodes[i].min.w);
        if (t != INTERIOR && t < n)
            leaf_of[t] = i;
    }
    std::vector<PhxBvhNode> sah(2 * (size_t)n - 1);
    out.order.resize(n);
    uint32_t got = phx_build_bvh_sah(tv, n, MAX_SAFE_DEPTH, leaf_size, sah.data(), out.order.data());
    if (!got)
        return false;
    out.pairs.resize(std::max<size_t>(1, got / 2));
    out.root_ref = 0;
    if (phx_build_bvh_pairs(sah.data(), got, out.pairs.data(), &out.root_ref) != (int)(got / 2))
        return false;
    /* triangles and their tie-break keys in the tree's order */
    out.key.resize(n);
    out.tvo.resize(n);
    for (uint32_t i = 0; i < n; i++) {
        out.key[i] = leaf_of[out.order[i]];
        out.tvo[i] = tv[out.order[i]];
    }
    return true;
}

hipError_t upload_pair_tree(const PairTreeHost &h, GeometryBuffers &g, hipStream_t s) {
    uint32_t n = (uint32_t)h.order.size();
    hipError_t e;
    if ((e = copy_to_dev_buf(g.d_bvh_pairs, h.pairs.data(), h.pairs.size(), sizeof(PhxBvhPair), s)) != hipSuccess)
        return e;
    if ((e = copy_to_dev_buf(g.d_tri_key, h.key.data(), n, sizeof(uint32_t), s)) != hipSuccess)
        return e;
    if ((e = copy_to_dev_buf(g.d_tri_order, h.order.data(), n, sizeof(uint32_t), s)) != hipSuccess)
        return e;
    if ((e = copy_to_dev_buf(g.d_tri_verts_ord, h.tvo.data(), n, sizeof(PhxTriangleVertices), s)) != hipSuccess)
        return e;
    g.root_ref = h.root_ref;
    return hipSuccess;
}

/* ---- the registry of attached pair trees ---- */
namespace {
struct Attached {
    int device = 0;
    const PhxTriangleVertices *tri_verts = nullptr;
    uint32_t n_tris = 0;
    GeometryBuffers g;
};
std::mutex g_mu;
std::map<std::pair<int, const void *>, std::unique_ptr<Attached>> g_attached;
} // namespace

void apply_attached_tree(SceneArgs &s) {
    int dev = 0;
    if (!s.nodes || hipGetDevice(&dev) != hipSuccess)
        return;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_attached.find({dev, (const void *)s.nodes});
    if (it == g_attached.end())
        return;
    const Attached &a = *it->second;
    if (a.tri_verts != s.tv || a.n_tris != s.n_tris)
        return;
    s.near_first = 1;
    s.root_ref = a.g.root_ref;
    s.pairs = a.g.d_bvh_pairs.as<PhxBvhPair>();
    s.tri_key = a.g.d_tri_key.as<uint32_t>();
    s.tri_order = a.g.d_tri_order.as<uint32_t>();
    s.tvo = a.g.d_tri_verts_ord.as<PhxTriangleVertices>();
}

} // namespace phx

using namespace phx;

extern "C" {

int phx_accel_eligible(const PhxBvhNode *host_tree, uint32_t n_nodes, const PhxTriangleVertices *host_tri_verts,
                       uint32_t n_triangles) {
    if (!host_tree || !host_tri_verts)
        return fail(PHX_ERR_INVALID_ARG, "phx_accel_eligible: null pointer argument");
    std::string why;
    bool ok = pair_tree_eligible(host_tree, n_nodes, host_tri_verts, n_triangles, why);
    if (!ok)
        set_error("phx_accel_eligible: " + why);
    return ok ? 1 : 0;
}

int phx_accel_attach(const PhxBvhNode *device_tree, const PhxTriangleVertices *device_tri_verts,
                     const PhxBvhNode *host_tree, uint32_t n_nodes, const PhxTriangleVertices *host_tri_verts,
                     uint32_t n_triangles, uint32_t leaf_size, uint32_t *accelerated) {
    if (!device_tree || !device_tri_verts || !host_tree || !host_tri_verts || !accelerated)
        return fail(PHX_ERR_INVALID_ARG, "phx_accel_attach: null pointer argument");
    if (leaf_size > PHX_PAIR_MAX_LEAF)
        return fail(PHX_ERR_INVALID_ARG, "phx_accel_attach: leaf_size beyond PHX_PAIR_MAX_LEAF");
    *accelerated = 0;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return hip_status(e, "phx_accel_attach");
    phx_accel_detach(device_tree);
    std::string why;
    if (!pair_tree_eligible(host_tree, n_nodes, host_tri_verts, n_triangles, why)) {
        set_error("phx_accel_attach: reference tree kept: " + why);
        return PHX_OK;
    }
    /* the device arrays must hold what the host copies describe */
    std::vector<PhxBvhNode> dn(n_nodes);
    std::vector<PhxTriangleVertices> dv(n_triangles);
    if ((e = hipMemcpy(dn.data(), device_tree, n_nodes * sizeof(PhxBvhNode), hipMemcpyDeviceToHost)) != hipSuccess ||
        (e = hipMemcpy(dv.data(), device_tri_verts, (size_t)n_triangles * sizeof(PhxTriangleVertices),
                       hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_status(e, "phx_accel_attach");
    if (std::memcmp(dn.data(), host_tree, n_nodes * sizeof(PhxBvhNode)) != 0 ||
        std::memcmp(dv.data(), host_tri_verts, (size_t)n_triangles * sizeof(PhxTriangleVertices)) != 0)
        return fail(PHX_ERR_INVALID_ARG, "phx_accel_attach: device tree or triangles differ from the host copies");
    PairTreeHost h;
    if (!build_pair_tree(host_tree, n_nodes, host_tri_verts, n_triangles, leaf_size ? leaf_size : 3u, h)) {
        set_error("phx_accel_attach: reference tree kept: pair tree build failed");
        return PHX_OK;
    }
    auto a = std::make_unique<Attached>();
    a->device = dev;
    a->tri_verts = device_tri_verts;
    a->n_tris = n_triangles;
    if ((e = upload_pair_tree(h, a->g, nullptr)) != hipSuccess || (e = hipDeviceSynchronize()) != hipSuccess)
        return hip_status(e, "phx_accel_attach");
    std::lock_guard<std::mutex> lk(g_mu);
    g_attached[{dev, (const void *)device_tree}] = std::move(a);
    *accelerated = 1;
    return PHX_OK;
}

int phx_accel_detach(const PhxBvhNode *device_tree) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return PHX_OK;
    std::unique_ptr<Attached> gone;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_attached.find({dev, (const void *)device_tree});
        if (it == g_attached.end())
            return PHX_OK;
        gone = std::move(it->second);
        g_attached.erase(it);
    }
    /* no launch may still read the buffers freed when `gone` ends */
    return hip_status(hipDeviceSynchronize(), "phx_accel_detach");
}

} // extern "C"
