/*
 * bvh.cpp — host BVH build in the reference's node format.
 *
 * src/renderer/bvh.cpp:18-84: every node's triangle range is sorted by the
 * (BBOX_EPS-expanded) triangle bbox minimum along the longest axis of the
 * node bbox and split at the median; leaves hold one triangle; nodes are
 * flattened depth-first with the left child at index + 1 and the right child
 * index in max.w (bounding_box.h:48-90 for the bbox helpers).
 *
 * The per-node sorts act on disjoint index ranges, so the ranges of the two
 * children are processed concurrently (std::async down to a size cutoff); the
 * permutation is the same as the sequential recursion because each range is
 * sorted by std::sort with the same comparator on the same input sequence.
 */
#include <algorithm>
#include <cstring>
#include <future>
#include <numeric>
#include <vector>

#include "../../include/phosphor_hip.h"
#include "phx_math.h"

namespace {

constexpr float BBOX_EPS = 1e-3f;
constexpr float BIG = 1e9f; /* constants.h INF */

struct Box {
    float lo[3], hi[3];
};

Box triangle_box(const PhxTriangleVertices &t) {
    Box b;
    for (int a = 0; a < 3; a++) {
        b.lo[a] = BIG;
        b.hi[a] = -BIG;
    }
    const phx_float4 *v[3] = {&t.v0, &t.v1, &t.v2};
    for (const phx_float4 *p : v) {
        const float c[3] = {p->x, p->y, p->z};
        for (int a = 0; a < 3; a++) {
            b.lo[a] = std::min(b.lo[a], c[a]);
            b.hi[a] = std::max(b.hi[a], c[a]);
        }
    }
    for (int a = 0; a < 3; a++) {
        b.lo[a] -= BBOX_EPS;
        b.hi[a] += BBOX_EPS;
    }
    return b;
}

struct Node {
    Box box;
    int tri;                 /* -1 for interior */
    unsigned left, right;    /* child node ids */
};

struct Builder {
    const std::vector<Box> &boxes;
    std::vector<unsigned> &order;
    explicit Builder(const std::vector<Box> &b, std::vector<unsigned> &o) : boxes(b), order(o) {}

    /* Builds the subtree of order[start, end) into its own node list
     * (preorder), returning it; children are built concurrently. */
    std::vector<Node> build(unsigned start, unsigned end, int depth) {
        std::vector<Node> out;
        if (end == start + 1) {
            out.push_back(Node{boxes[order[start]], (int)order[start], 0u, 0u});
            return out;
        }
        Box bb = boxes[order[start]];
        for (unsigned i = start + 1; i < end; i++) {
            const Box &b = boxes[order[i]];
            for (int a = 0; a < 3; a++) {
                bb.lo[a] = std::min(bb.lo[a], b.lo[a]);
                bb.hi[a] = std::max(bb.hi[a], b.hi[a]);
            }
        }
        float ex = bb.hi[0] - bb.lo[0], ey = bb.hi[1] - bb.lo[1], ez = bb.hi[2] - bb.lo[2];
        int axis = 2;
        if (ex > ey) {
            if (ex > ez)
                axis = 0;
        } else if (ey > ez) {
            axis = 1;
        }
        const std::vector<Box> &bx = boxes;
        std::sort(order.begin() + start, order.begin() + end,
                  [&bx, axis](unsigned i, unsigned j) { return bx[i].lo[axis] < bx[j].lo[axis]; });
        unsigned mid = start + (end - start) / 2;
        std::vector<Node> left, right;
        if (depth < 6 && end - start > 4096) {
            auto fut = std::async(std::launch::async, [this, start, mid, depth] { return build(start, mid, depth + 1); });
            right = build(mid, end, depth + 1);
            left = fut.get();
        } else {
            left = build(start, mid, depth + 1);
            right = build(mid, end, depth + 1);
        }
        out.reserve(1 + left.size() + right.size());
        out.push_back(Node{bb, -1, 1u, (unsigned)(1 + left.size())});
        unsigned lofs = 1, rofs = (unsigned)(1 + left.size());
        for (Node n : left) {
            if (n.tri < 0) {
                n.left += lofs;
                n.right += lofs;
            }
            out.push_back(n);
        }
        for (Node n : right) {
            if (n.tri < 0) {
                n.left += rofs;
                n.right += rofs;
            }
            out.push_back(n);
        }
        return out;
    }
};

} // namespace

extern "C" uint32_t phx_build_bvh_median(const PhxTriangleVertices *tri_verts, uint32_t n_tris, PhxBvhNode *out_nodes) {
    if (!tri_verts || !out_nodes || n_tris == 0)
        return 0;
    std::vector<Box> boxes(n_tris);
    for (uint32_t i = 0; i < n_tris; i++)
        boxes[i] = triangle_box(tri_verts[i]);
    std::vector<unsigned> order(n_tris);
    std::iota(order.begin(), order.end(), 0u);
    Builder b(boxes, order);
    std::vector<Node> nodes = b.build(0, n_tris, 0);
    /* preorder with left = id + 1 is exactly the reference's flatten order */
    for (size_t i = 0; i < nodes.size(); i++) {
        const Node &n = nodes[i];
        PhxBvhNode &o = out_nodes[i];
        o.min.x = n.box.lo[0];
        o.min.y = n.box.lo[1];
        o.min.z = n.box.lo[2];
        o.max.x = n.box.hi[0];
        o.max.y = n.box.hi[1];
        o.max.z = n.box.hi[2];
        o.min.w = phx_as_float((uint32_t)n.tri);
        o.max.w = phx_as_float(n.tri < 0 ? n.right : 0u);
    }
    return (uint32_t)nodes.size();
}

extern "C" int phx_build_photon_hash_info(const PhxBvhNode *root, uint32_t grid_res, PhxPhotonHashData *out) {
    /* photon_hash.h:41-50 */
    if (!root || !out || grid_res == 0)
        return PHX_ERR_INVALID_ARG;
    PhxPhotonHashData d;
    std::memset(&d, 0, sizeof(d));
    d.grid_res = grid_res;
    d.origin.x = root->min.x;
    d.origin.y = root->min.y;
    d.origin.z = root->min.z;
    const float min_cell = 0.001f; /* constants.h MIN_CELL_SIZE */
    d.cell_sizes.x = std::max((root->max.x - root->min.x) / (float)grid_res, min_cell);
    d.cell_sizes.y = std::max((root->max.y - root->min.y) / (float)grid_res, min_cell);
    d.cell_sizes.z = std::max((root->max.z - root->min.z) / (float)grid_res, min_cell);
    *out = d;
    return PHX_OK;
}
