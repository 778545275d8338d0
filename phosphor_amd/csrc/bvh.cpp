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
 *
 * phx_build_bvh_sah is this library's own tree for tracing: the same leaves
 * (one expanded triangle box each) grouped by a binned surface-area
 * heuristic, within a depth limit, in the same node format.
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

/* Binned surface-area-heuristic build over the same expanded triangle boxes.
 * A range of at most `max_leaf` triangles becomes a leaf whose box is the
 * union of its triangles' boxes, so a closest-hit search over this tree meets
 * every candidate triangle the median tree's search meets; only the boxes
 * (and so the number of nodes visited) change.  Every split keeps the tree
 * within `max_depth` interior levels: a candidate split is taken only if both
 * sides can still be built balanced within the limit, otherwise the range is
 * split at its object median, which always can. */
struct SahNode {
    Box box;
    int first;      /* -1 for interior; a leaf's first position in `order` */
    unsigned count; /* a leaf's triangle count */
    unsigned right; /* right child id (interior) */
    unsigned hint;  /* split axis | 4 when the left child lies above the right */
};

unsigned ceil_log2(unsigned c) {
    unsigned k = 0;
    while ((1u << k) < c)
        k++;
    return k;
}

float half_area(const Box &b) {
    float e[3];
    for (int a = 0; a < 3; a++)
        e[a] = std::max(0.0f, b.hi[a] - b.lo[a]);
    return e[0] * e[1] + e[1] * e[2] + e[2] * e[0];
}

void grow(Box &b, const Box &c) {
    for (int a = 0; a < 3; a++) {
        b.lo[a] = std::min(b.lo[a], c.lo[a]);
        b.hi[a] = std::max(b.hi[a], c.hi[a]);
    }
}

Box empty_box() {
    Box b;
    for (int a = 0; a < 3; a++) {
        b.lo[a] = BIG;
        b.hi[a] = -BIG;
    }
    return b;
}

unsigned ceil_div(unsigned a, unsigned b) { return (a + b - 1) / b; }

struct SahBuilder {
    static constexpr int BINS = 32;
    const std::vector<Box> &boxes;
    const std::vector<float> *cen; /* [3] centroid per triangle */
    std::vector<unsigned> &order;
    unsigned max_depth;
    unsigned max_leaf;

    /* interior levels a range of c triangles needs below its own node */
    unsigned levels(unsigned c) const { return ceil_log2(ceil_div(c, max_leaf)); }

    std::vector<SahNode> build(unsigned start, unsigned end, unsigned depth) const {
        std::vector<SahNode> out;
        unsigned count = end - start;
        if (count <= max_leaf) {
            Box lb = empty_box();
            for (unsigned i = start; i < end; i++)
                grow(lb, boxes[order[i]]);
            out.push_back(SahNode{lb, (int)start, count, 0u, 0u});
            return out;
        }
        Box bb = empty_box(), cb = empty_box();
        for (unsigned i = start; i < end; i++) {
            unsigned t = order[i];
            grow(bb, boxes[t]);
            for (int a = 0; a < 3; a++) {
                cb.lo[a] = std::min(cb.lo[a], cen[a][t]);
                cb.hi[a] = std::max(cb.hi[a], cen[a][t]);
            }
        }
        /* best binned split: least area-weighted count, among those that keep
           both children buildable within max_depth */
        int best_axis = -1, best_bin = 0;
        float best_cost = 0.0f;
        for (int a = 0; a < 3; a++) {
            float ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0.0f))
                continue;
            float scale = BINS / ext;
            Box bin_box[BINS];
            unsigned bin_n[BINS] = {};
            for (int k = 0; k < BINS; k++)
                bin_box[k] = empty_box();
            for (unsigned i = start; i < end; i++) {
                unsigned t = order[i];
                int k = std::min(BINS - 1, (int)((cen[a][t] - cb.lo[a]) * scale));
                bin_n[k]++;
                grow(bin_box[k], boxes[t]);
            }
            float left_area[BINS];
            unsigned left_n[BINS];
            Box acc = empty_box();
            unsigned n_acc = 0;
            for (int k = 0; k < BINS; k++) {
                if (bin_n[k])
                    grow(acc, bin_box[k]);
                n_acc += bin_n[k];
                left_area[k] = half_area(acc);
                left_n[k] = n_acc;
            }
            acc = empty_box();
            n_acc = 0;
            for (int k = BINS - 1; k > 0; k--) {
                if (bin_n[k])
                    grow(acc, bin_box[k]);
                n_acc += bin_n[k];
                unsigned nl = left_n[k - 1], nr = n_acc;
                if (nl == 0 || nr == 0 || depth + levels(std::max(nl, nr)) > max_depth)
                    continue;
                float cost = left_area[k - 1] * (float)nl + half_area(acc) * (float)nr;
                if (best_axis < 0 || cost < best_cost) {
                    best_axis = a;
                    best_bin = k;
                    best_cost = cost;
                }
            }
        }
        unsigned mid;
        int axis;
        if (best_axis >= 0) {
            axis = best_axis;
            float lo = cb.lo[axis], scale = BINS / (cb.hi[axis] - cb.lo[axis]);
            const std::vector<float> &c = cen[axis];
            int bin = best_bin;
            auto it = std::stable_partition(order.begin() + start, order.begin() + end, [&](unsigned t) {
                return std::min(BINS - 1, (int)((c[t] - lo) * scale)) < bin;
            });
            mid = (unsigned)(it - order.begin());
        } else {
            /* object median along the longest centroid extent (index order
               breaks ties, so the build is deterministic) */
            float ex = cb.hi[0] - cb.lo[0], ey = cb.hi[1] - cb.lo[1], ez = cb.hi[2] - cb.lo[2];
            axis = ex >= ey ? (ex >= ez ? 0 : 2) : (ey >= ez ? 1 : 2);
            const std::vector<float> &c = cen[axis];
            std::sort(order.begin() + start, order.begin() + end,
                      [&c](unsigned i, unsigned j) { return c[i] < c[j] || (c[i] == c[j] && i < j); });
            mid = start + count / 2;
        }
        std::vector<SahNode> left, right;
        if (depth < 6 && count > 4096) {
            auto fut = std::async(std::launch::async, [this, start, mid, depth] { return build(start, mid, depth + 1); });
            right = build(mid, end, depth + 1);
            left = fut.get();
        } else {
            left = build(start, mid, depth + 1);
            right = build(mid, end, depth + 1);
        }
        const Box &lb = left[0].box, &rb = right[0].box;
        unsigned swap = (lb.lo[axis] + lb.hi[axis]) > (rb.lo[axis] + rb.hi[axis]) ? 4u : 0u;
        out.reserve(1 + left.size() + right.size());
        out.push_back(SahNode{bb, -1, 0u, (unsigned)(1 + left.size()), (unsigned)axis | swap});
        unsigned rofs = (unsigned)(1 + left.size());
        for (SahNode n : left) {
            if (n.first < 0)
                n.right += 1;
            out.push_back(n);
        }
        for (SahNode n : right) {
            if (n.first < 0)
                n.right += rofs;
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

extern "C" uint32_t phx_build_bvh_sah(const PhxTriangleVertices *tri_verts, uint32_t n_tris, uint32_t max_depth,
                                      uint32_t max_leaf, PhxBvhNode *out_nodes, uint32_t *out_order) {
    if (!tri_verts || !out_nodes || !out_order || n_tris == 0 || n_tris > PHX_PAIR_MAX_TRIS || max_leaf < 1 ||
        max_leaf > PHX_PAIR_MAX_LEAF)
        return 0;
    if (ceil_log2(ceil_div(n_tris, max_leaf)) > max_depth + 1)
        return 0; /* not even a balanced tree fits */
    std::vector<Box> boxes(n_tris);
    std::vector<float> cen[3];
    for (int a = 0; a < 3; a++)
        cen[a].resize(n_tris);
    for (uint32_t i = 0; i < n_tris; i++) {
        boxes[i] = triangle_box(tri_verts[i]);
        for (int a = 0; a < 3; a++)
            cen[a][i] = 0.5f * (boxes[i].lo[a] + boxes[i].hi[a]);
    }
    std::vector<unsigned> order(n_tris);
    std::iota(order.begin(), order.end(), 0u);
    SahBuilder b{boxes, cen, order, max_depth, max_leaf};
    std::vector<SahNode> nodes = b.build(0, n_tris, 0);
    for (size_t i = 0; i < nodes.size(); i++) {
        const SahNode &n = nodes[i];
        PhxBvhNode &o = out_nodes[i];
        o.min.x = n.box.lo[0];
        o.min.y = n.box.lo[1];
        o.min.z = n.box.lo[2];
        o.max.x = n.box.hi[0];
        o.max.y = n.box.hi[1];
        o.max.z = n.box.hi[2];
        o.min.w = phx_as_float(n.first < 0 ? 0xFFFFFFF8u | n.hint
                                           : PHX_PAIR_LEAF | ((uint32_t)n.first << 2) | (n.count - 1u));
        o.max.w = phx_as_float(n.first < 0 ? n.right : 0u);
    }
    for (uint32_t i = 0; i < n_tris; i++)
        out_order[i] = order[i];
    return (uint32_t)nodes.size();
}

extern "C" int phx_build_bvh_pairs(const PhxBvhNode *tree, uint32_t n_nodes, PhxBvhPair *out_pairs,
                                   uint32_t *root_ref) {
    if (!tree || !root_ref || n_nodes == 0 || (n_nodes & 1u) == 0 || (n_nodes > 1 && !out_pairs))
        return -1;
    /* pair index of every interior node, in preorder */
    std::vector<uint32_t> pair_of(n_nodes, 0xFFFFFFFFu);
    uint32_t n_pairs = 0;
    for (uint32_t i = 0; i < n_nodes; i++) {
        uint32_t w = phx_as_uint(tree[i].min.w);
        if (w >= 0xFFFFFFF8u)
            pair_of[i] = n_pairs++;
        else if (!(w & PHX_PAIR_LEAF))
            return -1; /* neither interior nor a leaf reference */
    }
    if (n_pairs != n_nodes / 2)
        return -1;
    auto ref = [&](uint32_t i) -> uint32_t {
        uint32_t w = phx_as_uint(tree[i].min.w);
        return w >= 0xFFFFFFF8u ? pair_of[i] : w;
    };
    for (uint32_t i = 0; i < n_nodes; i++) {
        if (pair_of[i] == 0xFFFFFFFFu)
            continue;
        uint32_t l = i + 1, r = phx_as_uint(tree[i].max.w);
        if (l >= n_nodes || r >= n_nodes || r <= i)
            return -1;
        PhxBvhPair &p = out_pairs[pair_of[i]];
        p.lmin = tree[l].min;
        p.lmax = tree[l].max;
        p.rmin = tree[r].min;
        p.rmax = tree[r].max;
        p.lmin.w = phx_as_float(ref(l));
        p.rmin.w = phx_as_float(ref(r));
        p.lmax.w = 0.0f;
        p.rmax.w = 0.0f;
    }
    *root_ref = ref(0);
    return (int)n_pairs;
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
