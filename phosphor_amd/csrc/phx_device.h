/*
 * phx_device.h — gfx950 device helpers of the SPPM pass.
 *
 * Restates the reference's device helpers (the headers of include/device/ and
 * the OpenCL halves of those of include/common/) for HIP.  Arithmetic follows
 * the reference expression by expression with IEEE division/sqrt, FP
 * contraction off and the elementary functions of phx_math.h, so a kernel
 * built from these helpers is a deterministic function of its inputs.
 * Vector builtins use one fixed evaluation order:
 * dot = ((x·x + y·y) + z·z) + w·w, length = sqrt(dot),
 * normalize(v) = v · (1/length(v)) with the zero vector left unchanged.
 *
 * Layout choices for gfx950: nodes, vertices and materials are read with
 * 16-byte vector loads straight from the reference's AoS structs (32-byte BVH
 * nodes = two dwordx4 loads); the traversal stack and the IOR stack live in
 * LDS in a [depth][lane] layout (consecutive lanes hit consecutive banks).
 */
#ifndef PHX_DEVICE_H
#define PHX_DEVICE_H

#include <hip/hip_runtime.h>

#include "../../include/phosphor_hip.h"
#include "internal.h"
#include "phx_math.h"

typedef uint32_t u32;
typedef int32_t i32;
typedef float f32;

#define PHX_DEV static __device__ __forceinline__

namespace phx {

/* constants.h:6-45 */
constexpr f32 EPS = 1e-5f;
constexpr f32 INF = 1e9f;
constexpr f32 PI = 3.141592654f;
constexpr f32 U32_MAX_F = 4294967296.0f;
constexpr u32 NO_TEXTURE = 0xFFFFFFFFu;
/* min.w of a BVH node: a triangle index below this marks a leaf.  The
 * reference marks interior nodes with 0xFFFFFFFF; uploaded scenes may carry
 * a visiting-order hint in the low three bits instead (0xFFFFFFF8 | hint). */
constexpr u32 BVH_INTERIOR = 0xFFFFFFF8u;
constexpr f32 AIR_IOR = 1.0f;
constexpr u32 MAX_RAY_BOUNCES = 5;
constexpr int BVH_STACK_SIZE = 24;
constexpr u32 IOR_STACK_SIZE = 8;
constexpr f32 MIN_ROUGHNESS = 0.001f;
constexpr u32 GGX_SAMPLING_ATTEMPTS = 8;
constexpr u32 RR_MIN_RAY_DEPTH = 2;
constexpr u32 RR_MIN_PHOTON_DEPTH = 2;
constexpr f32 RR_MIN_Q = 0.05f;
constexpr f32 RR_MAX_Q = 1.0f;
constexpr f32 GAUSS_ALPHA = 1.818f;
constexpr f32 GAUSS_BETA = 1.953f;
constexpr f32 GAUSS_DENOM = 0.85815f;
constexpr u32 BSDF_DIFFUSE = 0, BSDF_METALLIC = 1, BSDF_FRESNEL = 2, BSDF_TRANSMIT = 3;

/* ---------------- float4 with OpenCL semantics ---------------- */
struct F4 {
    f32 x, y, z, w;
};
PHX_DEV F4 mk(f32 x, f32 y, f32 z, f32 w) { return F4{x, y, z, w}; }
PHX_DEV F4 ld(const phx_float4 *p) {
    phx_float4 v = *p;
    return F4{v.x, v.y, v.z, v.w};
}
PHX_DEV F4 from(const phx_float4 &v) { return F4{v.x, v.y, v.z, v.w}; }
PHX_DEV void st(phx_float4 *p, F4 a) {
    phx_float4 v;
    v.x = a.x;
    v.y = a.y;
    v.z = a.z;
    v.w = a.w;
    *p = v;
}
PHX_DEV F4 operator+(F4 a, F4 b) { return F4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
PHX_DEV F4 operator-(F4 a, F4 b) { return F4{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
PHX_DEV F4 operator*(F4 a, F4 b) { return F4{a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
PHX_DEV F4 operator*(f32 s, F4 a) { return F4{s * a.x, s * a.y, s * a.z, s * a.w}; }
PHX_DEV F4 operator*(F4 a, f32 s) { return F4{a.x * s, a.y * s, a.z * s, a.w * s}; }
PHX_DEV F4 operator/(F4 a, f32 s) { return F4{a.x / s, a.y / s, a.z / s, a.w / s}; }
PHX_DEV F4 operator-(F4 a) { return F4{-a.x, -a.y, -a.z, -a.w}; }
PHX_DEV f32 dot(F4 a, F4 b) { return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w; }
PHX_DEV F4 cross(F4 a, F4 b) { return F4{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.0f}; }
PHX_DEV f32 length(F4 a) { return phx_sqrtf(dot(a, a)); }
PHX_DEV F4 normalize(F4 a) {
    f32 d = dot(a, a);
    if (d == 0.0f)
        return a;
    f32 inv = 1.0f / phx_sqrtf(d);
    return a * inv;
}
constexpr F4 ZERO = {0.0f, 0.0f, 0.0f, 0.0f};
constexpr F4 BLACK = {0.0f, 0.0f, 0.0f, 0.0f};
constexpr F4 WHITE = {1.0f, 1.0f, 1.0f, 0.0f};

/* Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
 * with its own L2.  Mapping dispatch order b to logical block
 * (b mod 8)·⌈n/8⌉-style chunks gives every XCD one contiguous range of
 * logical blocks, so blocks that touch neighbouring data share an L2.  The
 * map is a bijection on [0, n). */
PHX_DEV u32 xcd_block(u32 b, u32 n) {
    u32 q = n / 8u, r = n % 8u;
    u32 x = b % 8u, k = b / 8u;
    return x < r ? x * (q + 1u) + k : r * (q + 1u) + (x - r) * q + k;
}

/* ---------------- random.h:16-62 (PCG32, RXS-M-XS output) ---------------- */
PHX_DEV u32 pcg_seed(u32 seed) {
    u32 s = 0u;
    s = s * 747796405u + 2891336453u;
    s += seed;
    s = s * 747796405u + 2891336453u;
    return s;
}
PHX_DEV u32 random_u32(u32 &state) {
    u32 old = state;
    state = state * 747796405u + 2891336453u;
    return ((old >> ((old >> 28u) + 4u)) ^ old) * 277803737u;
}
PHX_DEV f32 random_float(u32 &state) { return (f32)random_u32(state) / U32_MAX_F; }

/* utils.h:59-71 (returned by value so the frame stays in registers) */
struct TBN {
    F4 t, b;
};
PHX_DEV TBN make_tbn(F4 n) {
    n.w = 0.0f;
    bool big_x = phx_fabsf(n.x) > 0.1f;
    /* cross(n, e_y) or cross(n, e_x), written out (same products and order) */
    F4 t = big_x ? cross(n, mk(0.0f, 1.0f, 0.0f, 0.0f)) : cross(n, mk(1.0f, 0.0f, 0.0f, 0.0f));
    if (length(t) < EPS)
        t = cross(n, mk(0.0f, 0.0f, 1.0f, 0.0f));
    t = normalize(t);
    return TBN{t, cross(n, t)};
}
/* random.h:40-46 */
PHX_DEV F4 random_unit_vector(u32 &rng) {
    f32 z = 2.0f * random_float(rng) - 1.0f;
    f32 r = phx_sqrtf(phx_fmaxf(0.0f, 1.0f - z * z));
    f32 phi = 2.0f * PI * random_float(rng);
    f32 s, c;
    phx_sincosf(phi, &s, &c);
    return mk(r * c, r * s, z, 0.0f);
}
/* random.h:48-62 */
PHX_DEV F4 random_in_unit_hemisphere(u32 &rng, F4 n) {
    n.w = 0.0f;
    f32 r1 = random_float(rng);
    f32 r2 = random_float(rng);
    f32 phi = 2.0f * PI * r1;
    f32 st_ = phx_sqrtf(r2);
    f32 ct = phx_sqrtf(1.0f - r2);
    f32 s, c;
    phx_sincosf(phi, &s, &c);
    F4 hl = mk(st_ * c, st_ * s, ct, 0.0f);
    TBN f = make_tbn(n);
    return normalize(hl.x * f.t + hl.y * f.b + hl.z * n);
}

/* camera.h:28-47 */
PHX_DEV void make_ray(const PhxCameraParams &cam, u32 px, u32 py, u32 w, u32 h, u32 &rng, F4 &o, F4 &d) {
    f32 u = random_float(rng) - 0.5f;
    f32 v = random_float(rng) - 0.5f;
    f32 s = ((f32)px + 0.5f + u) / (f32)w;
    f32 t = 1.0f - ((f32)py + 0.5f + v) / (f32)h;
    F4 disk = mk(0.0f, 0.0f, 0.0f, 0.0f);
    if (cam.defocus_angle > 0.0f) {
        f32 r1 = phx_sqrtf(random_float(rng));
        f32 r2 = 2.0f * PI * random_float(rng);
        f32 sn, c;
        phx_sincosf(r2, &sn, &c);
        disk = r1 * (c * from(cam.defocus_disk_u) + sn * from(cam.defocus_disk_v));
    }
    o = from(cam.position) + disk;
    F4 dir = from(cam.lower_left_corner) + s * from(cam.horizontal) + t * from(cam.vertical) - o;
    d = normalize(dir);
}

/* ---------------- encoding.h:10-74 ---------------- */
PHX_DEV u32 encode_rgbe(F4 p) {
    f32 m = phx_fmaxf(p.x, phx_fmaxf(p.y, p.z));
    if (m <= 0.0f)
        return 0u;
    int e = phx_frexp_exp(m);
    f32 scale = phx_ldexp1(e);
    u32 r = phx_f2u(p.x / scale * 256.0f);
    u32 g = phx_f2u(p.y / scale * 256.0f);
    u32 b = phx_f2u(p.z / scale * 256.0f);
    u32 E = (u32)(e + 128);
    r = r < 255u ? r : 255u;
    g = g < 255u ? g : 255u;
    b = b < 255u ? b : 255u;
    return r | (g << 8) | (b << 16) | (E << 24);
}
PHX_DEV F4 decode_rgbe(u32 rgbe) {
    u32 E = rgbe >> 24;
    if (E == 0)
        return mk(0.0f, 0.0f, 0.0f, 0.0f);
    f32 scale = phx_ldexp1((i32)E - 128 - 8);
    return mk((f32)(rgbe & 0xFF) * scale, (f32)((rgbe >> 8) & 0xFF) * scale, (f32)((rgbe >> 16) & 0xFF) * scale, 0.0f);
}
/* decode_rgbe with the hardware's exact ldexp: c·2^(E−136) is a multiple of
 * 2^-149 no larger than 255·2^119, representable exactly, so both forms give
 * the same bits (E = 0 decodes to zero) */
PHX_DEV F4 decode_rgbe_ldexp(u32 rgbe) {
    u32 E = rgbe >> 24;
    i32 k = E == 0 ? -200 : (i32)E - 136;
    return mk(__builtin_amdgcn_ldexpf((f32)(rgbe & 0xFF), k), __builtin_amdgcn_ldexpf((f32)((rgbe >> 8) & 0xFF), k),
              __builtin_amdgcn_ldexpf((f32)((rgbe >> 16) & 0xFF), k), 0.0f);
}
PHX_DEV u32 encode_oct(F4 dir) {
    dir = normalize(dir);
    f32 l = phx_fabsf(dir.x) + phx_fabsf(dir.y) + phx_fabsf(dir.z);
    f32 nx = dir.x / l, ny = dir.y / l;
    if (!(dir.z >= 0.0f)) {
        f32 wx = (1.0f - phx_fabsf(ny)) * (nx >= 0.0f ? 1.0f : -1.0f);
        f32 wy = (1.0f - phx_fabsf(nx)) * (ny >= 0.0f ? 1.0f : -1.0f);
        nx = wx;
        ny = wy;
    }
    u32 ux = phx_f2u(phx_clampf(nx * 0.5f + 0.5f, 0.0f, 1.0f) * 65535.0f);
    u32 uy = phx_f2u(phx_clampf(ny * 0.5f + 0.5f, 0.0f, 1.0f) * 65535.0f);
    return ux | (uy << 16);
}
PHX_DEV F4 decode_oct(u32 packed) {
    f32 fx = (f32)(packed & 0xFFFF) / 65535.0f;
    f32 fy = (f32)((packed >> 16) & 0xFFFF) / 65535.0f;
    fx = fx * 2.0f - 1.0f;
    fy = fy * 2.0f - 1.0f;
    F4 n = mk(fx, fy, 1.0f - phx_fabsf(fx) - phx_fabsf(fy), 0.0f);
    f32 t = phx_clampf(-n.z, 0.0f, 1.0f);
    n.x += (n.x >= 0.0f ? -t : t);
    n.y += (n.y >= 0.0f ? -t : t);
    return normalize(n);
}

/* registers and occupancy of one kernel (phx_kernel_info) */
template <typename K>
inline bool kernel_info(const char *name, K kernel, u32 block, PhxKernelInfo &out) {
    hipFuncAttributes fa;
    int blocks = 0;
    if (hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(kernel)) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, (int)block, 0) != hipSuccess)
        return false;
    for (u32 i = 0; i < sizeof(out.name); i++) {
        out.name[i] = name[i];
        if (!name[i])
            break;
    }
    out.name[sizeof(out.name) - 1] = 0;
    out.vgprs = (u32)fa.numRegs;
    out.lds_bytes = (u32)fa.sharedSizeBytes;
    out.scratch_bytes = (u32)fa.localSizeBytes;
    out.block = block;
    out.waves_per_simd = (u32)blocks * (block / 64u) / 4u;
    return true;
}
/* appends one kernel's entry to out[n] (when n < max), counting it in n;
 * clears ok when the runtime cannot describe it */
#define PHX_INFO(NAME, KERNEL, BLOCK)                                                                                 \
    if (kernel_info(NAME, KERNEL, BLOCK, k)) {                                                                         \
        if (n < max)                                                                                                   \
            out[n] = k;                                                                                                \
        n++;                                                                                                           \
    } else {                                                                                                           \
        ok = false;                                                                                                    \
    }

/* ---------------- photon_hash.h:21-37 ---------------- */
PHX_DEV u32 get_hash(f32 px, f32 py, f32 pz, const PhxPhotonHashData &d) {
    u32 g = d.grid_res;
    f32 fx = (px - d.origin.x) / d.cell_sizes.x;
    f32 fy = (py - d.origin.y) / d.cell_sizes.y;
    f32 fz = (pz - d.origin.z) / d.cell_sizes.z;
    if (fx < 0.0f || fy < 0.0f || fz < 0.0f || fx >= (f32)g || fy >= (f32)g || fz >= (f32)g)
        return 0;
    u32 x = (u32)fx, y = (u32)fy, z = (u32)fz;
    return g * g * x + g * y + z + 1;
}

/* get_hash with the half-cell octant of the point inside its cell:
 * 8·get_hash + (x half) + 2·(y half) + 4·(z half), a half being 1 where the
 * quotient's fraction f − ⌊f⌋ (exact for the same rounded quotient) is at
 * least 0.5; 0 outside the grid.  Sorting by this key keeps every cell's
 * records contiguous (the cell ranges are unchanged) and orders them by
 * octant inside the cell, so that the gather can skip octants. */
PHX_DEV u32 get_hash_octant(f32 px, f32 py, f32 pz, const PhxPhotonHashData &d) {
    u32 g = d.grid_res;
    f32 fx = (px - d.origin.x) / d.cell_sizes.x;
    f32 fy = (py - d.origin.y) / d.cell_sizes.y;
    f32 fz = (pz - d.origin.z) / d.cell_sizes.z;
    if (fx < 0.0f || fy < 0.0f || fz < 0.0f || fx >= (f32)g || fy >= (f32)g || fz >= (f32)g)
        return 0;
    u32 x = (u32)fx, y = (u32)fy, z = (u32)fz;
    u32 o = (fx - (f32)x >= 0.5f ? 1u : 0u) | (fy - (f32)y >= 0.5f ? 2u : 0u) | (fz - (f32)z >= 0.5f ? 4u : 0u);
    return (g * g * x + g * y + z + 1) * 8u + o;
}

/* scene pointers of one render (kernel argument block) */
typedef SceneArgs SceneView;

struct HitRecord {
    f32 t, u, v;
    u32 tri_index, mat_index;
};

/* hit.h:45-78 for the triangle with vertices *tv, recorded as tri_idx */
PHX_DEV bool is_triangle_hit(const PhxTriangleVertices *tv, u32 tri_idx, F4 o, F4 d, f32 t_min, f32 t_max,
                             HitRecord &rec) {
    F4 v0 = ld(&tv->v0), v1 = ld(&tv->v1), v2 = ld(&tv->v2);
    F4 e1 = v1 - v0;
    F4 e2 = v2 - v0;
    F4 h = cross(d, e2);
    f32 a = dot(e1, h);
    if (phx_fabsf(a) < EPS)
        return false;
    f32 f = 1.0f / a;
    F4 s = o - v0;
    f32 u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f)
        return false;
    F4 q = cross(s, e1);
    f32 v = f * dot(d, q);
    if (v < 0.0f || u + v > 1.0f)
        return false;
    f32 t = f * dot(e2, q);
    if (t < t_min || t > t_max)
        return false;
    rec.t = t;
    rec.u = u;
    rec.v = v;
    rec.tri_index = tri_idx;
    return true;
}

PHX_DEV bool is_triangle_hit(const SceneView &sc, u32 tri_idx, F4 o, F4 d, f32 t_min, f32 t_max, HitRecord &rec) {
    return is_triangle_hit(sc.tv + tri_idx, tri_idx, o, d, t_min, t_max, rec);
}

/* hit.h:80-133, one axis of the slab test; inv = 1/dir component (the same
 * value the reference recomputes at every node) */
PHX_DEV bool slab(f32 mn, f32 mx, f32 o, f32 inv, f32 &t_min, f32 &t_max) {
    f32 t0 = (mn - o) * inv;
    f32 t1 = (mx - o) * inv;
    if (t0 < t1) {
        if (t0 > t_min)
            t_min = t0;
        if (t1 < t_max)
            t_max = t1;
    } else {
        if (t1 > t_min)
            t_min = t1;
        if (t0 < t_max)
            t_max = t0;
    }
    return t_max > t_min;
}

/* hit.h:135-172.  Depth-first, left child first, no near/far ordering.
 * The reference pushes the right child, pushes the left child and pops the
 * left child at once; here the left child stays in a register and only the
 * right child is pushed, which visits nodes in the same order.  With q the
 * index of the top entry of this stack, the reference's stack pointer after
 * a pop equals q, so its overflow rule (drop both children when p + 2 >= 24)
 * is applied as q + 2 >= 24.  `stk` points at this lane's column of an LDS
 * array laid out [depth][lanes]; `stride` is the number of lanes.
 * ANY: stop at the first accepted triangle (used only where any hit decides
 * the result, see the shadow kernel). */
template <bool ANY>
PHX_DEV bool traverse(const SceneView &sc, F4 o, F4 d, f32 t_min, f32 t_max, HitRecord &rec, u32 *stk,
                      u32 stride) {
    f32 ix = 1.0f / d.x, iy = 1.0f / d.y, iz = 1.0f / d.z;
    int q = -1;
    u32 index = 0;
    bool found = false;
    f32 closest = t_max;
    for (;;) {
        const PhxBvhNode *node = sc.nodes + index;
        F4 nmin = ld(&node->min);
        F4 nmax = ld(&node->max);
        f32 lo = t_min, hi = closest;
        bool box = slab(nmin.x, nmax.x, o.x, ix, lo, hi) && slab(nmin.y, nmax.y, o.y, iy, lo, hi) &&
                   slab(nmin.z, nmax.z, o.z, iz, lo, hi);
        if (box) {
            u32 tri = phx_as_uint(nmin.w);
            if (tri < BVH_INTERIOR) {
                HitRecord cand;
                if (is_triangle_hit(sc, tri, o, d, t_min, closest, cand)) {
                    closest = cand.t;
                    rec = cand;
                    found = true;
                    if (ANY)
                        break;
                }
            } else if (q + 2 < BVH_STACK_SIZE) {
                stk[(q + 1) * stride] = phx_as_uint(nmax.w);
                q++;
                index = index + 1;
                continue;
            }
        }
        if (q < 0)
            break;
        index = stk[q * stride];
        q--;
    }
    if (found)
        rec.mat_index = sc.tmat[rec.tri_index];
    return found;
}

/* The hit traverse<ANY> finds on the reference tree, found on the library's
 * SAH tree over the same triangles (bvh.cpp phx_build_bvh_sah), stored as
 * pairs: each step loads one 64-byte node holding both children's boxes,
 * tests both, enters the one with the nearer entry point and pushes the
 * other.  A leaf child names one to four consecutive triangles of the tree's
 * order (vertices in sc.tvo), tested when its turn comes; a hit records the
 * position, mapped to the triangle index (sc.tri_order) at the end.
 *
 * Why the result is the same.  The reference's depth-first search with its
 * strict box test (hi > lo) returns the hit of least t; among several hits at
 * that t the one met last, i.e. in the leaf of largest preorder index
 * (is_triangle_hit accepts t <= t_max, so a later equal hit overwrites an
 * earlier one, and nodes are stored in preorder) -- except at t == t_min,
 * where the first one met wins, because once closest == t_min every further
 * box fails the strict test.  The reference's leaf boxes are the triangle
 * boxes expanded by 1e-3 and every leaf box of the SAH tree contains its
 * triangles' expanded boxes, so a leaf holding a hit at t in [t_min, closest]
 * has entry < t < exit by that margin and passes the inclusive test below
 * whenever it is tested; a subtree is dropped only when no point of its box
 * lies in [t_min, closest].  So every candidate triangle is met whatever the
 * order, and "least t; at equal t the larger reference leaf index (tri_key),
 * at t == t_min the smaller" is the reference's hit -- provided the reference
 * loses nothing to its stack-overflow rule (hit.h:163).  The host builds the
 * SAH tree only when no interior node of the reference tree is deeper than
 * 22, and builds it no deeper than 22 itself, so neither search overflows (a
 * pair search holds at most one pending child per interior level), and only
 * for scenes within |coordinate| <= 1024, where the box test's rounding (a few
 * 1e-4 at most) stays inside the 1e-3 leaf-box margin.  ANY (any hit in
 * [t_min, t_max], for shadow rays in opaque scenes): the reference's
 * closest-hit loop finds a hit exactly when one exists, and so does this
 * search.
 *
 * The box test is the usual fused form t = box * inv + (-origin * inv) with
 * min/max; it only decides which nodes are visited, never the hit record,
 * which is_triangle_hit computes exactly as the reference. */
struct RayInv {
    f32 ix, iy, iz, ox, oy, oz;
};

PHX_DEV RayInv ray_inv(F4 o, F4 d) {
    const f32 tiny = 1e-20f;
    RayInv r;
    r.ix = 1.0f / (phx_fabsf(d.x) > tiny ? d.x : __builtin_copysignf(tiny, d.x));
    r.iy = 1.0f / (phx_fabsf(d.y) > tiny ? d.y : __builtin_copysignf(tiny, d.y));
    r.iz = 1.0f / (phx_fabsf(d.z) > tiny ? d.z : __builtin_copysignf(tiny, d.z));
    r.ox = -o.x * r.ix;
    r.oy = -o.y * r.iy;
    r.oz = -o.z * r.iz;
    return r;
}

/* entry and exit of [t_min, closest] through a box; it is hit when lo <= hi */
PHX_DEV void box_span(const RayInv &r, F4 mn, F4 mx, f32 t_min, f32 closest, f32 &lo, f32 &hi) {
    f32 x0 = __builtin_fmaf(mn.x, r.ix, r.ox), x1 = __builtin_fmaf(mx.x, r.ix, r.ox);
    f32 y0 = __builtin_fmaf(mn.y, r.iy, r.oy), y1 = __builtin_fmaf(mx.y, r.iy, r.oy);
    f32 z0 = __builtin_fmaf(mn.z, r.iz, r.oz), z1 = __builtin_fmaf(mx.z, r.iz, r.oz);
    lo = __builtin_fmaxf(__builtin_fmaxf(__builtin_fminf(x0, x1), __builtin_fminf(y0, y1)),
                         __builtin_fmaxf(__builtin_fminf(z0, z1), t_min));
    hi = __builtin_fminf(__builtin_fminf(__builtin_fmaxf(x0, x1), __builtin_fmaxf(y0, y1)),
                         __builtin_fminf(__builtin_fmaxf(z0, z1), closest));
}

/* GROUP_LEAVES: interior steps run until the lane holds a leaf (or has
 * nothing left) and the lanes of a wave then test their triangles together;
 * otherwise each iteration takes one step or one triangle test.  Same hits
 * either way; grouping pays for incoherent rays (photon paths), the plain
 * loop for coherent ones (a pixel's shadow rays, camera rays). */
template <bool ANY, bool GROUP_LEAVES>
PHX_DEV bool traverse_pairs(const SceneView &sc, F4 o, F4 d, f32 t_min, f32 t_max, HitRecord &rec, u32 *stk,
                            u32 stride) {
    if (!(t_min < t_max))
        return false; /* the reference's root test fails */
    RayInv ri = ray_inv(o, d);
    f32 closest = t_max;
    bool found = false;
    int q = -1;
    u32 ref = sc.root_ref;
#define PHX_PAIR_LEAF_TEST(BREAK_STMT)                                                                                 \
    {                                                                                                                  \
        const u32 j0 = (ref & ~PHX_PAIR_LEAF) >> 2, j1 = j0 + (ref & 3u);                                              \
        for (u32 j = j0; j <= j1; j++) {                                                                               \
            HitRecord cand;                                                                                            \
            if (is_triangle_hit(sc.tvo + j, j, o, d, t_min, closest, cand)) {                                          \
                bool take = ANY || !found || cand.t < closest;                                                         \
                if (!take) { /* cand.t == closest: the reference's order decides */                                    \
                    u32 kc = sc.tri_key[j], kb = sc.tri_key[rec.tri_index];                                            \
                    take = (cand.t == t_min) ? kc < kb : kc > kb;                                                      \
                }                                                                                                      \
                if (take) {                                                                                            \
                    closest = cand.t;                                                                                  \
                    rec = cand;                                                                                        \
                    found = true;                                                                                      \
                }                                                                                                      \
                if (ANY)                                                                                               \
                    break;                                                                                             \
            }                                                                                                          \
        }                                                                                                              \
        if (ANY && found)                                                                                              \
            BREAK_STMT;                                                                                                \
    }
#define PHX_PAIR_STEP(NEAR_FAR, ONE, NONE)                                                                             \
    {                                                                                                                  \
        const PhxBvhPair *p = sc.pairs + ref;                                                                          \
        F4 lmin = ld(&p->lmin), lmax = ld(&p->lmax), rmin = ld(&p->rmin), rmax = ld(&p->rmax);                         \
        f32 llo, lhi, rlo, rhi;                                                                                        \
        box_span(ri, lmin, lmax, t_min, closest, llo, lhi);                                                            \
        box_span(ri, rmin, rmax, t_min, closest, rlo, rhi);                                                            \
        bool hl = llo <= lhi, hr = rlo <= rhi;                                                                         \
        u32 lref = phx_as_uint(lmin.w), rref = phx_as_uint(rmin.w);                                                    \
        if (hl && hr) {                                                                                                \
            bool right_first = rlo < llo;                                                                              \
            q++;                                                                                                       \
            stk[q * stride] = right_first ? lref : rref;                                                               \
            ref = right_first ? rref : lref;                                                                           \
            NEAR_FAR;                                                                                                  \
        } else if (hl || hr) {                                                                                         \
            ref = hl ? lref : rref;                                                                                    \
            ONE;                                                                                                       \
        } else {                                                                                                       \
            NONE;                                                                                                      \
        }                                                                                                              \
    }
    if (GROUP_LEAVES) {
        /* DONE has the leaf bit but names no triangle position (a tree of
           pairs orders at most PHX_PAIR_MAX_TRIS triangles) */
        constexpr u32 DONE = 0xFFFFFFFFu;
        for (;;) {
            while (!(ref & PHX_PAIR_LEAF)) {
                PHX_PAIR_STEP((void)0, (void)0, {
                    if (q >= 0) {
                        ref = stk[q * stride];
                        q--;
                    } else {
                        ref = DONE;
                    }
                })
            }
            if (ref == DONE)
                break;
            PHX_PAIR_LEAF_TEST(break)
            if (q < 0)
                break;
            ref = stk[q * stride];
            q--;
        }
    } else {
        for (;;) {
            if (ref & PHX_PAIR_LEAF) {
                PHX_PAIR_LEAF_TEST(break)
            } else {
                PHX_PAIR_STEP(continue, continue, (void)0)
            }
            if (q < 0)
                break;
            ref = stk[q * stride];
            q--;
        }
    }
#undef PHX_PAIR_STEP
#undef PHX_PAIR_LEAF_TEST
    if (found) {
        rec.tri_index = sc.tri_order[rec.tri_index];
        rec.mat_index = sc.tmat[rec.tri_index];
    }
    return found;
}

/* The closest-hit search of traverse_pairs<false, ...> held as lane state
 * and advanced one operation at a time (one pair step, or one triangle test),
 * so that a persistent trace kernel can give a lane a new ray as soon as its
 * own search ends instead of idling until the slowest ray of its wave is
 * done, and can run the lanes that need a pair step and those that need a
 * triangle test in separate rounds.  The hit is the same whatever the order
 * of the operations (the argument of traverse_pairs: every leaf holding a
 * candidate passes every box test that includes [t_min, closest]); leaf tests
 * done later than the plain search would do them only leave `closest` larger
 * meanwhile, which costs visits, not results.
 * While the search runs rec.tri_index is the hit's position in the tree's
 * triangle order (sc.tri_order maps it to the triangle index). */
struct PairSearch {
    RayInv ri;
    F4 o, d;
    f32 t_min, closest;
    HitRecord rec;
    int q;
    u32 ref;
    bool found;
};

PHX_DEV void pair_search_begin(PairSearch &s, const SceneView &sc, F4 o, F4 d, f32 t_min, f32 t_max) {
    s.ri = ray_inv(o, d);
    s.o = o;
    s.d = d;
    s.t_min = t_min;
    s.closest = t_max;
    s.found = false;
    s.q = -1;
    s.ref = sc.root_ref;
    s.rec.t = 0.0f;
    s.rec.u = 0.0f;
    s.rec.v = 0.0f;
    s.rec.tri_index = 0;
    s.rec.mat_index = 0;
}

/* Pops the next pending child; false when none is left (the search ends). */
PHX_DEV bool pair_search_pop(PairSearch &s, u32 *stk, u32 stride) {
    if (s.q < 0)
        return false;
    s.ref = stk[s.q * stride];
    s.q--;
    return true;
}

/* One step at the pair s.ref names (s.ref has no leaf bit): both children's
 * boxes tested, the nearer hit child entered and the other pushed; false when
 * the search has ended.  Only called with t_min < t_max (the reference's
 * root test; callers check it). */
PHX_DEV bool pair_search_node(PairSearch &s, const SceneView &sc, u32 *stk, u32 stride) {
    const PhxBvhPair *p = sc.pairs + s.ref;
    F4 lmin = ld(&p->lmin), lmax = ld(&p->lmax), rmin = ld(&p->rmin), rmax = ld(&p->rmax);
    f32 llo, lhi, rlo, rhi;
    box_span(s.ri, lmin, lmax, s.t_min, s.closest, llo, lhi);
    box_span(s.ri, rmin, rmax, s.t_min, s.closest, rlo, rhi);
    bool hl = llo <= lhi, hr = rlo <= rhi;
    u32 lref = phx_as_uint(lmin.w), rref = phx_as_uint(rmin.w);
    if (hl && hr) {
        bool right_first = rlo < llo;
        s.q++;
        stk[s.q * stride] = right_first ? lref : rref;
        s.ref = right_first ? rref : lref;
        return true;
    }
    if (hl || hr) {
        s.ref = hl ? lref : rref;
        return true;
    }
    return pair_search_pop(s, stk, stride);
}

/* The first triangle left of the leaf s.ref names (PHX_PAIR_LEAF | j << 2 |
 * count − 1): tested with traverse_pairs' rules, then s.ref names the rest of
 * the leaf or the next pending child; false when the search has ended.  ANY:
 * the search ends at the first hit (traverse_pairs<true, ...>). */
template <bool ANY = false>
PHX_DEV bool pair_search_tri(PairSearch &s, const SceneView &sc, u32 *stk, u32 stride) {
    const u32 j = (s.ref & ~PHX_PAIR_LEAF) >> 2, rest = s.ref & 3u;
    HitRecord cand;
    if (is_triangle_hit(sc.tvo + j, j, s.o, s.d, s.t_min, s.closest, cand)) {
        if (ANY) {
            s.rec = cand;
            s.found = true;
            return false;
        }
        bool take = !s.found || cand.t < s.closest;
        if (!take) { /* cand.t == closest: the reference's order decides */
            u32 kc = sc.tri_key[j], kb = sc.tri_key[s.rec.tri_index];
            take = (cand.t == s.t_min) ? kc < kb : kc > kb;
        }
        if (take) {
            s.closest = cand.t;
            s.rec = cand;
            s.found = true;
        }
    }
    if (rest > 0) {
        s.ref = PHX_PAIR_LEAF | ((j + 1u) << 2) | (rest - 1u);
        return true;
    }
    return pair_search_pop(s, stk, stride);
}

/* traverse_pairs<ANY, false> run as rounds of pair steps or rounds of
 * triangle tests, whichever more lanes of the wave are waiting for (lanes
 * whose searches end stay idle until the wave's last search ends).  Same
 * hits as traverse_pairs (the argument of PairSearch). */
template <bool ANY>
PHX_DEV bool traverse_pairs_rounds(const SceneView &sc, F4 o, F4 d, f32 t_min, f32 t_max, HitRecord &rec, u32 *stk,
                                   u32 stride) {
    PairSearch sr;
    bool act = t_min < t_max; /* the reference's root test */
    if (act)
        pair_search_begin(sr, sc, o, d, t_min, t_max);
    for (;;) {
        const bool at_leaf = act && (sr.ref & PHX_PAIR_LEAF) != 0;
        const u32 n_act = (u32)__popcll(__ballot(act));
        if (n_act == 0)
            break;
        const u32 n_tri = (u32)__popcll(__ballot(at_leaf));
        if (n_tri > n_act - n_tri) {
            if (at_leaf)
                act = pair_search_tri<ANY>(sr, sc, stk, stride);
        } else if (act && !at_leaf) {
            act = pair_search_node(sr, sc, stk, stride);
        }
    }
    if (!(t_min < t_max) || !sr.found)
        return false;
    rec = sr.rec;
    rec.tri_index = sc.tri_order[rec.tri_index];
    rec.mat_index = sc.tmat[rec.tri_index];
    return true;
}

#ifndef PHX_ROUNDS_TRAVERSAL
#define PHX_ROUNDS_TRAVERSAL 0
#endif

/* any-hit test of shadow rays in opaque scenes: the reference tree
 * depth-first, or the SAH tree */
PHX_DEV bool scene_any_hit(const SceneView &sc, F4 o, F4 d, f32 t_min, f32 t_max, HitRecord &rec, u32 *stk,
                           u32 stride) {
    if (sc.near_first)
        return PHX_ROUNDS_TRAVERSAL ? traverse_pairs_rounds<true>(sc, o, d, t_min, t_max, rec, stk, stride)
                                    : traverse_pairs<true, false>(sc, o, d, t_min, t_max, rec, stk, stride);
    return traverse<true>(sc, o, d, t_min, t_max, rec, stk, stride);
}

template <bool GROUP_LEAVES = false>
PHX_DEV bool scene_intersect(const SceneView &sc, F4 o, F4 d, f32 t_min, f32 t_max, HitRecord &rec, u32 *stk,
                             u32 stride) {
    if (sc.near_first)
        return (PHX_ROUNDS_TRAVERSAL && !GROUP_LEAVES)
                   ? traverse_pairs_rounds<false>(sc, o, d, t_min, t_max, rec, stk, stride)
                   : traverse_pairs<false, GROUP_LEAVES>(sc, o, d, t_min, t_max, rec, stk, stride);
    return traverse<false>(sc, o, d, t_min, t_max, rec, stk, stride);
}

struct SurfaceHit {
    F4 position, normal;
    f32 uvx, uvy;
    f32 t;
    u32 mat_index, tri_index;
    bool front_face;
};

/* hit.h:31-43 */
PHX_DEV F4 compute_bary(f32 bx, f32 by, F4 v0, F4 v1, F4 v2) {
    f32 w0 = 1.0f - bx - by;
    f32 w1 = bx, w2 = by;
    return mk(w0 * v0.x + w1 * v1.x + w2 * v2.x, w0 * v0.y + w1 * v1.y + w2 * v2.y, w0 * v0.z + w1 * v1.z + w2 * v2.z,
              0.0f);
}

/* hit.h:174-197 */
PHX_DEV SurfaceHit process_hit(const SceneView &sc, const HitRecord &rec, F4 o, F4 d) {
    SurfaceHit sh;
    sh.t = rec.t;
    sh.mat_index = rec.mat_index;
    sh.tri_index = rec.tri_index;
    sh.position = o + d * rec.t;
    f32 u = rec.u, v = rec.v;
    f32 w = 1.0f - u - v;
    const PhxTriangleNormals *tn = sc.tn + rec.tri_index;
    F4 outward = normalize(compute_bary(u, v, ld(&tn->n0), ld(&tn->n1), ld(&tn->n2)));
    bool front = dot(d, outward) < 0.0f;
    sh.front_face = front;
    sh.normal = front ? outward : -outward;
    const PhxTriangleUvs *uv = sc.tuv + rec.tri_index;
    phx_float2 a = uv->uv0, b = uv->uv1, c = uv->uv2;
    sh.uvx = w * a.x + u * b.x + v * c.x;
    sh.uvy = w * a.y + u * b.y + v * c.y;
    return sh;
}

/* ---------------- material.h:66-149 (textures) ---------------- */
PHX_DEV F4 srgb_to_linear(F4 c) { return mk(phx_powf(c.x, 2.2f), phx_powf(c.y, 2.2f), phx_powf(c.z, 2.2f), c.w); }
PHX_DEV F4 texel4(const SceneView &sc, i32 idx) {
    return mk((f32)sc.atlas[idx], (f32)sc.atlas[idx + 1], (f32)sc.atlas[idx + 2], 255.0f) / 255.0f;
}
PHX_DEV F4 naive_sample_texture(const SceneView &sc, u32 index, f32 ux, f32 uy) {
    if (index == NO_TEXTURE)
        return WHITE;
    PhxTextureMeta m = sc.tex_meta[index];
    u32 px = phx_f2u(ux * (f32)m.width);
    px = px < m.width - 1 ? px : m.width - 1;
    u32 py = phx_f2u(uy * (f32)m.height);
    py = py < m.height - 1 ? py : m.height - 1;
    u32 texel = m.atlas_offset + (py * m.width + px) * 3u;
    return mk((f32)sc.atlas[texel + 0] / 255.0f, (f32)sc.atlas[texel + 1] / 255.0f, (f32)sc.atlas[texel + 2] / 255.0f,
              1.0f);
}
PHX_DEV F4 mix4(F4 a, F4 b, f32 t) { return a + (b - a) * t; }
PHX_DEV i32 imin(i32 a, i32 b) { return a < b ? a : b; }
PHX_DEV F4 sample_texture(const SceneView &sc, u32 index, f32 ux, f32 uy) {
    if (index == NO_TEXTURE)
        return WHITE;
    PhxTextureMeta m = sc.tex_meta[index];
    i32 w = (i32)m.width, h = (i32)m.height;
    ux = ux - phx_floorf(ux);
    uy = uy - phx_floorf(uy);
    uy = 1.0f - uy;
    f32 x = phx_fminf(ux * (f32)w, (f32)(w - 1));
    f32 y = phx_fminf(uy * (f32)h, (f32)(h - 1));
    i32 x1 = imin((i32)x, w - 1);
    i32 x2 = imin((i32)phx_ceilf(x), w - 1);
    i32 y1 = imin((i32)y, h - 1);
    i32 y2 = imin((i32)phx_ceilf(y), h - 1);
    i32 ch = (i32)m.channels, off = (i32)m.atlas_offset;
    F4 q11 = texel4(sc, off + (y1 * w + x1) * ch);
    F4 q12 = texel4(sc, off + (y2 * w + x1) * ch);
    F4 q21 = texel4(sc, off + (y1 * w + x2) * ch);
    F4 q22 = texel4(sc, off + (y2 * w + x2) * ch);
    if (x2 == x1 && y2 == y1)
        return naive_sample_texture(sc, index, ux, uy);
    if (x2 == x1)
        return mix4(q11, q12, y - (f32)y1);
    if (y2 == y1)
        return mix4(q11, q21, x - (f32)x1);
    f32 denom = (f32)(x2 - x1) * (f32)(y2 - y1);
    f32 w11 = ((f32)x2 - x) * ((f32)y2 - y) / denom;
    f32 w12 = ((f32)x2 - x) * (y - (f32)y1) / denom;
    f32 w21 = (x - (f32)x1) * ((f32)y2 - y) / denom;
    f32 w22 = (x - (f32)x1) * (y - (f32)y1) / denom;
    return w11 * q11 + w12 * q12 + w21 * q21 + w22 * q22;
}
PHX_DEV F4 sample_texture_uv(const SceneView &sc, u32 mat_index, u32 tex_index, f32 ux, f32 uy, u32 slot) {
    const PhxUvTransform *t = &sc.uvt[mat_index].diff + slot;
    f32 sx = ux * t->uv_scale.x, sy = uy * t->uv_scale.y;
    f32 sn, c;
    phx_sincosf(t->uv_rotation, &sn, &c);
    f32 rx = c * sx - sn * sy, ry = sn * sx + c * sy;
    return sample_texture(sc, tex_index, rx + t->uv_offset.x, ry + t->uv_offset.y);
}

/* ---------------- shading.h ---------------- */
struct Ctx {
    F4 base_color, emissive, shading_normal;
    f32 metallic, roughness, transmission, occlusion, ior;
};

/* IOR stack of shading.h:25-53 in LDS: ior[k * stride] for depth k */
struct IorStack {
    f32 *ior;
    u32 stride;
    u32 cur;
};
/* Every stack starts with air (shading.h:30-33) and entry 0 is never
 * overwritten (a push writes above it), so the bottom entry is read as the
 * constant: a path that never enters a medium reads no stack memory. */
PHX_DEV void init_ior_stack(IorStack &s) {
    s.cur = 0;
    s.ior[0] = AIR_IOR;
}
PHX_DEV f32 current_ior(const IorStack &s) { return s.cur == 0 ? AIR_IOR : s.ior[s.cur * s.stride]; }
PHX_DEV f32 target_ior(const IorStack &s, f32 inside, bool front) {
    if (front)
        return inside;
    return (s.cur <= 1) ? AIR_IOR : s.ior[(s.cur - 1) * s.stride];
}
PHX_DEV void update_ior_stack(IorStack &s, f32 inside, bool front) {
    if (front) {
        if (s.cur < IOR_STACK_SIZE - 1) {
            s.cur++;
            s.ior[s.cur * s.stride] = inside;
        }
    } else {
        if (s.cur > 0)
            s.cur--;
    }
}
PHX_DEV F4 apply_normal_map(F4 ms, F4 gn, F4 t, F4 b) {
    F4 n = mk(ms.x * 2.0f - 1.0f, ms.y * 2.0f - 1.0f, ms.z * 2.0f - 1.0f, 0.0f);
    F4 p = n.x * t + n.y * b + n.z * gn;
    f32 len = length(p);
    return (len > EPS) ? (p / len) : gn;
}

/* trace_rays.cl:83-93 / emit_photons.cl:66-76: tangent frame at the hit.  The
 * frame only feeds the normal map; it is computed only when one is bound
 * (the result of evaluate_material is the same either way). */
PHX_DEV TBN hit_tbn(const SceneView &sc, const HitRecord &rec, const SurfaceHit &sh) {
    f32 bw = 1.0f - rec.u - rec.v;
    const PhxTriangleTangents *tt = sc.tt + sh.tri_index;
    F4 tr = bw * ld(&tt->t0) + rec.u * ld(&tt->t1) + rec.v * ld(&tt->t2);
    F4 br = (length(tr) > EPS) ? cross(sh.normal, tr) : ZERO;
    if (length(br) > EPS) {
        F4 b = normalize(br);
        return TBN{cross(b, sh.normal), b};
    }
    return make_tbn(sh.normal);
}

/* shading.h:74-123 */
PHX_DEV Ctx evaluate_material(const SceneView &sc, const PhxMaterial &mat, u32 mi, const HitRecord &rec,
                              const SurfaceHit &sh) {
    Ctx c;
    f32 ux = sh.uvx, uy = sh.uvy;
    c.base_color = from(mat.base_color);
    if (mat.diff_index != NO_TEXTURE)
        c.base_color = c.base_color * srgb_to_linear(sample_texture_uv(sc, mi, mat.diff_index, ux, uy, 0));
    c.emissive = from(mat.emissive);
    if (mat.emis_index != NO_TEXTURE)
        c.emissive = c.emissive * srgb_to_linear(sample_texture_uv(sc, mi, mat.emis_index, ux, uy, 1));
    c.metallic = mat.metallic;
    c.roughness = mat.roughness;
    if (mat.metal_rough_index != NO_TEXTURE) {
        F4 mr = sample_texture_uv(sc, mi, mat.metal_rough_index, ux, uy, 4);
        c.metallic *= mr.z;
        c.roughness *= mr.y;
    }
    c.roughness = phx_clampf(c.roughness, MIN_ROUGHNESS, 1.0f);
    c.transmission = mat.transmission;
    if (mat.transmission_index != NO_TEXTURE)
        c.transmission *= sample_texture_uv(sc, mi, mat.transmission_index, ux, uy, 5).x;
    c.occlusion = 1.0f;
    if (mat.occlusion_index != NO_TEXTURE)
        c.occlusion = sample_texture_uv(sc, mi, mat.occlusion_index, ux, uy, 3).x;
    c.ior = mat.ior;
    c.shading_normal = sh.normal;
    if (mat.norm_index != NO_TEXTURE) {
        TBN f = hit_tbn(sc, rec, sh);
        c.shading_normal =
            apply_normal_map(sample_texture_uv(sc, mi, mat.norm_index, ux, uy, 2), sh.normal, f.t, f.b);
    }
    return c;
}

/* shading.h:125-133 */
PHX_DEV F4 beer_lambert(F4 att, f32 dist, f32 travel) {
    if (dist < EPS)
        return WHITE;
    F4 sigma =
        mk(-phx_logf(phx_fmaxf(att.x, EPS)), -phx_logf(phx_fmaxf(att.y, EPS)), -phx_logf(phx_fmaxf(att.z, EPS)), 0.0f) /
        dist;
    return mk(phx_expf(-sigma.x * travel), phx_expf(-sigma.y * travel), phx_expf(-sigma.z * travel), 0.0f);
}

/* ---------------- bsdfs.h ---------------- */
PHX_DEV F4 reflect(F4 i, F4 n) { return normalize(i - 2.0f * dot(i, n) * n); }
PHX_DEV F4 safe_reflect(F4 d, F4 n, F4 g) {
    F4 r = reflect(d, n);
    if (dot(r, g) <= 0.0f) {
        r = r - dot(r, g) * g;
        r = normalize(r);
    }
    return r;
}
/* bsdfs.h:27-72 (Heitz 2018 VNDF sampling with ≤ 8 rejection attempts) */
PHX_DEV F4 ggx_sample_vndf(u32 &rng, F4 normal, F4 geom_normal, F4 view, f32 roughness, u32 &fell_back) {
    f32 alpha = roughness * roughness;
    TBN f = make_tbn(normal);
    F4 t = f.t, b = f.b;
    F4 vl = normalize(mk(dot(view, t), dot(view, b), dot(view, normal), 0.0f));
    F4 vh = normalize(mk(alpha * vl.x, alpha * vl.y, vl.z, 0.0f));
    f32 len_sq = vh.x * vh.x + vh.y * vh.y;
    F4 T1 = (len_sq > 0.0f) ? mk(-vh.y, vh.x, 0.0f, 0.0f) * (1.0f / phx_sqrtf(len_sq)) : mk(1.0f, 0.0f, 0.0f, 0.0f);
    F4 T2 = cross(vh, T1);
    f32 s = 0.5f * (1.0f + vh.z);
    for (u32 a = 0; a < GGX_SAMPLING_ATTEMPTS; a++) {
        f32 u1 = random_float(rng);
        f32 u2 = random_float(rng);
        f32 r = phx_sqrtf(u1);
        f32 phi = 2.0f * PI * u2;
        f32 sn, cs;
        phx_sincosf(phi, &sn, &cs);
        f32 t1 = r * cs;
        f32 t2 = r * sn;
        t2 = (1.0f - s) * phx_sqrtf(phx_fmaxf(0.0f, 1.0f - t1 * t1)) + s * t2;
        f32 nz2 = phx_fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2);
        F4 nh = t1 * T1 + t2 * T2 + phx_sqrtf(nz2) * vh;
        F4 nl = normalize(mk(alpha * nh.x, alpha * nh.y, phx_fmaxf(0.0f, nh.z), 0.0f));
        F4 ne = normalize(nl.x * t + nl.y * b + nl.z * normal);
        F4 refl = reflect(-view, ne);
        if (dot(refl, geom_normal) >= 0.0f) {
            fell_back = 0u;
            return ne;
        }
    }
    fell_back = 1u;
    return normal;
}
PHX_DEV f32 smith_g1_ggx(f32 c, f32 alpha) {
    if (c <= 0.0f)
        return 0.0f;
    f32 a2 = alpha * alpha;
    return 2.0f * c / (c + phx_sqrtf(a2 + (1.0f - a2) * c * c));
}
/* pow(x, 5) of bsdfs.h:88,93 evaluated as ((x·x)·(x·x))·x */
PHX_DEV f32 pow5(f32 x) {
    f32 x2 = x * x;
    return (x2 * x2) * x;
}
PHX_DEV f32 fresnel(f32 r0, F4 i, F4 n) {
    f32 c = phx_fabsf(dot(i, n));
    return r0 + (1.0f - r0) * pow5(1.0f - c);
}
PHX_DEV F4 fresnel4(F4 r0, F4 i, F4 n) {
    f32 c = phx_fabsf(dot(i, n));
    f32 p = pow5(1.0f - c);
    return r0 + (mk(1.0f, 1.0f, 1.0f, 1.0f) - r0) * p;
}
PHX_DEV f32 fresnel_refracted(f32 i1, f32 i2, F4 i, F4 n) {
    f32 q = (i1 - i2) / (i1 + i2);
    return fresnel(q * q, i, n);
}
PHX_DEV F4 refract(F4 i, F4 n, f32 e1, f32 e2, bool &tir) {
    f32 eta = e1 / e2;
    f32 ndi = dot(i, n);
    f32 st2 = eta * eta * (1.0f - ndi * ndi);
    if (st2 > 1.0f) {
        tir = true;
        return reflect(i, n);
    }
    tir = false;
    F4 r = eta * i - (eta * ndi + phx_sqrtf(1.0f - st2)) * n;
    return normalize(r);
}
PHX_DEV f32 ggx_ndf(f32 ndh, f32 alpha) {
    f32 a2 = alpha * alpha;
    f32 d = ndh * ndh * (a2 - 1.0f) + 1.0f;
    return a2 / phx_fmaxf(PI * d * d, EPS);
}

struct Bsdf {
    F4 dir, throughput;
    F4 h;          /* the sampled GGX half vector (read by the adjoint) */
    u32 event;
    u32 fell_back; /* every draw of the half vector was rejected: h is the normal (read by the adjoint) */
};
/* shading.h:150-213 (the order of random draws is part of the contract) */
PHX_DEV Bsdf sample_bsdf(u32 &rng, const Ctx &c, F4 sn, F4 gn, F4 view, IorStack &ior, bool front) {
    Bsdf s;
    F4 h = ggx_sample_vndf(rng, sn, gn, view, c.roughness, s.fell_back);
    s.h = h;
    f32 alpha = c.roughness * c.roughness;
    if (random_float(rng) < c.metallic) {
        s.dir = safe_reflect(-view, h, gn);
        f32 cl = phx_fmaxf(0.0f, dot(s.dir, sn));
        f32 g1 = smith_g1_ggx(cl, alpha);
        s.throughput = fresnel4(c.base_color, -view, h) * g1;
        s.event = BSDF_METALLIC;
        return s;
    }
    f32 i1 = current_ior(ior);
    f32 i2 = target_ior(ior, c.ior, front);
    f32 fr = fresnel_refracted(i1, i2, -view, h);
    f32 rd = random_float(rng);
    if (rd < fr) {
        s.dir = safe_reflect(-view, h, gn);
        f32 cl = phx_fmaxf(0.0f, dot(s.dir, sn));
        f32 g1 = smith_g1_ggx(cl, alpha);
        s.throughput = mk(g1, g1, g1, 0.0f);
        s.event = BSDF_FRESNEL;
        return s;
    }
    f32 rt = (rd - fr) / phx_fmaxf(1.0f - fr, EPS);
    if (rt < c.transmission) {
        bool tir;
        F4 refr = refract(-view, h, i1, i2, tir);
        if (tir) {
            s.dir = safe_reflect(-view, h, gn);
            f32 cl = phx_fmaxf(0.0f, dot(s.dir, sn));
            f32 g1 = smith_g1_ggx(cl, alpha);
            s.throughput = mk(g1, g1, g1, 0.0f);
            s.event = BSDF_FRESNEL;
            return s;
        }
        if (dot(refr, h) < 0.0f)
            update_ior_stack(ior, c.ior, front);
        s.dir = refr;
        f32 cl = phx_fmaxf(0.0f, phx_fabsf(dot(refr, sn)));
        f32 g1 = smith_g1_ggx(cl, alpha);
        s.throughput = c.base_color * g1;
        s.event = BSDF_TRANSMIT;
        return s;
    }
    s.dir = random_in_unit_hemisphere(rng, sn);
    if (dot(s.dir, gn) <= 0.0f)
        s.dir = random_in_unit_hemisphere(rng, gn);
    s.throughput = c.base_color;
    s.event = BSDF_DIFFUSE;
    return s;
}

/* ---------------- light_sampling.h ---------------- */
struct EmissivePoint {
    u32 tri_idx;
    f32 bu, bv, bw;
    F4 pos, normal;
};
struct LightFragment {
    F4 radiance, dir;
    f32 dist, n_dot_dir;
    u32 light; /* index of the sampled light (read by the adjoint) */
};
PHX_DEV f32 spotlight_falloff(f32 ca, f32 ci, f32 co) {
    f32 t = phx_clampf((ca - co) / (ci - co + EPS), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
PHX_DEV void fill_light_vecs(F4 a, F4 b, F4 &dir, f32 &dist, f32 &dist_sq) {
    F4 d = b - a;
    dist_sq = phx_fmaxf(dot(d, d), EPS);
    dist = phx_sqrtf(dist_sq);
    dir = d / dist;
}
/* light_sampling.h:52-67 */
PHX_DEV u32 choose_emissive_triangle(u32 &rng, const SceneView &sc, u32 start, u32 n, f32 total_area) {
    if (n <= 1 || total_area <= EPS)
        return start;
    f32 r = random_float(rng) * total_area;
    f32 acc = 0.0f;
    for (u32 i = 0; i < n; i++) {
        const PhxTriangleVertices *ev = sc.etv + start + i;
        F4 v0 = ld(&ev->v0);
        acc += 0.5f * length(cross(ld(&ev->v1) - v0, ld(&ev->v2) - v0));
        if (r <= acc)
            return start + i;
    }
    return start + n - 1;
}
PHX_DEV EmissivePoint sample_emissive_point(u32 &rng, const SceneView &sc, u32 start, u32 n, f32 total_area) {
    EmissivePoint p;
    p.tri_idx = choose_emissive_triangle(rng, sc, start, n, total_area);
    f32 r1 = random_float(rng);
    f32 r2 = random_float(rng);
    f32 s = phx_sqrtf(r1);
    p.bu = 1.0f - s;
    p.bv = r2 * s;
    p.bw = 1.0f - p.bu - p.bv;
    const PhxTriangleVertices *ev = sc.etv + p.tri_idx;
    F4 v0 = ld(&ev->v0), v1 = ld(&ev->v1), v2 = ld(&ev->v2);
    p.pos = p.bw * v0 + p.bu * v1 + p.bv * v2;
    const PhxTriangleNormals *en = sc.etn + p.tri_idx;
    F4 nn = p.bw * ld(&en->n0) + p.bu * ld(&en->n1) + p.bv * ld(&en->n2);
    if (length(nn) > EPS)
        p.normal = normalize(nn);
    else
        p.normal = normalize(cross(v1 - v0, v2 - v0));
    return p;
}
PHX_DEV F4 apply_emissive_texture(const SceneView &sc, F4 color, u32 tex, const EmissivePoint &p) {
    if (tex == NO_TEXTURE)
        return color;
    const PhxTriangleUvs *e = sc.etuv + p.tri_idx;
    f32 ux = p.bw * e->uv0.x + p.bu * e->uv1.x + p.bv * e->uv2.x;
    f32 uy = p.bw * e->uv0.y + p.bu * e->uv1.y + p.bv * e->uv2.y;
    u32 mi = sc.etmat[p.tri_idx];
    return color * srgb_to_linear(sample_texture_uv(sc, mi, tex, ux, uy, 1));
}
/* light_sampling.h:117-139 */
PHX_DEV u32 select_light(u32 &rng, const SceneView &sc, f32 &scale) {
    if (sc.n_lights == 0 || sc.total_lum < EPS) {
        scale = 1.0f;
        return 0;
    }
    f32 r = random_float(rng) * sc.total_lum;
    u32 lo = 0, hi = sc.n_lights;
    while (lo < hi) {
        u32 mid = (lo + hi) / 2;
        if (sc.lum_pref_sum[mid] < r)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo >= sc.n_lights)
        lo = sc.n_lights - 1;
    f32 lum = (lo == 0) ? sc.lum_pref_sum[0] : (sc.lum_pref_sum[lo] - sc.lum_pref_sum[lo - 1]);
    scale = phx_fmaxf(lum / sc.total_lum, EPS);
    return lo;
}
/* light_sampling.h:200-227 */
PHX_DEV void sample_light(u32 &rng, const SceneView &sc, F4 &origin, F4 &dir, F4 &power, u32 *light = nullptr) {
    if (sc.n_lights == 0 || sc.total_lum < EPS) {
        origin = mk(0.0f, 0.0f, 0.0f, 0.0f);
        dir = mk(0.0f, 1.0f, 0.0f, 0.0f);
        power = BLACK;
        if (light)
            *light = 0xFFFFFFFFu;
        return;
    }
    f32 scale;
    u32 idx = select_light(rng, sc, scale);
    if (light)
        *light = idx;
    const PhxLight &l = sc.lights[idx];
    u32 kind = l.kind;
    if (kind == PHX_LIGHT_POINT) {
        origin = from(l.position);
        dir = random_unit_vector(rng);
        power = from(l.power) * (4.0f * PI);
    } else if (kind == PHX_LIGHT_SPOT) {
        origin = from(l.position);
        f32 ci = phx_cosf(l.aux.x);
        f32 co = phx_cosf(l.aux.y);
        f32 r1 = random_float(rng);
        f32 r2 = random_float(rng);
        f32 ct = 1.0f - r1 * (1.0f - co);
        f32 st_ = phx_sqrtf(phx_fmaxf(0.0f, 1.0f - ct * ct));
        f32 phi = 2.0f * PI * r2;
        f32 sn, cs;
        phx_sincosf(phi, &sn, &cs);
        F4 ldl = mk(st_ * cs, st_ * sn, ct, 0.0f);
        F4 ldir = normalize(from(l.direction));
        TBN f = make_tbn(ldir);
        dir = normalize(ldl.x * f.t + ldl.y * f.b + ldl.z * ldir);
        f32 fo = spotlight_falloff(ct, ci, co);
        f32 sa = 2.0f * PI * (1.0f - co);
        power = from(l.power) * fo * sa;
    } else if (kind == PHX_LIGHT_DIRECTIONAL) {
        f32 r1 = sc.scene_radius * phx_sqrtf(random_float(rng));
        f32 r2 = 2.0f * PI * random_float(rng);
        f32 sn, cs;
        phx_sincosf(r2, &sn, &cs);
        F4 disk = r1 * (cs * from(l.tangent) + sn * from(l.bitangent));
        dir = from(l.direction);
        origin = from(sc.scene_center) - from(l.direction) * 2.0f * sc.scene_radius + disk;
        power = from(l.power) * PI * sc.scene_radius * sc.scene_radius;
    } else {
        u32 tex = phx_as_uint(l.aux.x);
        u32 start = phx_as_uint(l.aux.y);
        u32 n = phx_as_uint(l.aux.z);
        f32 area = l.aux.w;
        EmissivePoint p = sample_emissive_point(rng, sc, start, n, area);
        dir = random_in_unit_hemisphere(rng, p.normal);
        origin = p.pos + EPS * dir;
        power = apply_emissive_texture(sc, from(l.power), tex, p);
    }
    power = power / scale;
}
/* The PCG draws sample_light_fragment takes, without its geometry: one to
 * select the light (when there is light), then for an emissive-mesh light of
 * area >= EPS one to choose the triangle (more than one triangle, area > EPS,
 * choose_emissive_triangle) and two for the point (sample_emissive_point). */
PHX_DEV void skip_light_fragment(u32 &rng, const SceneView &sc) {
    if (sc.n_lights == 0 || sc.total_lum < EPS)
        return;
    f32 scale;
    u32 li = select_light(rng, sc, scale);
    const PhxLight &l = sc.lights[li];
    u32 kind = l.kind;
    if (kind == PHX_LIGHT_POINT || kind == PHX_LIGHT_SPOT || kind == PHX_LIGHT_DIRECTIONAL)
        return;
    f32 area = l.aux.w;
    if (area < EPS)
        return;
    if (phx_as_uint(l.aux.z) > 1u && area > EPS)
        (void)random_u32(rng);
    (void)random_u32(rng);
    (void)random_u32(rng);
}

/* The PCG state `steps` draws later, by the LCG jump-ahead of F. Brown,
 * "Random number generation with arbitrary strides" (1994): the map
 * x -> a·x + c applied k times is x -> A·x + C with A and C formed from the
 * binary digits of k by squaring the map (all mod 2^32). */
PHX_DEV u32 pcg_advance(u32 state, u32 steps) {
    u32 acc_mult = 1u, acc_plus = 0u, cur_mult = 747796405u, cur_plus = 2891336453u;
    while (steps) {
        if (steps & 1u) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1u) * cur_plus;
        cur_mult *= cur_mult;
        steps >>= 1;
    }
    return acc_mult * state + acc_plus;
}

/* skip_light_fragment `count` times.  With one light every fragment selects
 * it and takes the same number of draws (one, plus for an emissive mesh of
 * area >= EPS one for the triangle when it has several and two for the
 * point), so the state jumps by count times that many steps at once. */
PHX_DEV void skip_light_fragments(u32 &rng, const SceneView &sc, u32 count) {
    if (sc.n_lights == 0 || sc.total_lum < EPS)
        return;
    if (sc.n_lights == 1) {
        const PhxLight &l = sc.lights[0];
        u32 per = 1u;
        if (l.kind != PHX_LIGHT_POINT && l.kind != PHX_LIGHT_SPOT && l.kind != PHX_LIGHT_DIRECTIONAL &&
            !(l.aux.w < EPS))
            per += (phx_as_uint(l.aux.z) > 1u && l.aux.w > EPS) ? 3u : 2u;
        rng = pcg_advance(rng, count * per);
        return;
    }
    for (u32 i = 0; i < count; i++)
        skip_light_fragment(rng, sc);
}

/* light_sampling.h:231-300 */
PHX_DEV bool sample_light_fragment(u32 &rng, const SceneView &sc, F4 pos, F4 normal, LightFragment &out) {
    if (sc.n_lights == 0 || sc.total_lum < EPS)
        return false;
    f32 scale;
    u32 li = select_light(rng, sc, scale);
    const PhxLight &l = sc.lights[li];
    F4 flux = mk(l.power.x, l.power.y, l.power.z, 0.0f) / scale;
    F4 dir;
    f32 dist, dist_sq, ndd;
    F4 rad;
    u32 kind = l.kind;
    if (kind == PHX_LIGHT_POINT || kind == PHX_LIGHT_SPOT) {
        fill_light_vecs(pos, from(l.position), dir, dist, dist_sq);
        ndd = dot(dir, normal);
        if (ndd <= 0.0f)
            return false;
        f32 cone = 1.0f;
        if (kind == PHX_LIGHT_SPOT) {
            f32 cd = dot(-dir, normalize(from(l.direction)));
            cone = spotlight_falloff(cd, phx_cosf(l.aux.x), phx_cosf(l.aux.y));
            if (cone < EPS)
                return false;
        }
        rad = flux * cone * ndd / dist_sq;
    } else if (kind == PHX_LIGHT_DIRECTIONAL) {
        dir = normalize(-from(l.direction));
        dist = INF;
        ndd = dot(dir, normal);
        if (ndd <= 0.0f)
            return false;
        rad = flux * ndd;
    } else {
        u32 tex = phx_as_uint(l.aux.x);
        u32 start = phx_as_uint(l.aux.y);
        u32 n = phx_as_uint(l.aux.z);
        f32 area = l.aux.w;
        if (area < EPS)
            return false;
        EmissivePoint p = sample_emissive_point(rng, sc, start, n, area);
        fill_light_vecs(pos, p.pos, dir, dist, dist_sq);
        ndd = dot(dir, normal);
        if (ndd <= 0.0f)
            return false;
        F4 ef = apply_emissive_texture(sc, flux, tex, p);
        f32 cl = phx_fmaxf(dot(-dir, p.normal), 0.0f);
        rad = ef * ndd * cl / (PI * dist_sq);
    }
    if (dot(rad, rad) < EPS * EPS)
        return false;
    out.radiance = rad;
    out.dir = dir;
    out.dist = dist;
    out.n_dot_dir = ndd;
    out.light = li;
    return true;
}

/* envmap.h:9-20 */
PHX_DEV F4 sample_envmap(const SceneView &sc, F4 dir) {
    f32 theta = phx_acosf(phx_clampf(dir.z, -1.0f, 1.0f));
    f32 phi = phx_atan2f(dir.y, dir.x);
    f32 u = (phi + PI) / (2.0f * PI);
    f32 v = phx_clampf(1.0f - theta / PI, 0.0f, 1.0f);
    u = u - phx_floorf(u);
    u32 px = phx_f2u(u * (f32)sc.env_w);
    px = px < sc.env_w - 1 ? px : sc.env_w - 1;
    u32 py = phx_f2u(v * (f32)sc.env_h);
    py = py < sc.env_h - 1 ? py : sc.env_h - 1;
    u32 idx = 3 * (py * sc.env_w + px);
    return mk(sc.envmap[idx], sc.envmap[idx + 1], sc.envmap[idx + 2], 0.0f);
}

} // namespace phx

#endif /* PHX_DEVICE_H */
