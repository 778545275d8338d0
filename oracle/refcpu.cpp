/*
 * refcpu.cpp — CPU restatement of alfazet/phosphor's SPPM pass loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the CPU
 * throughput baseline.  Only the test suite, the smoke check and the
 * benchmark's CPU-baseline leg load it; the product library never links or
 * calls it.
 *
 * It restates, function by function, the reference's device code and host
 * orchestration (file:line cited at each function).  Deliberate, documented
 * choices where the reference leaves results implementation-defined or
 * nondeterministic:
 *   - elementary functions come from phosphor_amd/csrc/phx_math.h (the one
 *     header shared with the GPU build; it is pinned against libm by ULP
 *     tests), division and sqrt are IEEE, and FP contraction is off;
 *   - OpenCL vector builtins are evaluated in a fixed order:
 *     dot = ((x·x + y·y) + z·z) + w·w, length = sqrt(dot),
 *     normalize(v) = v · (1/length(v)) (zero vector stays zero);
 *   - photon deposits are ordered canonically (photon index, then bounce)
 *     instead of atomic_inc arrival order (emit_photons.cl:98);
 *   - the scatter is stable (input order kept inside a cell) instead of
 *     atomic_inc arrival order (scatter_photons.cl:13).
 * With these choices every stage of one pass is a deterministic function of
 * its inputs.
 *
 * Parity pins: PCG streams, RGBE/octahedral codecs, get_hash, scan, scatter,
 * BVH layout, closest hit and the SPPM update are checked against closed
 * forms, brute force and numpy restatements in tests/; BSDF sampling, quad,
 * point, spot and directional light sampling, NEE, photon emission,
 * Beer-Lambert attenuation of shadow rays and bilinear texture sampling
 * against closed forms and sampling statistics (tests/test_oracle_transport.py,
 * tests/test_oracle_lights_media.py).
 * Parity unpinned: the uv transform (rotation, scale, offset), the normal,
 * emission, metal/roughness, transmission and occlusion texture slots and
 * textured emissive lights (no closed-form test; no workload uses them).
 */
#include "../include/phosphor_hip.h"
#include "../phosphor_amd/csrc/phx_math.h"

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef uint32_t u32;
typedef int32_t i32;
typedef float f32;

namespace {

/* ---------------- constants.h:6-45 ---------------- */
const u32 SCAN_CHUNK = 4096;
const f32 EPS = 1e-5f;
const f32 BBOX_EPS = 1e-3f;
const f32 INF = 1e9f;
const f32 PI = 3.141592654f;
const f32 U32_MAX_F = 4294967296.0f;
const u32 NO_TEXTURE = 0xFFFFFFFFu;
const f32 AIR_IOR = 1.0f;
const u32 MAX_RAY_BOUNCES = 5;
const f32 MIN_CELL_SIZE = 0.001f;
const int BVH_STACK_SIZE = 24;
const u32 IOR_STACK_SIZE = 8;
const f32 MIN_ROUGHNESS = 0.001f;
const u32 GGX_SAMPLING_ATTEMPTS = 8;
const u32 RR_MIN_RAY_DEPTH = 2;
const u32 RR_MIN_PHOTON_DEPTH = 2;
const f32 RR_MIN_Q = 0.05f;
const f32 RR_MAX_Q = 1.0f;
const f32 GAUSS_ALPHA = 1.818f;
const f32 GAUSS_BETA = 1.953f;
const f32 GAUSS_DENOM = 0.85815f;

/* ---------------- OpenCL float4 semantics ---------------- */
struct V4 {
    f32 x, y, z, w;
};
inline V4 mk(f32 x, f32 y, f32 z, f32 w) { return V4{x, y, z, w}; }
inline V4 from(const phx_float4 &a) { return V4{a.x, a.y, a.z, a.w}; }
inline phx_float4 to(const V4 &a) {
    phx_float4 r;
    r.x = a.x;
    r.y = a.y;
    r.z = a.z;
    r.w = a.w;
    return r;
}
inline V4 operator+(V4 a, V4 b) { return V4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline V4 operator-(V4 a, V4 b) { return V4{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
inline V4 operator*(V4 a, V4 b) { return V4{a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
inline V4 operator*(f32 s, V4 a) { return V4{s * a.x, s * a.y, s * a.z, s * a.w}; }
inline V4 operator*(V4 a, f32 s) { return V4{a.x * s, a.y * s, a.z * s, a.w * s}; }
inline V4 operator/(V4 a, f32 s) { return V4{a.x / s, a.y / s, a.z / s, a.w / s}; }
inline V4 operator-(V4 a) { return V4{-a.x, -a.y, -a.z, -a.w}; }
inline f32 dot(V4 a, V4 b) { return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w; }
inline V4 cross(V4 a, V4 b) {
    return V4{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.0f};
}
inline f32 length(V4 a) { return phx_sqrtf(dot(a, a)); }
inline V4 normalize(V4 a) {
    f32 d = dot(a, a);
    if (d == 0.0f)
        return a;
    f32 inv = 1.0f / phx_sqrtf(d);
    return a * inv;
}
const V4 ZERO = {0.0f, 0.0f, 0.0f, 0.0f};
const V4 BLACK = {0.0f, 0.0f, 0.0f, 0.0f};
const V4 WHITE = {1.0f, 1.0f, 1.0f, 0.0f};

/* ---------------- random.h:16-62 ---------------- */
struct Rng {
    u32 state;
};
inline Rng pcg_seed(u32 seed) {
    Rng r;
    r.state = 0u;
    r.state = r.state * 747796405u + 2891336453u;
    r.state += seed;
    r.state = r.state * 747796405u + 2891336453u;
    return r;
}
inline u32 random_u32(Rng *rng) {
    u32 old = rng->state;
    rng->state = rng->state * 747796405u + 2891336453u;
    return ((old >> ((old >> 28u) + 4u)) ^ old) * 277803737u;
}
inline f32 random_float(Rng *rng) { return (f32)random_u32(rng) / U32_MAX_F; }

/* utils.h:59-71 (device make_tbn) */
inline void make_tbn(V4 normal, V4 *tangent, V4 *bitangent) {
    normal.w = 0.0f;
    V4 tmp = (phx_fabsf(normal.x) > 0.1f) ? mk(0.0f, 1.0f, 0.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f, 0.0f);
    *tangent = cross(normal, tmp);
    if (length(*tangent) < EPS) {
        tmp = mk(0.0f, 0.0f, 1.0f, 0.0f);
        *tangent = cross(normal, tmp);
    }
    *tangent = normalize(*tangent);
    *bitangent = cross(normal, *tangent);
}

/* random.h:40-46 */
inline V4 random_unit_vector(Rng *rng) {
    f32 z = 2.0f * random_float(rng) - 1.0f;
    f32 r = phx_sqrtf(phx_fmaxf(0.0f, 1.0f - z * z));
    f32 phi = 2.0f * PI * random_float(rng);
    return mk(r * phx_cosf(phi), r * phx_sinf(phi), z, 0.0f);
}

/* random.h:48-62 */
inline V4 random_in_unit_hemisphere(Rng *rng, V4 normal) {
    normal.w = 0.0f;
    f32 r1 = random_float(rng);
    f32 r2 = random_float(rng);
    f32 phi = 2.0f * PI * r1;
    f32 sin_theta = phx_sqrtf(r2);
    f32 cos_theta = phx_sqrtf(1.0f - r2);
    V4 hl = mk(sin_theta * phx_cosf(phi), sin_theta * phx_sinf(phi), cos_theta, 0.0f);
    V4 t, b;
    make_tbn(normal, &t, &b);
    return normalize(hl.x * t + hl.y * b + hl.z * normal);
}

/* camera.h:28-47 */
inline void make_ray(const PhxCameraParams *cam, u32 px, u32 py, u32 width, u32 height, Rng *rng, V4 *out_origin,
                     V4 *out_dir) {
    f32 u = random_float(rng) - 0.5f;
    f32 v = random_float(rng) - 0.5f;
    f32 s = ((f32)px + 0.5f + u) / (f32)width;
    f32 t = 1.0f - ((f32)py + 0.5f + v) / (f32)height;
    V4 disk = mk(0.0f, 0.0f, 0.0f, 0.0f);
    if (cam->defocus_angle > 0.0f) {
        f32 r1 = phx_sqrtf(random_float(rng));
        f32 r2 = 2.0f * PI * random_float(rng);
        f32 c = phx_cosf(r2);
        f32 sn = phx_sinf(r2);
        disk = r1 * (c * from(cam->defocus_disk_u) + sn * from(cam->defocus_disk_v));
    }
    *out_origin = from(cam->position) + disk;
    V4 dir = from(cam->lower_left_corner) + s * from(cam->horizontal) + t * from(cam->vertical) - *out_origin;
    *out_dir = normalize(dir);
}

/* ---------------- encoding.h:10-74 ---------------- */
inline u32 encode_rgbe(V4 p) {
    f32 m = phx_fmaxf(p.x, phx_fmaxf(p.y, p.z));
    if (m <= 0.0f)
        return 0u;
    int e = phx_frexp_exp(m);
    f32 scale = phx_ldexp1(e);
    u32 r = phx_f2u(p.x / scale * 256.0f);
    u32 g = phx_f2u(p.y / scale * 256.0f);
    u32 b = phx_f2u(p.z / scale * 256.0f);
    u32 E = (u32)(e + 128);
    r = std::min(r, 255u);
    g = std::min(g, 255u);
    b = std::min(b, 255u);
    return r | (g << 8) | (b << 16) | (E << 24);
}
inline V4 decode_rgbe(u32 rgbe) {
    u32 E = rgbe >> 24;
    if (E == 0)
        return mk(0.0f, 0.0f, 0.0f, 0.0f);
    f32 scale = phx_ldexp1((i32)E - 128 - 8);
    return mk((f32)(rgbe & 0xFF) * scale, (f32)((rgbe >> 8) & 0xFF) * scale, (f32)((rgbe >> 16) & 0xFF) * scale,
              0.0f);
}
inline u32 encode_oct(V4 dir) {
    dir = normalize(dir);
    f32 l = phx_fabsf(dir.x) + phx_fabsf(dir.y) + phx_fabsf(dir.z);
    f32 nx = dir.x / l, ny = dir.y / l;
    if (!(dir.z >= 0.0f)) {
        /* oct_wrap: (1 - |v.yx|) * sign(v) */
        f32 wx = (1.0f - phx_fabsf(ny)) * (nx >= 0.0f ? 1.0f : -1.0f);
        f32 wy = (1.0f - phx_fabsf(nx)) * (ny >= 0.0f ? 1.0f : -1.0f);
        nx = wx;
        ny = wy;
    }
    u32 ux = phx_f2u(phx_clampf(nx * 0.5f + 0.5f, 0.0f, 1.0f) * 65535.0f);
    u32 uy = phx_f2u(phx_clampf(ny * 0.5f + 0.5f, 0.0f, 1.0f) * 65535.0f);
    return ux | (uy << 16);
}
inline V4 decode_oct(u32 packed) {
    f32 fx = (f32)(packed & 0xFFFF) / 65535.0f;
    f32 fy = (f32)((packed >> 16) & 0xFFFF) / 65535.0f;
    fx = fx * 2.0f - 1.0f;
    fy = fy * 2.0f - 1.0f;
    V4 n = mk(fx, fy, 1.0f - phx_fabsf(fx) - phx_fabsf(fy), 0.0f);
    f32 t = phx_clampf(-n.z, 0.0f, 1.0f);
    n.x += (n.x >= 0.0f ? -t : t);
    n.y += (n.y >= 0.0f ? -t : t);
    return normalize(n);
}

/* ---------------- photon_hash.h:21-37 ---------------- */
inline u32 get_hash(V4 pos, const PhxPhotonHashData &d) {
    u32 g = d.grid_res;
    f32 fx = (pos.x - d.origin.x) / d.cell_sizes.x;
    f32 fy = (pos.y - d.origin.y) / d.cell_sizes.y;
    f32 fz = (pos.z - d.origin.z) / d.cell_sizes.z;
    if (fx < 0.0f || fy < 0.0f || fz < 0.0f || fx >= (f32)g || fy >= (f32)g || fz >= (f32)g)
        return 0;
    u32 x = (u32)fx, y = (u32)fy, z = (u32)fz;
    return g * g * x + g * y + z + 1;
}

/* Half-cell octant of a point in its cell: bit 0/1/2 set where the x/y/z
 * quotient's fraction (exact for the same rounded quotient as get_hash) is at
 * least 0.5; 0 outside the grid.  The pass orders each cell's records by
 * octant (stable), a canonical instance of the reference's in-cell order,
 * which is atomic arrival (scatter_photons.cl:13). */
inline u32 octant_of(V4 pos, const PhxPhotonHashData &d) {
    u32 g = d.grid_res;
    f32 fx = (pos.x - d.origin.x) / d.cell_sizes.x;
    f32 fy = (pos.y - d.origin.y) / d.cell_sizes.y;
    f32 fz = (pos.z - d.origin.z) / d.cell_sizes.z;
    if (fx < 0.0f || fy < 0.0f || fz < 0.0f || fx >= (f32)g || fy >= (f32)g || fz >= (f32)g)
        return 0;
    u32 x = (u32)fx, y = (u32)fy, z = (u32)fz;
    return (fx - (f32)x >= 0.5f ? 1u : 0u) | (fy - (f32)y >= 0.5f ? 2u : 0u) | (fz - (f32)z >= 0.5f ? 4u : 0u);
}

/* ---------------- hit.h ---------------- */
struct HitRecord {
    f32 t, u, v;
    u32 tri_index, mat_index;
};
struct SurfaceHit {
    V4 position, normal;
    f32 uvx, uvy;
    f32 t;
    u32 mat_index, tri_index;
    bool front_face;
};

struct Scene {
    const PhxTriangleVertices *tv;
    const PhxTriangleNormals *tn;
    const PhxTriangleTangents *tt;
    const PhxTriangleUvs *tuv;
    const u32 *tmat;
    u32 n_tris;
    const PhxBvhNode *nodes;
    const PhxTriangleVertices *etv;
    const PhxTriangleNormals *etn;
    const PhxTriangleUvs *etuv;
    const u32 *etmat;
    const PhxMaterial *materials;
    const PhxUvTransformBlock *uvt;
    const PhxTextureMeta *tex_meta;
    const uint8_t *atlas;
    const PhxLight *lights;
    u32 n_lights;
    const f32 *lum_pref_sum;
    f32 total_lum;
    V4 scene_center;
    f32 scene_radius;
    const f32 *envmap;
    u32 env_w, env_h, has_env;
};

/* hit.h:45-78 */
inline bool is_triangle_hit(const Scene &sc, u32 tri_idx, V4 origin, V4 dir, f32 t_min, f32 t_max, HitRecord *rec) {
    const PhxTriangleVertices &tv = sc.tv[tri_idx];
    V4 v0 = from(tv.v0), v1 = from(tv.v1), v2 = from(tv.v2);
    V4 e1 = v1 - v0;
    V4 e2 = v2 - v0;
    V4 h = cross(dir, e2);
    f32 a = dot(e1, h);
    if (phx_fabsf(a) < EPS)
        return false;
    f32 f = 1.0f / a;
    V4 s = origin - v0;
    f32 u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f)
        return false;
    V4 q = cross(s, e1);
    f32 v = f * dot(dir, q);
    if (v < 0.0f || u + v > 1.0f)
        return false;
    f32 t = f * dot(e2, q);
    if (t < t_min || t > t_max)
        return false;
    rec->t = t;
    rec->u = u;
    rec->v = v;
    rec->tri_index = tri_idx;
    rec->mat_index = sc.tmat[tri_idx];
    return true;
}

/* hit.h:80-133 */
inline bool bbox_hit(const PhxBvhNode &n, V4 o, V4 d, f32 t_min, f32 t_max) {
    f32 inv_dx = 1.0f / d.x;
    f32 tx0 = (n.min.x - o.x) * inv_dx;
    f32 tx1 = (n.max.x - o.x) * inv_dx;
    if (tx0 < tx1) {
        if (tx0 > t_min)
            t_min = tx0;
        if (tx1 < t_max)
            t_max = tx1;
    } else {
        if (tx1 > t_min)
            t_min = tx1;
        if (tx0 < t_max)
            t_max = tx0;
    }
    if (t_max <= t_min)
        return false;
    f32 inv_dy = 1.0f / d.y;
    f32 ty0 = (n.min.y - o.y) * inv_dy;
    f32 ty1 = (n.max.y - o.y) * inv_dy;
    if (ty0 < ty1) {
        if (ty0 > t_min)
            t_min = ty0;
        if (ty1 < t_max)
            t_max = ty1;
    } else {
        if (ty1 > t_min)
            t_min = ty1;
        if (ty0 < t_max)
            t_max = ty0;
    }
    if (t_max <= t_min)
        return false;
    f32 inv_dz = 1.0f / d.z;
    f32 tz0 = (n.min.z - o.z) * inv_dz;
    f32 tz1 = (n.max.z - o.z) * inv_dz;
    if (tz0 < tz1) {
        if (tz0 > t_min)
            t_min = tz0;
        if (tz1 < t_max)
            t_max = tz1;
    } else {
        if (tz1 > t_min)
            t_min = tz1;
        if (tz0 < t_max)
            t_max = tz0;
    }
    if (t_max <= t_min)
        return false;
    return true;
}

/* scene_intersect calls of this thread while visits are counted (the rays
 * a pass traces; the caller takes differences around each work item) */
static thread_local uint64_t tl_ray_calls = 0;

/* hit.h:135-172 (the stack holds 24 entries; a push that would overflow drops the children) */
inline bool scene_intersect(const Scene &sc, V4 origin, V4 dir, f32 t_min, f32 t_max, HitRecord *rec,
                            uint64_t *node_visits) {
    if (node_visits)
        tl_ray_calls++;
    int p = 0;
    u32 stack[BVH_STACK_SIZE];
    stack[0] = 0;
    bool found = false;
    f32 closest = t_max;
    while (p >= 0) {
        u32 index = stack[p];
        const PhxBvhNode &node = sc.nodes[index];
        p--;
        if (node_visits)
            (*node_visits)++;
        if (!bbox_hit(node, origin, dir, t_min, closest))
            continue;
        u32 tri = phx_as_uint(node.min.w);
        if (tri != 0xFFFFFFFFu) {
            HitRecord cand;
            if (is_triangle_hit(sc, tri, origin, dir, t_min, closest, &cand)) {
                closest = cand.t;
                *rec = cand;
                found = true;
            }
            continue;
        }
        if (p + 2 >= BVH_STACK_SIZE)
            continue;
        u32 right = phx_as_uint(node.max.w);
        stack[++p] = right;
        stack[++p] = index + 1;
    }
    return found;
}

/* hit.h:31-43 */
inline V4 compute_bary(f32 bx, f32 by, V4 v0, V4 v1, V4 v2) {
    f32 w0 = 1.0f - bx - by;
    f32 w1 = bx, w2 = by;
    V4 r;
    r.x = w0 * v0.x + w1 * v1.x + w2 * v2.x;
    r.y = w0 * v0.y + w1 * v1.y + w2 * v2.y;
    r.z = w0 * v0.z + w1 * v1.z + w2 * v2.z;
    r.w = 0.0f;
    return r;
}

/* hit.h:174-197 */
inline SurfaceHit process_hit(const Scene &sc, const HitRecord &rec, V4 origin, V4 dir) {
    SurfaceHit sh;
    sh.t = rec.t;
    sh.mat_index = rec.mat_index;
    sh.tri_index = rec.tri_index;
    sh.position = origin + dir * rec.t;
    f32 u = rec.u, v = rec.v;
    f32 w = 1.0f - u - v;
    const PhxTriangleNormals &tn = sc.tn[rec.tri_index];
    V4 outward = normalize(compute_bary(u, v, from(tn.n0), from(tn.n1), from(tn.n2)));
    bool front = dot(dir, outward) < 0.0f;
    sh.front_face = front;
    sh.normal = front ? outward : -outward;
    const PhxTriangleUvs &uv = sc.tuv[rec.tri_index];
    sh.uvx = w * uv.uv0.x + u * uv.uv1.x + v * uv.uv2.x;
    sh.uvy = w * uv.uv0.y + u * uv.uv1.y + v * uv.uv2.y;
    return sh;
}

/* ---------------- material.h:66-149 (textures) ---------------- */
inline V4 srgb_to_linear(V4 c) {
    return mk(phx_powf(c.x, 2.2f), phx_powf(c.y, 2.2f), phx_powf(c.z, 2.2f), c.w);
}
inline void apply_uv_transform(f32 ux, f32 uy, const PhxUvTransform &t, f32 *ox, f32 *oy) {
    f32 sx = ux * t.uv_scale.x, sy = uy * t.uv_scale.y;
    f32 c = phx_cosf(t.uv_rotation), s = phx_sinf(t.uv_rotation);
    f32 rx = c * sx - s * sy, ry = s * sx + c * sy;
    *ox = rx + t.uv_offset.x;
    *oy = ry + t.uv_offset.y;
}
inline V4 naive_sample_texture(const Scene &sc, u32 index, f32 ux, f32 uy) {
    if (index == NO_TEXTURE)
        return WHITE;
    const PhxTextureMeta &m = sc.tex_meta[index];
    u32 px = std::min(phx_f2u(ux * (f32)m.width), m.width - 1);
    u32 py = std::min(phx_f2u(uy * (f32)m.height), m.height - 1);
    u32 texel = m.atlas_offset + (py * m.width + px) * 3u;
    return mk((f32)sc.atlas[texel + 0] / 255.0f, (f32)sc.atlas[texel + 1] / 255.0f, (f32)sc.atlas[texel + 2] / 255.0f,
              1.0f);
}
inline V4 texel4(const Scene &sc, i32 idx) {
    return mk((f32)sc.atlas[idx], (f32)sc.atlas[idx + 1], (f32)sc.atlas[idx + 2], 255.0f) / 255.0f;
}
inline V4 mix4(V4 a, V4 b, f32 t) { return a + (b - a) * t; }
inline V4 sample_texture(const Scene &sc, u32 index, f32 ux, f32 uy) {
    if (index == NO_TEXTURE)
        return WHITE;
    const PhxTextureMeta &m = sc.tex_meta[index];
    i32 w = (i32)m.width, h = (i32)m.height;
    ux = ux - phx_floorf(ux);
    uy = uy - phx_floorf(uy);
    uy = 1.0f - uy;
    f32 x = phx_fminf(ux * (f32)w, (f32)(w - 1));
    f32 y = phx_fminf(uy * (f32)h, (f32)(h - 1));
    i32 x1 = std::min((i32)x, w - 1);
    i32 x2 = std::min((i32)phx_ceilf(x), w - 1);
    i32 y1 = std::min((i32)y, h - 1);
    i32 y2 = std::min((i32)phx_ceilf(y), h - 1);
    i32 ch = (i32)m.channels, off = (i32)m.atlas_offset;
    V4 q11 = texel4(sc, off + (y1 * w + x1) * ch);
    V4 q12 = texel4(sc, off + (y2 * w + x1) * ch);
    V4 q21 = texel4(sc, off + (y1 * w + x2) * ch);
    V4 q22 = texel4(sc, off + (y2 * w + x2) * ch);
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
inline V4 sample_texture_uv(const Scene &sc, u32 mat_index, u32 tex_index, f32 ux, f32 uy, u32 slot) {
    const PhxUvTransform *t = &sc.uvt[mat_index].diff + slot;
    f32 tx, ty;
    apply_uv_transform(ux, uy, *t, &tx, &ty);
    return sample_texture(sc, tex_index, tx, ty);
}

/* ---------------- shading.h ---------------- */
struct Ctx {
    V4 base_color, emissive, shading_normal;
    f32 metallic, roughness, transmission, occlusion, ior;
};
struct IorStack {
    f32 ior[IOR_STACK_SIZE];
    u32 cur;
};
inline void init_ior_stack(IorStack *s, f32 v) {
    s->cur = 0;
    s->ior[0] = v;
}
inline f32 current_ior(const IorStack *s) { return s->ior[s->cur]; }
inline f32 target_ior(const IorStack *s, f32 inside, bool front) {
    if (front)
        return inside;
    return (s->cur == 0) ? AIR_IOR : s->ior[s->cur - 1];
}
inline void update_ior_stack(IorStack *s, f32 inside, bool front) {
    if (front) {
        if (s->cur < IOR_STACK_SIZE - 1) {
            s->cur++;
            s->ior[s->cur] = inside;
        }
    } else {
        if (s->cur > 0)
            s->cur--;
    }
}
inline V4 apply_normal_map(V4 ms, V4 gn, V4 t, V4 b) {
    V4 n = mk(ms.x * 2.0f - 1.0f, ms.y * 2.0f - 1.0f, ms.z * 2.0f - 1.0f, 0.0f);
    V4 p = n.x * t + n.y * b + n.z * gn;
    f32 len = length(p);
    return (len > EPS) ? (p / len) : gn;
}
/* shading.h:74-123 */
inline Ctx evaluate_material(const Scene &sc, const PhxMaterial &mat, u32 mi, f32 ux, f32 uy, V4 gn, V4 t, V4 b) {
    Ctx c;
    c.base_color = from(mat.base_color);
    if (mat.diff_index != NO_TEXTURE)
        c.base_color = c.base_color * srgb_to_linear(sample_texture_uv(sc, mi, mat.diff_index, ux, uy, 0));
    c.emissive = from(mat.emissive);
    if (mat.emis_index != NO_TEXTURE)
        c.emissive = c.emissive * srgb_to_linear(sample_texture_uv(sc, mi, mat.emis_index, ux, uy, 1));
    c.metallic = mat.metallic;
    c.roughness = mat.roughness;
    if (mat.metal_rough_index != NO_TEXTURE) {
        V4 mr = sample_texture_uv(sc, mi, mat.metal_rough_index, ux, uy, 4);
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
    c.shading_normal = gn;
    if (mat.norm_index != NO_TEXTURE)
        c.shading_normal = apply_normal_map(sample_texture_uv(sc, mi, mat.norm_index, ux, uy, 2), gn, t, b);
    return c;
}
/* shading.h:125-133 */
inline V4 beer_lambert(V4 att, f32 dist, f32 travel) {
    if (dist < EPS)
        return WHITE;
    V4 sigma = mk(-phx_logf(phx_fmaxf(att.x, EPS)), -phx_logf(phx_fmaxf(att.y, EPS)), -phx_logf(phx_fmaxf(att.z, EPS)),
                  0.0f) /
               dist;
    return mk(phx_expf(-sigma.x * travel), phx_expf(-sigma.y * travel), phx_expf(-sigma.z * travel), 0.0f);
}

/* ---------------- bsdfs.h ---------------- */
inline V4 reflect(V4 i, V4 n) { return normalize(i - 2.0f * dot(i, n) * n); }
inline V4 safe_reflect(V4 d, V4 n, V4 g) {
    V4 r = reflect(d, n);
    if (dot(r, g) <= 0.0f) {
        r = r - dot(r, g) * g;
        r = normalize(r);
    }
    return r;
}
/* bsdfs.h:27-72 */
inline V4 ggx_sample_vndf(Rng *rng, V4 normal, V4 geom_normal, V4 view, f32 roughness) {
    f32 alpha = roughness * roughness;
    V4 t, b;
    make_tbn(normal, &t, &b);
    V4 vl = normalize(mk(dot(view, t), dot(view, b), dot(view, normal), 0.0f));
    V4 vh = normalize(mk(alpha * vl.x, alpha * vl.y, vl.z, 0.0f));
    f32 len_sq = vh.x * vh.x + vh.y * vh.y;
    V4 T1 = (len_sq > 0.0f) ? mk(-vh.y, vh.x, 0.0f, 0.0f) * (1.0f / phx_sqrtf(len_sq)) : mk(1.0f, 0.0f, 0.0f, 0.0f);
    V4 T2 = cross(vh, T1);
    f32 s = 0.5f * (1.0f + vh.z);
    for (u32 a = 0; a < GGX_SAMPLING_ATTEMPTS; a++) {
        f32 u1 = random_float(rng);
        f32 u2 = random_float(rng);
        f32 r = phx_sqrtf(u1);
        f32 phi = 2.0f * PI * u2;
        f32 t1 = r * phx_cosf(phi);
        f32 t2 = r * phx_sinf(phi);
        t2 = (1.0f - s) * phx_sqrtf(phx_fmaxf(0.0f, 1.0f - t1 * t1)) + s * t2;
        f32 nz2 = phx_fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2);
        V4 nh = t1 * T1 + t2 * T2 + phx_sqrtf(nz2) * vh;
        V4 nl = normalize(mk(alpha * nh.x, alpha * nh.y, phx_fmaxf(0.0f, nh.z), 0.0f));
        V4 ne = normalize(nl.x * t + nl.y * b + nl.z * normal);
        V4 refl = reflect(-view, ne);
        if (dot(refl, geom_normal) >= 0.0f)
            return ne;
    }
    return normal;
}
inline f32 smith_g1_ggx(f32 c, f32 alpha) {
    if (c <= 0.0f)
        return 0.0f;
    f32 a2 = alpha * alpha;
    return 2.0f * c / (c + phx_sqrtf(a2 + (1.0f - a2) * c * c));
}
/* pow(x, 5) of bsdfs.h:88,93 evaluated as ((x·x)·(x·x))·x */
inline f32 pow5(f32 x) {
    f32 x2 = x * x;
    return (x2 * x2) * x;
}
inline f32 fresnel(f32 r0, V4 i, V4 n) {
    f32 c = phx_fabsf(dot(i, n));
    return r0 + (1.0f - r0) * pow5(1.0f - c);
}
inline V4 fresnel4(V4 r0, V4 i, V4 n) {
    f32 c = phx_fabsf(dot(i, n));
    f32 p = pow5(1.0f - c);
    return r0 + (mk(1.0f, 1.0f, 1.0f, 1.0f) - r0) * p;
}
inline f32 fresnel_refracted(f32 i1, f32 i2, V4 i, V4 n) {
    f32 q = (i1 - i2) / (i1 + i2);
    return fresnel(q * q, i, n);
}
inline V4 refract(V4 i, V4 n, f32 e1, f32 e2, bool *tir) {
    f32 eta = e1 / e2;
    f32 ndi = dot(i, n);
    f32 st2 = eta * eta * (1.0f - ndi * ndi);
    if (st2 > 1.0f) {
        *tir = true;
        return reflect(i, n);
    }
    *tir = false;
    V4 r = eta * i - (eta * ndi + phx_sqrtf(1.0f - st2)) * n;
    return normalize(r);
}
inline f32 ggx_ndf(f32 ndh, f32 alpha) {
    f32 a2 = alpha * alpha;
    f32 d = ndh * ndh * (a2 - 1.0f) + 1.0f;
    return a2 / phx_fmaxf(PI * d * d, EPS);
}

const u32 BSDF_DIFFUSE = 0, BSDF_METALLIC = 1, BSDF_FRESNEL = 2, BSDF_TRANSMIT = 3;
struct Bsdf {
    V4 dir, throughput;
    u32 event;
};
/* shading.h:150-213 */
inline Bsdf sample_bsdf(Rng *rng, const Ctx &c, V4 sn, V4 gn, V4 view, IorStack *ior, bool front) {
    Bsdf s;
    V4 h = ggx_sample_vndf(rng, sn, gn, view, c.roughness);
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
        V4 refr = refract(-view, h, i1, i2, &tir);
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
    V4 pos, normal;
};
struct LightFragment {
    V4 radiance, dir;
    f32 dist, n_dot_dir;
};
inline f32 spotlight_falloff(f32 ca, f32 ci, f32 co) {
    f32 t = phx_clampf((ca - co) / (ci - co + EPS), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
inline V4 light_flux(const PhxLight &l, f32 scale) { return mk(l.power.x, l.power.y, l.power.z, 0.0f) / scale; }
inline void fill_light_vecs(V4 from_, V4 to_, V4 *dir, f32 *dist, f32 *dist_sq) {
    V4 d = to_ - from_;
    *dist_sq = phx_fmaxf(dot(d, d), EPS);
    *dist = phx_sqrtf(*dist_sq);
    *dir = d / *dist;
}
/* light_sampling.h:52-67 */
inline u32 choose_emissive_triangle(Rng *rng, const Scene &sc, u32 start, u32 n, f32 total_area) {
    if (n <= 1 || total_area <= EPS)
        return start;
    f32 r = random_float(rng) * total_area;
    f32 acc = 0.0f;
    for (u32 i = 0; i < n; i++) {
        const PhxTriangleVertices &ev = sc.etv[start + i];
        acc += 0.5f * length(cross(from(ev.v1) - from(ev.v0), from(ev.v2) - from(ev.v0)));
        if (r <= acc)
            return start + i;
    }
    return start + n - 1;
}
inline EmissivePoint sample_emissive_point(Rng *rng, const Scene &sc, u32 start, u32 n, f32 total_area) {
    EmissivePoint p;
    p.tri_idx = choose_emissive_triangle(rng, sc, start, n, total_area);
    f32 r1 = random_float(rng);
    f32 r2 = random_float(rng);
    f32 s = phx_sqrtf(r1);
    p.bu = 1.0f - s;
    p.bv = r2 * s;
    p.bw = 1.0f - p.bu - p.bv;
    const PhxTriangleVertices &ev = sc.etv[p.tri_idx];
    p.pos = p.bw * from(ev.v0) + p.bu * from(ev.v1) + p.bv * from(ev.v2);
    const PhxTriangleNormals &en = sc.etn[p.tri_idx];
    V4 nn = p.bw * from(en.n0) + p.bu * from(en.n1) + p.bv * from(en.n2);
    if (length(nn) > EPS)
        p.normal = normalize(nn);
    else
        p.normal = normalize(cross(from(ev.v1) - from(ev.v0), from(ev.v2) - from(ev.v0)));
    return p;
}
inline V4 apply_emissive_texture(const Scene &sc, V4 color, u32 tex, const EmissivePoint &p) {
    if (tex == NO_TEXTURE)
        return color;
    const PhxTriangleUvs &e = sc.etuv[p.tri_idx];
    f32 ux = p.bw * e.uv0.x + p.bu * e.uv1.x + p.bv * e.uv2.x;
    f32 uy = p.bw * e.uv0.y + p.bu * e.uv1.y + p.bv * e.uv2.y;
    u32 mi = sc.etmat[p.tri_idx];
    return color * srgb_to_linear(sample_texture_uv(sc, mi, tex, ux, uy, 1));
}
/* light_sampling.h:117-139 */
inline u32 select_light(Rng *rng, const Scene &sc, f32 *scale) {
    if (sc.n_lights == 0 || sc.total_lum < EPS) {
        *scale = 1.0f;
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
    *scale = phx_fmaxf(lum / sc.total_lum, EPS);
    return lo;
}
/* light_sampling.h:200-227 */
inline void sample_light(Rng *rng, const Scene &sc, V4 *origin, V4 *dir, V4 *power) {
    if (sc.n_lights == 0 || sc.total_lum < EPS) {
        *origin = mk(0.0f, 0.0f, 0.0f, 0.0f);
        *dir = mk(0.0f, 1.0f, 0.0f, 0.0f);
        *power = BLACK;
        return;
    }
    f32 scale;
    u32 idx = select_light(rng, sc, &scale);
    const PhxLight &l = sc.lights[idx];
    if (l.kind == PHX_LIGHT_POINT) {
        *origin = from(l.position);
        *dir = random_unit_vector(rng);
        *power = from(l.power) * (4.0f * PI);
    } else if (l.kind == PHX_LIGHT_SPOT) {
        *origin = from(l.position);
        f32 ci = phx_cosf(l.aux.x);
        f32 co = phx_cosf(l.aux.y);
        f32 r1 = random_float(rng);
        f32 r2 = random_float(rng);
        f32 ct = 1.0f - r1 * (1.0f - co);
        f32 st = phx_sqrtf(phx_fmaxf(0.0f, 1.0f - ct * ct));
        f32 phi = 2.0f * PI * r2;
        V4 ld = mk(st * phx_cosf(phi), st * phx_sinf(phi), ct, 0.0f);
        V4 t, b;
        V4 ldir = normalize(from(l.direction));
        make_tbn(ldir, &t, &b);
        *dir = normalize(ld.x * t + ld.y * b + ld.z * ldir);
        f32 fo = spotlight_falloff(ct, ci, co);
        f32 sa = 2.0f * PI * (1.0f - co);
        *power = from(l.power) * fo * sa;
    } else if (l.kind == PHX_LIGHT_DIRECTIONAL) {
        f32 r1 = sc.scene_radius * phx_sqrtf(random_float(rng));
        f32 r2 = 2.0f * PI * random_float(rng);
        V4 disk = r1 * (phx_cosf(r2) * from(l.tangent) + phx_sinf(r2) * from(l.bitangent));
        *dir = from(l.direction);
        *origin = sc.scene_center - from(l.direction) * 2.0f * sc.scene_radius + disk;
        *power = from(l.power) * PI * sc.scene_radius * sc.scene_radius;
    } else {
        u32 tex = phx_as_uint(l.aux.x);
        u32 start = phx_as_uint(l.aux.y);
        u32 n = phx_as_uint(l.aux.z);
        f32 area = l.aux.w;
        EmissivePoint p = sample_emissive_point(rng, sc, start, n, area);
        *dir = random_in_unit_hemisphere(rng, p.normal);
        *origin = p.pos + EPS * *dir;
        *power = apply_emissive_texture(sc, from(l.power), tex, p);
    }
    *power = *power / scale;
}
/* light_sampling.h:231-300 */
inline bool sample_light_fragment(Rng *rng, const Scene &sc, V4 pos, V4 normal, LightFragment *out) {
    if (sc.n_lights == 0 || sc.total_lum < EPS)
        return false;
    f32 scale;
    u32 li = select_light(rng, sc, &scale);
    const PhxLight &l = sc.lights[li];
    V4 flux = light_flux(l, scale);
    V4 dir;
    f32 dist, dist_sq, ndd;
    V4 rad;
    if (l.kind == PHX_LIGHT_POINT || l.kind == PHX_LIGHT_SPOT) {
        fill_light_vecs(pos, from(l.position), &dir, &dist, &dist_sq);
        ndd = dot(dir, normal);
        if (ndd <= 0.0f)
            return false;
        f32 cone = 1.0f;
        if (l.kind == PHX_LIGHT_SPOT) {
            f32 cd = dot(-dir, normalize(from(l.direction)));
            cone = spotlight_falloff(cd, phx_cosf(l.aux.x), phx_cosf(l.aux.y));
            if (cone < EPS)
                return false;
        }
        rad = flux * cone * ndd / dist_sq;
    } else if (l.kind == PHX_LIGHT_DIRECTIONAL) {
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
        fill_light_vecs(pos, p.pos, &dir, &dist, &dist_sq);
        ndd = dot(dir, normal);
        if (ndd <= 0.0f)
            return false;
        V4 ef = apply_emissive_texture(sc, flux, tex, p);
        f32 cl = phx_fmaxf(dot(-dir, p.normal), 0.0f);
        rad = ef * ndd * cl / (PI * dist_sq);
    }
    if (dot(rad, rad) < EPS * EPS)
        return false;
    out->radiance = rad;
    out->dir = dir;
    out->dist = dist;
    out->n_dot_dir = ndd;
    return true;
}
/* light_sampling.h:302-333 */
inline V4 shadow_transmittance(const Scene &sc, V4 origin, V4 dir, f32 dist, uint64_t *visits) {
    V4 trans = WHITE;
    f32 t_min = EPS, t_max = dist - EPS;
    for (u32 depth = 0; depth < MAX_RAY_BOUNCES; depth++) {
        HitRecord hr;
        if (!scene_intersect(sc, origin, dir, t_min, t_max, &hr, visits))
            break;
        const PhxMaterial &mat = sc.materials[hr.mat_index];
        if (mat.transmission < EPS)
            return BLACK;
        V4 bc = from(mat.base_color);
        if (mat.diff_index != NO_TEXTURE) {
            SurfaceHit sh = process_hit(sc, hr, origin, dir);
            bc = bc * srgb_to_linear(sample_texture_uv(sc, hr.mat_index, mat.diff_index, sh.uvx, sh.uvy, 0));
        }
        if (mat.att_dist > EPS)
            trans = trans * beer_lambert(from(mat.att_color), mat.att_dist, hr.t - t_min);
        trans = trans * (bc * mat.transmission);
        if (dot(trans, trans) < EPS)
            return BLACK;
        t_min = hr.t + EPS;
    }
    return trans;
}
inline V4 direct_lighting_diffuse(Rng *rng, const Scene &sc, V4 pos, V4 normal, V4 base_color, uint64_t *visits) {
    LightFragment f;
    if (!sample_light_fragment(rng, sc, pos, normal, &f))
        return BLACK;
    V4 tr = shadow_transmittance(sc, pos, f.dir, f.dist, visits);
    return f.radiance * (base_color / PI) * tr;
}
inline V4 direct_lighting_specular(Rng *rng, const Scene &sc, V4 pos, V4 normal, V4 view, V4 f0, f32 roughness,
                                   uint64_t *visits) {
    f32 ndv = dot(view, normal);
    if (ndv < EPS)
        return BLACK;
    LightFragment f;
    if (!sample_light_fragment(rng, sc, pos, normal, &f))
        return BLACK;
    if (f.n_dot_dir < EPS)
        return BLACK;
    f32 alpha = roughness * roughness;
    V4 h = normalize(f.dir + view);
    f32 ndh = phx_fmaxf(dot(normal, h), 0.0f);
    f32 ggx = ggx_ndf(ndh, alpha);
    V4 fr = fresnel4(f0, view, h);
    f32 sm = smith_g1_ggx(f.n_dot_dir, alpha) * smith_g1_ggx(ndv, alpha);
    V4 spec = f.radiance * ggx * fr * sm / (4.0f * f.n_dot_dir * ndv);
    if (dot(spec, spec) < EPS * EPS)
        return BLACK;
    V4 tr = shadow_transmittance(sc, pos, f.dir, f.dist, visits);
    return spec * tr;
}
/* envmap.h:9-20 */
inline V4 sample_envmap(const Scene &sc, V4 dir) {
    f32 theta = phx_acosf(phx_clampf(dir.z, -1.0f, 1.0f));
    f32 phi = phx_atan2f(dir.y, dir.x);
    f32 u = (phi + PI) / (2.0f * PI);
    f32 v = phx_clampf(1.0f - theta / PI, 0.0f, 1.0f);
    u = u - phx_floorf(u);
    u32 px = std::min(phx_f2u(u * (f32)sc.env_w), sc.env_w - 1);
    u32 py = std::min(phx_f2u(v * (f32)sc.env_h), sc.env_h - 1);
    u32 idx = 3 * (py * sc.env_w + px);
    return mk(sc.envmap[idx], sc.envmap[idx + 1], sc.envmap[idx + 2], 0.0f);
}

/* shared TBN block of trace_rays.cl:83-93 / emit_photons.cl:66-76 */
inline void hit_tbn(const Scene &sc, const HitRecord &rec, const SurfaceHit &sh, V4 *t, V4 *b) {
    f32 bw = 1.0f - rec.u - rec.v;
    const PhxTriangleTangents &tt = sc.tt[sh.tri_index];
    V4 tr = bw * from(tt.t0) + rec.u * from(tt.t1) + rec.v * from(tt.t2);
    V4 br = (length(tr) > EPS) ? cross(sh.normal, tr) : ZERO;
    if (length(br) > EPS) {
        *b = normalize(br);
        *t = cross(*b, sh.normal);
    } else {
        make_tbn(sh.normal, t, b);
    }
}

/* ---------------- kernels ---------------- */

/* trace_rays.cl:42-168, one pixel */
inline void trace_pixel(const Scene &sc, const PhxCameraParams *cam, u32 W, u32 H, u32 seed, u32 direct_samples,
                        u32 tid, V4 out[6], uint64_t *visits) {
    u32 px = tid % W, py = tid / W;
    Rng rng = pcg_seed(seed + tid);
    V4 origin, dir;
    make_ray(cam, px, py, W, H, &rng, &origin, &dir);
    V4 position = ZERO, normal = ZERO, throughput = WHITE, base_color = BLACK, direct = BLACK, emission = BLACK;
    IorStack ior;
    init_ior_stack(&ior, AIR_IOR);
    for (i32 depth = 0; depth < (i32)MAX_RAY_BOUNCES; depth++) {
        HitRecord rec;
        bool hit = scene_intersect(sc, origin, dir, EPS, INF, &rec, visits);
        if (!hit) {
            if (sc.has_env)
                emission = emission + throughput * sample_envmap(sc, dir);
            break;
        }
        SurfaceHit sh = process_hit(sc, rec, origin, dir);
        const PhxMaterial &mat = sc.materials[sh.mat_index];
        V4 vol = WHITE;
        f32 curr_ior = current_ior(&ior);
        if (curr_ior != AIR_IOR && mat.thickness > EPS && mat.att_dist > EPS) {
            vol = beer_lambert(from(mat.att_color), mat.att_dist, sh.t);
            throughput = throughput * vol;
        }
        V4 tg, bt;
        hit_tbn(sc, rec, sh, &tg, &bt);
        Ctx c = evaluate_material(sc, mat, sh.mat_index, sh.uvx, sh.uvy, sh.normal, tg, bt);
        emission = emission + throughput * c.emissive;
        V4 view = -dir;
        Bsdf bs = sample_bsdf(&rng, c, c.shading_normal, sh.normal, view, &ior, sh.front_face);
        if (bs.event == BSDF_DIFFUSE) {
            V4 sum = BLACK;
            for (u32 d = 0; d < direct_samples; d++)
                sum = sum + direct_lighting_diffuse(&rng, sc, sh.position, c.shading_normal, c.base_color, visits);
            emission = mk(emission.x, emission.y, emission.z, 1.0f);
            position = sh.position;
            normal = sh.normal;
            base_color = c.base_color;
            direct = throughput * c.occlusion * (sum / (f32)direct_samples);
            break;
        }
        if (bs.event == BSDF_METALLIC) {
            V4 sum = BLACK;
            for (u32 d = 0; d < direct_samples; d++)
                sum = sum + direct_lighting_specular(&rng, sc, sh.position, c.shading_normal, view, c.base_color,
                                                     c.roughness, visits);
            emission = emission + throughput * (sum / (f32)direct_samples);
        }
        if (bs.event == BSDF_FRESNEL) {
            f32 ratio = (curr_ior - c.ior) / (curr_ior + c.ior);
            f32 r0 = ratio * ratio;
            V4 f0 = mk(r0, r0, r0, 0.0f);
            V4 sum = BLACK;
            for (u32 d = 0; d < direct_samples; d++)
                sum = sum + direct_lighting_specular(&rng, sc, sh.position, c.shading_normal, view, f0, c.roughness,
                                                     visits);
            emission = emission + throughput * (sum / (f32)direct_samples);
        }
        throughput = throughput * bs.throughput;
        if ((u32)depth >= RR_MIN_RAY_DEPTH) {
            f32 q = phx_fmaxf(throughput.x, phx_fmaxf(throughput.y, throughput.z));
            q = phx_clampf(q, RR_MIN_Q, RR_MAX_Q);
            if (random_float(&rng) > q)
                break;
            throughput = throughput * (1.0f / q);
        }
        V4 side = (dot(bs.dir, sh.normal) > 0.0f) ? sh.normal : -sh.normal;
        origin = sh.position + bs.dir * EPS + side * EPS;
        dir = bs.dir;
    }
    out[0] = position;
    out[1] = normal;
    out[2] = throughput;
    out[3] = base_color;
    out[4] = direct;
    out[5] = emission;
}

struct Deposit {
    V4 pos;
    u32 power, dir;
};

/* emit_photons.cl:34-113, one photon; deposits appended to `out` */
inline u32 emit_photon(const Scene &sc, u32 tid, u32 seed, u32 max_bounces, Deposit *out, uint64_t *visits) {
    Rng rng = pcg_seed(seed + tid);
    V4 origin, dir, power;
    sample_light(&rng, sc, &origin, &dir, &power);
    IorStack ior;
    init_ior_stack(&ior, AIR_IOR);
    V4 throughput = mk(1.0f, 1.0f, 1.0f, 0.0f);
    u32 n = 0;
    for (u32 depth = 0; depth < max_bounces; depth++) {
        HitRecord rec;
        if (!scene_intersect(sc, origin, dir, EPS, INF, &rec, visits))
            return n;
        SurfaceHit sh = process_hit(sc, rec, origin, dir);
        const PhxMaterial &mat = sc.materials[sh.mat_index];
        V4 vol = WHITE;
        f32 curr_ior = current_ior(&ior);
        if (curr_ior != AIR_IOR && mat.thickness > EPS && mat.att_dist > EPS) {
            vol = beer_lambert(from(mat.att_color), mat.att_dist, sh.t);
            power = power * vol;
            throughput = throughput * vol;
        }
        V4 tg, bt;
        hit_tbn(sc, rec, sh, &tg, &bt);
        Ctx c = evaluate_material(sc, mat, sh.mat_index, sh.uvx, sh.uvy, sh.normal, tg, bt);
        if (depth >= RR_MIN_PHOTON_DEPTH) {
            f32 q = phx_fmaxf(throughput.x, phx_fmaxf(throughput.y, throughput.z));
            q = phx_clampf(q, RR_MIN_Q, RR_MAX_Q);
            if (random_float(&rng) > q)
                return n;
            f32 iq = 1.0f / q;
            throughput = throughput * iq;
            power = power * iq;
        }
        V4 view = -dir;
        Bsdf bs = sample_bsdf(&rng, c, c.shading_normal, sh.normal, view, &ior, sh.front_face);
        if (bs.event == BSDF_DIFFUSE && dot(dir, sh.normal) < 0.0f) {
            out[n].pos = sh.position;
            out[n].power = encode_rgbe(power);
            out[n].dir = encode_oct(dir);
            n++;
        }
        power = power * bs.throughput;
        throughput = throughput * bs.throughput;
        V4 side = (dot(bs.dir, sh.normal) > 0.0f) ? sh.normal : -sh.normal;
        origin = sh.position + bs.dir * EPS + side * EPS;
        dir = bs.dir;
    }
    return n;
}

/* photon_hash.h:56-103 + gather_flux.cl:24-77, one pixel */
inline void gather_pixel(u32 tid, const phx_float4 *hp_pos, const phx_float4 *hp_normal,
                         const phx_float4 *hp_throughput, const phx_float4 *hp_base_color,
                         const phx_float4 *hp_direct, const phx_float4 *hp_emission, const phx_float4 *ppos,
                         const u32 *ppow, const u32 *pdir, const u32 *cell_start, const u32 *cell_end,
                         const PhxPhotonHashData &data, f32 alpha, phx_float4 *flux, f32 *photon_weight,
                         phx_float4 *radiance, uint64_t *visits) {
    V4 ge = from(hp_emission[tid]);
    V4 rsf = from(radiance[tid]);
    V4 r = mk(ge.x, ge.y, ge.z, 0.0f);
    if (ge.w == 0.0f) {
        radiance[tid] = to(rsf + r);
        return;
    }
    V4 dr = from(hp_direct[tid]);
    r.x += dr.x;
    r.y += dr.y;
    r.z += dr.z;
    radiance[tid] = to(rsf + r);
    V4 fsf = from(flux[tid]);
    f32 wsf = photon_weight[tid];
    f32 r_sq = fsf.w;
    if (r_sq < EPS) {
        f32 hx = data.cell_sizes.x * 0.5f, hy = data.cell_sizes.y * 0.5f, hz = data.cell_sizes.z * 0.5f;
        r_sq = hx * hx + hy * hy + hz * hz;
    }
    V4 nf = BLACK;
    f32 found = 0.0f;
    V4 pos = from(hp_pos[tid]), sn = from(hp_normal[tid]);
    uint64_t v = 0;
    for (u32 i = 0; i < 27; i++) {
        i32 cx = (i32)(i % 3) - 1, cy = (i32)((i / 3) % 3) - 1, cz = (i32)((i / 9) % 3) - 1;
        V4 p2 = pos;
        p2.x += (f32)cx * data.cell_sizes.x;
        p2.y += (f32)cy * data.cell_sizes.y;
        p2.z += (f32)cz * data.cell_sizes.z;
        u32 h = get_hash(p2, data);
        if (h == 0)
            continue;
        u32 s = cell_start[h], e = cell_end[h];
        if (s == e)
            continue;
        for (u32 k = s; k < e; k++) {
            v++;
            V4 d = pos - from(ppos[k]);
            f32 dsq = dot(d, d);
            if (dsq >= r_sq)
                continue;
            V4 pd = decode_oct(pdir[k]);
            if (dot(pd, sn) > 0.0f)
                continue;
            f32 w = GAUSS_ALPHA * (1.0f - (1.0f - phx_expf(-GAUSS_BETA * dsq / (2.0f * r_sq))) / GAUSS_DENOM);
            nf = nf + w * decode_rgbe(ppow[k]);
            found += w;
        }
    }
    if (visits)
        *visits += v;
    if (found > 0.0f) {
        f32 npw = wsf + alpha * found;
        f32 ratio = npw / (wsf + found);
        f32 nr = r_sq * ratio;
        V4 albedo = from(hp_base_color[tid]);
        V4 contrib = from(hp_throughput[tid]) * (albedo / PI) * nf;
        V4 tmp = (fsf + contrib) * ratio;
        flux[tid] = to(mk(tmp.x, tmp.y, tmp.z, nr));
        photon_weight[tid] = npw;
    }
}

/* ---------------- host: bounding_box.h + bvh.cpp ---------------- */
struct BBox {
    f32 mn[3], mx[3];
};
inline BBox tri_bbox(const PhxTriangleVertices &t) {
    BBox b;
    const phx_float4 *v[3] = {&t.v0, &t.v1, &t.v2};
    b.mn[0] = b.mn[1] = b.mn[2] = INF;
    b.mx[0] = b.mx[1] = b.mx[2] = -INF;
    for (int k = 0; k < 3; k++) {
        f32 c[3] = {v[k]->x, v[k]->y, v[k]->z};
        for (int a = 0; a < 3; a++) {
            b.mn[a] = std::min(b.mn[a], c[a]);
            b.mx[a] = std::max(b.mx[a], c[a]);
        }
    }
    for (int a = 0; a < 3; a++) {
        b.mn[a] -= BBOX_EPS;
        b.mx[a] += BBOX_EPS;
    }
    return b;
}
inline BBox merge(const BBox &a, const BBox &b) {
    BBox r;
    for (int k = 0; k < 3; k++) {
        r.mn[k] = std::min(a.mn[k], b.mn[k]);
        r.mx[k] = std::max(a.mx[k], b.mx[k]);
    }
    return r;
}
inline u32 longest_axis(const BBox &b) {
    f32 x = b.mx[0] - b.mn[0], y = b.mx[1] - b.mn[1], z = b.mx[2] - b.mn[2];
    if (x > y) {
        if (x > z)
            return 0;
    } else {
        if (y > z)
            return 1;
    }
    return 2;
}
struct TmpNode {
    BBox bbox;
    i32 tri;
    u32 left, right;
};
u32 build_tmp(const std::vector<BBox> &tb, std::vector<u32> &idx, u32 start, u32 end, std::vector<TmpNode> &tmp) {
    u32 id = (u32)tmp.size();
    tmp.push_back({});
    if (end == start + 1) {
        tmp[id].tri = (i32)idx[start];
        tmp[id].bbox = tb[idx[start]];
        tmp[id].left = tmp[id].right = 0;
        return id;
    }
    BBox bb = tb[idx[start]];
    for (u32 i = start + 1; i < end; i++)
        bb = merge(bb, tb[idx[i]]);
    tmp[id].bbox = bb;
    tmp[id].tri = -1;
    u32 axis = longest_axis(bb);
    std::sort(idx.begin() + start, idx.begin() + end, [&](u32 i, u32 j) { return tb[i].mn[axis] < tb[j].mn[axis]; });
    u32 mid = start + (end - start) / 2;
    u32 l = build_tmp(tb, idx, start, mid, tmp);
    u32 r = build_tmp(tb, idx, mid, end, tmp);
    tmp[id].left = l;
    tmp[id].right = r;
    return id;
}
void flatten(const std::vector<TmpNode> &tmp, u32 ti, std::vector<PhxBvhNode> &out) {
    u32 id = (u32)out.size();
    out.push_back({});
    const BBox &b = tmp[ti].bbox;
    out[id].min.x = b.mn[0];
    out[id].min.y = b.mn[1];
    out[id].min.z = b.mn[2];
    out[id].max.x = b.mx[0];
    out[id].max.y = b.mx[1];
    out[id].max.z = b.mx[2];
    out[id].min.w = phx_as_float((u32)tmp[ti].tri);
    out[id].max.w = phx_as_float(tmp[ti].right);
    if (tmp[ti].tri == -1) {
        flatten(tmp, tmp[ti].left, out);
        out[id].max.w = phx_as_float((u32)out.size());
        flatten(tmp, tmp[ti].right, out);
    }
}

} // namespace

/* ======================================================================
 * extern "C" surface used by the tests and the benchmark's CPU leg
 * ====================================================================== */
extern "C" {

int ref_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void ref_set_num_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* primitives (KAT hooks) */
u32 ref_pcg_seed(u32 seed) { return pcg_seed(seed).state; }
u32 ref_random_u32(u32 *state) {
    Rng r{*state};
    u32 v = random_u32(&r);
    *state = r.state;
    return v;
}
f32 ref_random_float(u32 *state) {
    Rng r{*state};
    f32 v = random_float(&r);
    *state = r.state;
    return v;
}
u32 ref_encode_rgbe(const f32 *p) { return encode_rgbe(mk(p[0], p[1], p[2], p[3])); }
void ref_decode_rgbe(u32 v, f32 *out) {
    V4 r = decode_rgbe(v);
    out[0] = r.x, out[1] = r.y, out[2] = r.z, out[3] = r.w;
}
u32 ref_encode_oct(const f32 *d) { return encode_oct(mk(d[0], d[1], d[2], d[3])); }
void ref_decode_oct(u32 v, f32 *out) {
    V4 r = decode_oct(v);
    out[0] = r.x, out[1] = r.y, out[2] = r.z, out[3] = r.w;
}
u32 ref_get_hash(const f32 *p, const PhxPhotonHashData *d) { return get_hash(mk(p[0], p[1], p[2], p[3]), *d); }
void ref_math(int which, const f32 *x, const f32 *y, f32 *out, u32 n) {
    for (u32 i = 0; i < n; i++) {
        switch (which) {
        case 0: out[i] = phx_sinf(x[i]); break;
        case 1: out[i] = phx_cosf(x[i]); break;
        case 2: out[i] = phx_expf(x[i]); break;
        case 3: out[i] = phx_logf(x[i]); break;
        case 4: out[i] = phx_powf(x[i], y[i]); break;
        case 5: out[i] = phx_atan2f(x[i], y[i]); break;
        case 6: out[i] = phx_acosf(x[i]); break;
        default: out[i] = 0.0f;
        }
    }
}

/* bvh.cpp:74-84 */
u32 ref_build_bvh(const PhxTriangleVertices *tv, u32 n, PhxBvhNode *out_nodes) {
    if (n == 0)
        return 0;
    std::vector<BBox> tb(n);
    for (u32 i = 0; i < n; i++)
        tb[i] = tri_bbox(tv[i]);
    std::vector<u32> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    std::vector<TmpNode> tmp;
    tmp.reserve(2 * (size_t)n);
    build_tmp(tb, idx, 0, n, tmp);
    std::vector<PhxBvhNode> nodes;
    nodes.reserve(2 * (size_t)n - 1);
    flatten(tmp, 0, nodes);
    std::memcpy(out_nodes, nodes.data(), nodes.size() * sizeof(PhxBvhNode));
    return (u32)nodes.size();
}

/* hit.h:135-172 as a query: returns 1 on hit with out = {t, u, v, tri, mat} */
int ref_scene_intersect(const PhxBvhNode *nodes, const PhxTriangleVertices *tv, const u32 *tmat, u32 n_tris,
                        const f32 *o, const f32 *d, f32 t_min, f32 t_max, f32 *out5) {
    Scene sc{};
    sc.tv = tv;
    sc.tmat = tmat;
    sc.n_tris = n_tris;
    sc.nodes = nodes;
    HitRecord rec{};
    bool hit = scene_intersect(sc, mk(o[0], o[1], o[2], o[3]), mk(d[0], d[1], d[2], d[3]), t_min, t_max, &rec, nullptr);
    if (hit) {
        out5[0] = rec.t;
        out5[1] = rec.u;
        out5[2] = rec.v;
        out5[3] = phx_as_float(rec.tri_index);
        out5[4] = phx_as_float(rec.mat_index);
    }
    return hit ? 1 : 0;
}

/* shading.h:150-213 as a query for statistical tests: material on a surface
 * with normal n, incoming view direction; returns event, writes dir and
 * throughput and the new RNG state. */
u32 ref_sample_bsdf(u32 *state, const PhxMaterial *mat, const f32 *normal, const f32 *view, int front_face,
                    f32 *out_dir, f32 *out_thr) {
    Rng rng{*state};
    Ctx c;
    c.base_color = from(mat->base_color);
    c.emissive = from(mat->emissive);
    c.metallic = mat->metallic;
    c.roughness = phx_clampf(mat->roughness, MIN_ROUGHNESS, 1.0f);
    c.transmission = mat->transmission;
    c.occlusion = 1.0f;
    c.ior = mat->ior;
    V4 n = mk(normal[0], normal[1], normal[2], 0.0f);
    c.shading_normal = n;
    IorStack ior;
    init_ior_stack(&ior, AIR_IOR);
    Bsdf b = sample_bsdf(&rng, c, n, n, mk(view[0], view[1], view[2], 0.0f), &ior, front_face != 0);
    *state = rng.state;
    out_dir[0] = b.dir.x, out_dir[1] = b.dir.y, out_dir[2] = b.dir.z, out_dir[3] = b.dir.w;
    out_thr[0] = b.throughput.x, out_thr[1] = b.throughput.y, out_thr[2] = b.throughput.z,
    out_thr[3] = b.throughput.w;
    return b.event;
}

/* shading.h:150-213 with an explicit IOR stack (ior[8], depth cur), read and
 * written back, so that entering, leaving and nested media can be checked. */
u32 ref_sample_bsdf_ior(u32 *state, const PhxMaterial *mat, const f32 *normal, const f32 *view, int front_face,
                        f32 *ior_stack, u32 *ior_cur, f32 *out_dir, f32 *out_thr) {
    Rng rng{*state};
    Ctx c;
    c.base_color = from(mat->base_color);
    c.emissive = from(mat->emissive);
    c.metallic = mat->metallic;
    c.roughness = phx_clampf(mat->roughness, MIN_ROUGHNESS, 1.0f);
    c.transmission = mat->transmission;
    c.occlusion = 1.0f;
    c.ior = mat->ior;
    V4 n = mk(normal[0], normal[1], normal[2], 0.0f);
    c.shading_normal = n;
    IorStack ior;
    for (u32 k = 0; k < IOR_STACK_SIZE; k++)
        ior.ior[k] = ior_stack[k];
    ior.cur = std::min(*ior_cur, IOR_STACK_SIZE - 1);
    Bsdf b = sample_bsdf(&rng, c, n, n, mk(view[0], view[1], view[2], 0.0f), &ior, front_face != 0);
    *state = rng.state;
    for (u32 k = 0; k < IOR_STACK_SIZE; k++)
        ior_stack[k] = ior.ior[k];
    *ior_cur = ior.cur;
    out_dir[0] = b.dir.x, out_dir[1] = b.dir.y, out_dir[2] = b.dir.z, out_dir[3] = b.dir.w;
    out_thr[0] = b.throughput.x, out_thr[1] = b.throughput.y, out_thr[2] = b.throughput.z,
    out_thr[3] = b.throughput.w;
    return b.event;
}

/* photon_hash.h:41-50 */
void ref_build_photon_hash_info(const PhxBvhNode *root, u32 grid_res, PhxPhotonHashData *out) {
    PhxPhotonHashData d;
    std::memset(&d, 0, sizeof(d));
    d.grid_res = grid_res;
    d.origin.x = root->min.x;
    d.origin.y = root->min.y;
    d.origin.z = root->min.z;
    d.origin.w = 0.0f;
    d.cell_sizes.x = std::max((root->max.x - root->min.x) / (f32)grid_res, MIN_CELL_SIZE);
    d.cell_sizes.y = std::max((root->max.y - root->min.y) / (f32)grid_res, MIN_CELL_SIZE);
    d.cell_sizes.z = std::max((root->max.z - root->min.z) / (f32)grid_res, MIN_CELL_SIZE);
    d.cell_sizes.w = 0.0f;
    *out = d;
}

/* hash_photons.cl */
void ref_hash_photons(const phx_float4 *pos, u32 n, const PhxPhotonHashData *d, u32 *hashes, u32 *counts) {
    for (u32 i = 0; i < n; i++) {
        u32 h = get_hash(from(pos[i]), *d);
        hashes[i] = h;
        counts[h]++;
    }
}
/* scan_small/big/add + copy (renderer.cpp:134-147): exclusive scan */
void ref_scan_cells(const u32 *counts, u32 n, u32 *start, u32 *end) {
    u32 s = 0;
    for (u32 i = 0; i < n; i++) {
        start[i] = s;
        end[i] = s;
        s += counts[i];
    }
}
/* Within every cell [cstart, cend), a stable partition of the records by
 * octant (octant_of): the photon map's order becomes (cell, octant, input
 * order). */
static void order_cells_by_octant(phx_float4 *pos, u32 *pw, u32 *dir, const u32 *cstart, const u32 *cend, u32 n_cells,
                                  const PhxPhotonHashData &d) {
#pragma omp parallel
    {
        std::vector<phx_float4> tp;
        std::vector<u32> tw, td, oc;
#pragma omp for schedule(dynamic, 256)
        for (int64_t h = 0; h < (int64_t)n_cells; h++) {
            u32 a = cstart[h], b = cend[h];
            if (b - a < 2)
                continue;
            u32 m = b - a;
            tp.assign(pos + a, pos + b);
            tw.assign(pw + a, pw + b);
            td.assign(dir + a, dir + b);
            oc.resize(m);
            u32 cnt[9] = {0};
            for (u32 i = 0; i < m; i++) {
                oc[i] = octant_of(from(tp[i]), d);
                cnt[oc[i] + 1]++;
            }
            for (u32 q = 1; q < 9; q++)
                cnt[q] += cnt[q - 1];
            for (u32 i = 0; i < m; i++) {
                u32 dst = a + cnt[oc[i]]++;
                pos[dst] = tp[i];
                pw[dst] = tw[i];
                dir[dst] = td[i];
            }
        }
    }
}

/* scatter_photons.cl with a stable (input-order) cursor walk */
void ref_sort_photons(const phx_float4 *pos_in, const u32 *pow_in, const u32 *dir_in, const u32 *hashes,
                      u32 *cell_end, u32 n, phx_float4 *pos_out, u32 *pow_out, u32 *dir_out) {
    for (u32 i = 0; i < n; i++) {
        u32 dst = cell_end[hashes[i]]++;
        pos_out[dst] = pos_in[i];
        pow_out[dst] = pow_in[i];
        dir_out[dst] = dir_in[i];
    }
}
void ref_gather_flux(const phx_float4 *hp_pos, const phx_float4 *hp_normal, const phx_float4 *hp_throughput,
                     const phx_float4 *hp_base_color, const phx_float4 *hp_direct, const phx_float4 *hp_emission,
                     u32 n_pixels, const phx_float4 *ppos, const u32 *ppow, const u32 *pdir, const u32 *cell_start,
                     const u32 *cell_end, const PhxPhotonHashData *data, f32 alpha, phx_float4 *flux,
                     f32 *photon_weight, phx_float4 *radiance) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n_pixels; i++)
        gather_pixel((u32)i, hp_pos, hp_normal, hp_throughput, hp_base_color, hp_direct, hp_emission, ppos, ppow,
                     pdir, cell_start, cell_end, *data, alpha, flux, photon_weight, radiance, nullptr);
}

/* ---------------- pass driver (renderer.cpp:22-161) ---------------- */
struct RefRenderer {
    PhxRenderArgs args;
    Scene sc;
    /* owned scene copies */
    std::vector<PhxTriangleVertices> tv, etv;
    std::vector<PhxTriangleNormals> tn, etn;
    std::vector<PhxTriangleTangents> tt;
    std::vector<PhxTriangleUvs> tuv, etuv;
    std::vector<u32> tmat, etmat;
    std::vector<PhxMaterial> mats;
    std::vector<PhxUvTransformBlock> uvt;
    std::vector<PhxLight> lights;
    std::vector<f32> lum;
    std::vector<PhxTextureMeta> tex_meta;
    std::vector<uint8_t> atlas;
    std::vector<f32> env;
    std::vector<PhxBvhNode> nodes;
    PhxCameraParams cam;
    PhxPhotonHashData hash;
    u32 bucket_count, photon_buf_size, n_pixels;
    Rng rng;
    u32 rounds = 0;
    uint64_t total_photons = 0;
    u32 last_count = 0, last_n = 0;
    std::vector<phx_float4> hp[6];
    std::vector<phx_float4> flux, radiance;
    std::vector<f32> weight;
    std::vector<phx_float4> ppos, spos;
    std::vector<u32> ppow, pdir, spow, sdir, hashes, counts, cstart, cend;
    double stage_ms[4] = {0, 0, 0, 0};
    uint64_t visits_trace = 0, visits_emit = 0, visits_gather = 0;
    uint64_t rays_trace = 0, rays_emit = 0;
    bool count_visits = false;
    /* photon map in (cell, octant, deposit) order, the library's pass; off:
       (cell, deposit), the order of phx_sort_photons and the stage-level ABI */
    bool octant_order = true;
};

void ref_renderer_destroy(RefRenderer *r) { delete r; }

RefRenderer *ref_renderer_create(const PhxRenderArgs *a, const PhxSceneDesc *s) {
    RefRenderer *r = new RefRenderer();
    r->args = *a;
    u32 n = s->n_triangles, ne = s->n_e_triangles;
    r->tv.assign(s->tri_verts, s->tri_verts + n);
    r->tn.assign(s->tri_normals, s->tri_normals + n);
    r->tt.assign(s->tri_tangents, s->tri_tangents + n);
    r->tuv.assign(s->tri_uvs, s->tri_uvs + n);
    r->tmat.assign(s->tri_mat_index, s->tri_mat_index + n);
    r->etv.assign(s->etri_verts, s->etri_verts + ne);
    r->etn.assign(s->etri_normals, s->etri_normals + ne);
    r->etuv.assign(s->etri_uvs, s->etri_uvs + ne);
    r->etmat.assign(s->etri_mat_index, s->etri_mat_index + ne);
    r->mats.assign(s->materials, s->materials + s->n_materials);
    r->uvt.assign(s->uv_transforms, s->uv_transforms + s->n_materials);
    r->lights.assign(s->lights, s->lights + s->n_lights);
    if (s->n_textures) {
        r->tex_meta.assign(s->tex_meta, s->tex_meta + s->n_textures);
        r->atlas.assign(s->tex_atlas, s->tex_atlas + s->tex_atlas_bytes);
    } else {
        r->tex_meta.resize(1);
        std::memset(r->tex_meta.data(), 0, sizeof(PhxTextureMeta));
        r->atlas.assign(1, 0);
    }
    bool has_env = s->envmap && s->envmap_width && s->envmap_height;
    if (has_env)
        r->env.assign(s->envmap, s->envmap + 3ull * s->envmap_width * s->envmap_height);
    else
        r->env.assign(1, 0.0f);
    /* scene_data.cpp:14-23 */
    f32 running = 0.0f;
    for (auto &l : r->lights) {
        f32 lum = l.power.x * 0.2126f + l.power.y * 0.7152f + l.power.z * 0.0722f;
        running += std::max(lum, 0.0f);
        r->lum.push_back(running);
    }
    if (r->lum.empty())
        r->lum.push_back(0.0f);
    /* bvh.cpp */
    if (s->bvh_nodes && s->n_bvh_nodes) {
        r->nodes.assign(s->bvh_nodes, s->bvh_nodes + s->n_bvh_nodes);
    } else {
        r->nodes.resize(2 * (size_t)n - 1);
        ref_build_bvh(r->tv.data(), n, r->nodes.data());
    }
    r->cam = *s->camera;
    Scene &sc = r->sc;
    sc.tv = r->tv.data();
    sc.tn = r->tn.data();
    sc.tt = r->tt.data();
    sc.tuv = r->tuv.data();
    sc.tmat = r->tmat.data();
    sc.n_tris = n;
    sc.nodes = r->nodes.data();
    sc.etv = r->etv.data();
    sc.etn = r->etn.data();
    sc.etuv = r->etuv.data();
    sc.etmat = r->etmat.data();
    sc.materials = r->mats.data();
    sc.uvt = r->uvt.data();
    sc.tex_meta = r->tex_meta.data();
    sc.atlas = r->atlas.data();
    sc.lights = r->lights.data();
    sc.n_lights = (u32)r->lights.size();
    sc.lum_pref_sum = r->lum.data();
    sc.total_lum = r->lights.empty() ? 0.0f : r->lum.back();
    /* shading.cpp:45-58 scene center and radius from the BVH root bbox */
    const PhxBvhNode &root = r->nodes[0];
    f32 cx = 0.5f * (root.min.x + root.max.x), cy = 0.5f * (root.min.y + root.max.y),
        cz = 0.5f * (root.min.z + root.max.z);
    sc.scene_center = mk(cx, cy, cz, 0.0f);
    f32 rad = 0.0f;
    for (int c = 0; c < 8; c++) {
        f32 px = (c & 1) ? root.max.x : root.min.x;
        f32 py = (c & 2) ? root.max.y : root.min.y;
        f32 pz = (c & 4) ? root.max.z : root.min.z;
        f32 dx = px - cx, dy = py - cy, dz = pz - cz;
        rad = std::max(rad, phx_sqrtf(dx * dx + dy * dy + dz * dz));
    }
    sc.scene_radius = rad;
    sc.envmap = r->env.data();
    sc.env_w = has_env ? s->envmap_width : 0;
    sc.env_h = has_env ? s->envmap_height : 0;
    sc.has_env = has_env ? 1 : 0;
    /* renderer.cpp:37-56 */
    ref_build_photon_hash_info(&root, a->grid_res, &r->hash);
    r->bucket_count = a->grid_res * a->grid_res * a->grid_res + 1;
    r->photon_buf_size = a->photons_per_round * a->max_photon_bounces;
    r->n_pixels = a->width * a->height;
    for (int k = 0; k < 6; k++)
        r->hp[k].assign(r->n_pixels, phx_float4{0, 0, 0, 0});
    r->flux.assign(r->n_pixels, phx_float4{0, 0, 0, 0});
    if (a->initial_radius > 0.0f)
        for (auto &f : r->flux)
            f.w = a->initial_radius * a->initial_radius;
    r->radiance.assign(r->n_pixels, phx_float4{0, 0, 0, 0});
    r->weight.assign(r->n_pixels, 0.0f);
    r->counts.assign(r->bucket_count, 0);
    r->cstart.assign(r->bucket_count, 0);
    r->cend.assign(r->bucket_count, 0);
    r->rng = pcg_seed(a->seed);
    return r;
}

void ref_renderer_count_visits(RefRenderer *r, int on) { r->count_visits = on != 0; }
void ref_renderer_octant_order(RefRenderer *r, int on) { r->octant_order = on != 0; }

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

/* Trace only the given pixel range of a pass with `seed` (for bounded CPU samples). */
void ref_trace_range(RefRenderer *r, u32 seed, u32 first, u32 count) {
    const Scene &sc = r->sc;
    uint64_t visits = 0, rays = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : visits, rays)
    for (int64_t i = first; i < (int64_t)first + count; i++) {
        V4 o[6];
        uint64_t v = 0, c0 = tl_ray_calls;
        trace_pixel(sc, &r->cam, r->args.width, r->args.height, seed, r->args.direct_samples, (u32)i, o,
                    r->count_visits ? &v : nullptr);
        visits += v;
        rays += tl_ray_calls - c0;
        for (int k = 0; k < 6; k++)
            r->hp[k][i] = to(o[k]);
    }
    r->visits_trace += visits;
    r->rays_trace += rays;
}

/* Emit photons [first, first+count) of a pass with `seed`; returns deposits. */
u32 ref_emit_range(RefRenderer *r, u32 seed, u32 first, u32 count, phx_float4 *pos_out, u32 *pow_out, u32 *dir_out,
                   u32 max_out) {
    u32 B = r->args.max_photon_bounces;
    std::vector<Deposit> dep((size_t)count * B);
    std::vector<u32> cnt(count);
    uint64_t visits = 0, rays = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : visits, rays)
    for (int64_t i = 0; i < (int64_t)count; i++) {
        uint64_t v = 0, c0 = tl_ray_calls;
        cnt[i] = emit_photon(r->sc, first + (u32)i, seed, B, &dep[(size_t)i * B], r->count_visits ? &v : nullptr);
        visits += v;
        rays += tl_ray_calls - c0;
    }
    r->visits_emit += visits;
    r->rays_emit += rays;
    u32 total = 0;
    for (u32 i = 0; i < count; i++)
        for (u32 k = 0; k < cnt[i]; k++) {
            if (total < max_out && pos_out) {
                pos_out[total] = to(dep[(size_t)i * B + k].pos);
                pow_out[total] = dep[(size_t)i * B + k].power;
                dir_out[total] = dep[(size_t)i * B + k].dir;
            }
            total++;
        }
    return total;
}

/* One SPPM round (renderer.cpp:64-78).  Returns 0, or PHX_ERR_NO_PHOTONS. */
int ref_renderer_run_pass(RefRenderer *r) {
    const PhxRenderArgs &a = r->args;
    u32 seed = random_u32(&r->rng);
    double t0 = now_ms();
    ref_trace_range(r, seed, 0, r->n_pixels);
    double t1 = now_ms();
    /* emit_photons (renderer.cpp:97-121): batches with a global offset */
    r->ppos.resize(r->photon_buf_size);
    r->ppow.resize(r->photon_buf_size);
    r->pdir.resize(r->photon_buf_size);
    u32 per_batch = std::min(a.photons_per_round, a.max_photons_per_batch ? a.max_photons_per_batch : 1u);
    uint64_t count = 0;
    for (u32 off = 0; off < a.photons_per_round; off += per_batch) {
        u32 to_emit = std::min(per_batch, a.photons_per_round - off);
        u32 room = count < r->photon_buf_size ? (u32)(r->photon_buf_size - count) : 0u;
        u32 got = ref_emit_range(r, seed, off, to_emit, r->ppos.data() + std::min<uint64_t>(count, r->photon_buf_size),
                                 r->ppow.data() + std::min<uint64_t>(count, r->photon_buf_size),
                                 r->pdir.data() + std::min<uint64_t>(count, r->photon_buf_size), room);
        count += got;
    }
    u32 n = (u32)std::min<uint64_t>(count, r->photon_buf_size);
    r->last_count = (u32)count;
    r->last_n = n;
    double t2 = now_ms();
    if (n == 0)
        return PHX_ERR_NO_PHOTONS;
    /* hash_and_sort_photons (renderer.cpp:123-153) */
    std::fill(r->counts.begin(), r->counts.end(), 0u);
    r->hashes.resize(n);
    ref_hash_photons(r->ppos.data(), n, &r->hash, r->hashes.data(), r->counts.data());
    ref_scan_cells(r->counts.data(), r->bucket_count, r->cstart.data(), r->cend.data());
    r->spos.resize(n);
    r->spow.resize(n);
    r->sdir.resize(n);
    ref_sort_photons(r->ppos.data(), r->ppow.data(), r->pdir.data(), r->hashes.data(), r->cend.data(), n,
                     r->spos.data(), r->spow.data(), r->sdir.data());
    if (r->octant_order)
        order_cells_by_octant(r->spos.data(), r->spow.data(), r->sdir.data(), r->cstart.data(), r->cend.data(),
                              r->bucket_count, r->hash);
    double t3 = now_ms();
    /* gather_flux (renderer.cpp:155-161) */
    uint64_t visits = 0;
    bool cv = r->count_visits;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : visits)
    for (int64_t i = 0; i < (int64_t)r->n_pixels; i++) {
        uint64_t v = 0;
        gather_pixel((u32)i, r->hp[0].data(), r->hp[1].data(), r->hp[2].data(), r->hp[3].data(), r->hp[4].data(),
                     r->hp[5].data(), r->spos.data(), r->spow.data(), r->sdir.data(), r->cstart.data(),
                     r->cend.data(), r->hash, a.sppm_alpha, r->flux.data(), r->weight.data(), r->radiance.data(),
                     cv ? &v : nullptr);
        visits += v;
    }
    r->visits_gather += visits;
    double t4 = now_ms();
    r->stage_ms[0] = t1 - t0;
    r->stage_ms[1] = t2 - t1;
    r->stage_ms[2] = t3 - t2;
    r->stage_ms[3] = t4 - t3;
    r->rounds++;
    r->total_photons += a.photons_per_round;
    return 0;
}

/* sample_light (light_sampling.h:200-227) of photon `tid` with pass seed `seed`:
 * writes origin, dir, power (4 floats each); returns the RNG state after. */
u32 ref_renderer_sample_light(RefRenderer *r, u32 seed, u32 tid, f32 *o4, f32 *d4, f32 *p4) {
    Rng rng = pcg_seed(seed + tid);
    V4 o, d, p;
    sample_light(&rng, r->sc, &o, &d, &p);
    std::memcpy(o4, &o, 16);
    std::memcpy(d4, &d, 16);
    std::memcpy(p4, &p, 16);
    return rng.state;
}

/* camera.h:28-47 for pixel `tid` of a pass with `seed` (trace_rays.cl:49):
 * the camera ray and the PCG state after it. */
u32 ref_renderer_camera_ray(RefRenderer *r, u32 seed, u32 tid, f32 *o4, f32 *d4) {
    Rng rng = pcg_seed(seed + tid);
    V4 o, d;
    make_ray(&r->cam, tid % r->args.width, tid / r->args.width, r->args.width, r->args.height, &rng, &o, &d);
    std::memcpy(o4, &o, 16);
    std::memcpy(d4, &d, 16);
    return rng.state;
}

/* emit_photons.cl:34-113 for photons [first, first+count) of a pass with
 * `seed`: deposits of each photon in bounce order, and per-photon counts. */
u32 ref_emit_photons_counted(RefRenderer *r, u32 seed, u32 first, u32 count, u32 *counts, phx_float4 *pos_out,
                             u32 *pow_out, u32 *dir_out, u32 max_out) {
    u32 B = r->args.max_photon_bounces;
    std::vector<Deposit> dep((size_t)count * B);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)count; i++)
        counts[i] = emit_photon(r->sc, first + (u32)i, seed, B, &dep[(size_t)i * B], nullptr);
    u32 total = 0;
    for (u32 i = 0; i < count; i++)
        for (u32 k = 0; k < counts[i]; k++, total++)
            if (total < max_out) {
                pos_out[total] = to(dep[(size_t)i * B + k].pos);
                pow_out[total] = dep[(size_t)i * B + k].power;
                dir_out[total] = dep[(size_t)i * B + k].dir;
            }
    return total;
}

/* Skip a round's seed without rendering (keeps the seed schedule aligned). */
u32 ref_renderer_next_seed(RefRenderer *r) { return random_u32(&r->rng); }

void ref_renderer_stage_ms(RefRenderer *r, double *out4) {
    for (int k = 0; k < 4; k++)
        out4[k] = r->stage_ms[k];
}
void ref_renderer_visits(RefRenderer *r, uint64_t *out3) {
    out3[0] = r->visits_trace;
    out3[1] = r->visits_emit;
    out3[2] = r->visits_gather;
}
/* scene_intersect calls (camera, shadow and photon rays) while counting is on */
void ref_renderer_rays(RefRenderer *r, uint64_t *out2) {
    out2[0] = r->rays_trace;
    out2[1] = r->rays_emit;
}
u32 ref_renderer_last_photon_count(RefRenderer *r) { return r->last_count; }
u32 ref_renderer_rounds_done(RefRenderer *r) { return r->rounds; }
void ref_renderer_hash_info(RefRenderer *r, PhxPhotonHashData *out) { *out = r->hash; }
u32 ref_renderer_bvh_nodes(RefRenderer *r, PhxBvhNode *out, u32 cap) {
    u32 n = (u32)std::min<size_t>(cap, r->nodes.size());
    if (out)
        std::memcpy(out, r->nodes.data(), n * sizeof(PhxBvhNode));
    return (u32)r->nodes.size();
}
void ref_renderer_scene_sphere(RefRenderer *r, f32 *center4, f32 *radius) {
    center4[0] = r->sc.scene_center.x;
    center4[1] = r->sc.scene_center.y;
    center4[2] = r->sc.scene_center.z;
    center4[3] = 0.0f;
    *radius = r->sc.scene_radius;
}
void ref_renderer_read_state(RefRenderer *r, phx_float4 *flux, f32 *weight, phx_float4 *radiance) {
    std::memcpy(flux, r->flux.data(), r->n_pixels * sizeof(phx_float4));
    std::memcpy(weight, r->weight.data(), r->n_pixels * sizeof(f32));
    std::memcpy(radiance, r->radiance.data(), r->n_pixels * sizeof(phx_float4));
}
void ref_renderer_write_state(RefRenderer *r, const phx_float4 *flux, const f32 *weight, const phx_float4 *radiance) {
    std::memcpy(r->flux.data(), flux, r->n_pixels * sizeof(phx_float4));
    std::memcpy(r->weight.data(), weight, r->n_pixels * sizeof(f32));
    std::memcpy(r->radiance.data(), radiance, r->n_pixels * sizeof(phx_float4));
}
void ref_renderer_read_hitpoints(RefRenderer *r, phx_float4 *hp6) {
    for (u32 i = 0; i < r->n_pixels; i++)
        for (int k = 0; k < 6; k++)
            hp6[(size_t)i * 6 + k] = r->hp[k][i];
}
/* unsorted deposits in canonical order */
u32 ref_renderer_read_deposits(RefRenderer *r, phx_float4 *pos, u32 *pw, u32 *dir) {
    if (pos) {
        std::memcpy(pos, r->ppos.data(), r->last_n * sizeof(phx_float4));
        std::memcpy(pw, r->ppow.data(), r->last_n * sizeof(u32));
        std::memcpy(dir, r->pdir.data(), r->last_n * sizeof(u32));
    }
    return r->last_n;
}
u32 ref_renderer_read_photon_map(RefRenderer *r, phx_float4 *pos, u32 *pw, u32 *dir, u32 *cs, u32 *ce) {
    if (pos) {
        std::memcpy(pos, r->spos.data(), r->last_n * sizeof(phx_float4));
        std::memcpy(pw, r->spow.data(), r->last_n * sizeof(u32));
        std::memcpy(dir, r->sdir.data(), r->last_n * sizeof(u32));
    }
    if (cs) {
        std::memcpy(cs, r->cstart.data(), r->bucket_count * sizeof(u32));
        std::memcpy(ce, r->cend.data(), r->bucket_count * sizeof(u32));
    }
    return r->last_n;
}
/* output.cpp:52-83 before tonemapping: linear estimate per pixel (RGB) */
void ref_renderer_estimate(RefRenderer *r, f32 *rgb) {
    f32 inv_ph = r->total_photons > 0 ? 1.0f / (f32)r->total_photons : 0.0f;
    f32 inv_rd = r->rounds > 0 ? 1.0f / (f32)r->rounds : 0.0f;
    for (u32 p = 0; p < r->n_pixels; p++) {
        phx_float4 f = r->flux[p];
        f32 area = PI * f.w;
        f32 ir = 0, ig = 0, ib = 0;
        if (area > 0.0f) {
            f32 pa = inv_ph / area;
            ir = f.x * pa;
            ig = f.y * pa;
            ib = f.z * pa;
        }
        const phx_float4 &rd = r->radiance[p];
        rgb[3 * p + 0] = rd.x * inv_rd + ir;
        rgb[3 * p + 1] = rd.y * inv_rd + ig;
        rgb[3 * p + 2] = rd.z * inv_rd + ib;
    }
}

} /* extern "C" */
