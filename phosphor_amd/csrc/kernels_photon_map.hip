/*
 * kernels_photon_map.hip — spatial-grid photon map: hash, scan, scatter/sort,
 * and the per-pixel flux gather with the SPPM update.
 *
 * scan: one kernel replaces scan_small_chunks + scan_big_chunks +
 *   scan_add_offsets + the cell_start→cell_end copy (renderer.cpp:134-147).
 *   256-thread tiles of 4096 elements; a thread scans 16 consecutive values,
 *   a wave64 shuffle scan combines thread totals, LDS carries combine the 4
 *   waves, and tiles are chained with a decoupled look-back over a
 *   (flag, value) status word per tile.  Tiles are numbered by an atomic
 *   ticket, so a tile only ever waits on tiles that started before it.  There
 *   is no 4096² element limit.
 * sort: stable LSD radix sort of the cell keys (8-bit digits), the payload
 *   being the input index.  Digit ranks inside a tile come from wave64
 *   ballot matching (peers with the same digit) plus per-wave running counts
 *   in LDS, so equal keys keep their input order; the last pass moves the
 *   24-byte photon records.  Result: the scatter of scatter_photons.cl with
 *   records inside a cell in input order.
 * gather: gather_flux.cl + photon_hash.h:56-103, one lane per pixel (body in
 *   pass_bodies.h).
 */
#include "pass_bodies.h"

#include <mutex>

namespace phx {

/* ---------------- fill ---------------- */
__global__ void k_fill_u32(u32 *p, u32 v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride)
        p[i] = v;
}

hipError_t launch_fill_u32(u32 *p, u32 v, size_t n, hipStream_t s) {
    if (n == 0)
        return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096)
        blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((u32)blocks), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_copy16(phx_float4 *__restrict__ dst, const phx_float4 *__restrict__ src,
                                                size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t stride = (size_t)gridDim.x * 256;
    for (; i < n; i += stride)
        dst[i] = src[i];
}

hipError_t launch_copy16(phx_float4 *dst, const phx_float4 *src, size_t n, hipStream_t s) {
    if (n == 0)
        return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 16384)
        blocks = 16384;
    hipLaunchKernelGGL(k_copy16, dim3((u32)blocks), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

/* Round seeds of a pass (renderer.cpp:65, random_u32 of the host PCG), one
 * per scene, drawn on the device so a pass needs no host input. */
__global__ void k_next_round_seed(u32 *state, u32 *seed, u32 n) {
    u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        seed[i] = random_u32(state[i]);
}

hipError_t launch_next_round_seed(u32 *state, u32 *seed, u32 n, hipStream_t s) {
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_next_round_seed, dim3((n + 63) / 64), dim3(64), 0, s, state, seed, n);
    return hipGetLastError();
}

/* info[0] = the pass's deposit count, info[1] += 1 when it is zero */
__global__ void k_note_count(const u32 *count, u32 *info) {
    u32 c = *count;
    info[0] = c;
    if (c == 0)
        info[1] += 1u;
}

hipError_t launch_note_count(const u32 *count, u32 *info, hipStream_t s) {
    hipLaunchKernelGGL(k_note_count, dim3(1), dim3(1), 0, s, count, info);
    return hipGetLastError();
}

/* ---------------- scan ---------------- */
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 16;
constexpr unsigned long long FLAG_AGG = 1ull << 32;
constexpr unsigned long long FLAG_PREFIX = 2ull << 32;

__device__ __forceinline__ u32 pad_idx(u32 i) { return i + (i >> 4); }

__device__ __forceinline__ u32 wave_incl_scan(u32 x, u32 lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        u32 y = __shfl_up(x, d, 64);
        if (lane >= (u32)d)
            x += y;
    }
    return x;
}

__device__ __forceinline__ u32 wave_sum(u32 x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        x += __shfl_xor(x, d, 64);
    return x;
}

/* With d_n set, the length is min(*d_n, n) and the grid is sized for n:
 * tiles past the device length take a ticket and leave (no tile waits on a
 * later ticket). */
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan(const u32 *__restrict__ in, u32 n, const u32 *d_n,
                                                      u32 *__restrict__ out, u32 *__restrict__ out2,
                                                      u32 *__restrict__ total, unsigned long long *status,
                                                      u32 *ticket) {
    __shared__ u32 s[SCAN_TILE + SCAN_TILE / 16];
    __shared__ u32 s_wave[SCAN_BLOCK / 64];
    __shared__ u32 s_tile, s_prefix;
    u32 t = threadIdx.x, lane = t & 63u, w = t >> 6;
    if (d_n) {
        u32 c = *d_n;
        n = c < n ? c : n;
    }
    u32 n_tiles = (u32)(((size_t)n + SCAN_TILE - 1) / SCAN_TILE);
    if (t == 0)
        s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    u32 tile = s_tile;
    if (tile >= n_tiles)
        return;
    u32 base = tile * SCAN_TILE;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        u32 i = (u32)k * SCAN_BLOCK + t;
        u32 g = base + i;
        s[pad_idx(i)] = g < n ? in[g] : 0u;
    }
    __syncthreads();
    u32 v[SCAN_ITEMS];
    u32 sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = s[pad_idx(t * SCAN_ITEMS + k)];
        sum += v[k];
    }
    u32 incl = wave_incl_scan(sum, lane);
    if (lane == 63)
        s_wave[w] = incl;
    __syncthreads();
    u32 wpre = 0, btotal = 0;
#pragma unroll
    for (u32 i = 0; i < SCAN_BLOCK / 64; i++) {
        u32 c = s_wave[i];
        wpre += (i < w) ? c : 0u;
        btotal += c;
    }
    u32 texcl = wpre + incl - sum;
    if (w == 0) {
        u32 excl = 0;
        if (tile == 0) {
            if (lane == 0)
                __hip_atomic_store(&status[0], FLAG_PREFIX | btotal, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0)
                __hip_atomic_store(&status[tile], FLAG_AGG | btotal, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int pred = (int)tile - 1;
            for (;;) {
                int idx = pred - (int)lane;
                unsigned long long st_ =
                    (idx >= 0) ? __hip_atomic_load(&status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                               : FLAG_PREFIX;
                u32 flag = (u32)(st_ >> 32);
                unsigned long long pmask = __ballot(flag == 2u);
                unsigned long long zmask = __ballot(flag == 0u);
                int first = pmask ? (__ffsll((long long)pmask) - 1) : 64;
                unsigned long long need = (first >= 63) ? ~0ull : ((2ull << first) - 1ull);
                if (zmask & need)
                    continue;
                u32 val = ((int)lane <= first) ? (u32)st_ : 0u;
                excl += wave_sum(val);
                if (first < 64)
                    break;
                pred -= 64;
            }
            if (lane == 0)
                __hip_atomic_store(&status[tile], FLAG_PREFIX | (u32)(excl + btotal), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) {
            s_prefix = excl;
            if (total && tile == n_tiles - 1)
                *total = excl + btotal;
        }
    }
    __syncthreads();
    u32 run = s_prefix + texcl;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        s[pad_idx(t * SCAN_ITEMS + k)] = run;
        run += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        u32 i = (u32)k * SCAN_BLOCK + t;
        u32 g = base + i;
        if (g < n) {
            u32 r = s[pad_idx(i)];
            out[g] = r;
            if (out2)
                out2[g] = r;
        }
    }
}

__global__ void k_scan_init(u32 *ticket, unsigned long long *status, u32 n_tiles, u32 *total) {
    u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        *ticket = 0u;
        if (total)
            *total = 0u;
    }
    for (; i < n_tiles; i += gridDim.x * blockDim.x)
        status[i] = 0ull;
}

size_t scan_workspace_bytes(u32 n) {
    size_t tiles = ((size_t)n + SCAN_TILE - 1) / SCAN_TILE;
    return 256 + align_up(tiles * sizeof(unsigned long long), 256);
}

hipError_t launch_scan(const u32 *in, u32 n, u32 *out, u32 *out2, u32 *total, void *ws, hipStream_t s,
                       const u32 *d_n) {
    u32 n_tiles = (u32)(((size_t)n + SCAN_TILE - 1) / SCAN_TILE);
    u32 *ticket = (u32 *)ws;
    unsigned long long *status = (unsigned long long *)((char *)ws + 256);
    u32 init_blocks = n_tiles ? (n_tiles + 255) / 256 : 1;
    if (init_blocks > 1024)
        init_blocks = 1024;
    hipLaunchKernelGGL(k_scan_init, dim3(init_blocks), dim3(256), 0, s, ticket, status, n_tiles, total);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || n == 0)
        return e;
    hipLaunchKernelGGL(k_scan, dim3(n_tiles), dim3(SCAN_BLOCK), 0, s, in, n, d_n, out, out2, total, status, ticket);
    return hipGetLastError();
}

/* ---------------- scan scratch without allocation ---------------- */
/* Workspace of phx_scan_cells_noalloc: the ticket (256 B) and one 8-byte
   look-back status word per 4096-cell tile, 64 KiB in all.  Each GPU the
   library runs on holds its own copy until the process ends.  Host code uses
   the address hipGetSymbolAddress returns for the current device, never the
   array's name (which is a host-side stand-in). */
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCAN_SCRATCH_BYTES];

hipError_t noalloc_scan_scratch(void **out) {
    constexpr int MAX_DEVICES = 64;
    static std::mutex mu;
    static void *ptr[MAX_DEVICES] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    if (dev < 0 || dev >= MAX_DEVICES)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        ptr[dev] = p;
    }
    *out = ptr[dev];
    return hipSuccess;
}

/* ---------------- hash (hash_photons.cl) ---------------- */
__global__ __launch_bounds__(256) void k_hash(const phx_float4 *__restrict__ pos, u32 n, const u32 *d_n, u32 max_n,
                                              PhxPhotonHashData d, u32 *__restrict__ hashes,
                                              u32 *__restrict__ counts, u32 octants, u32 stride) {
    if (d_n) {
        u32 c = *d_n;
        n = c < max_n ? c : max_n;
    }
    for (u32 i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        phx_float4 p = pos[(size_t)i * stride];
        u32 h = octants ? get_hash_octant(p.x, p.y, p.z, d) : get_hash(p.x, p.y, p.z, d);
        hashes[i] = h;
        if (counts)
            atomicAdd(&counts[h], 1u);
    }
}

hipError_t launch_hash(const phx_float4 *pos, u32 n, const PhxPhotonHashData &d, u32 *hashes, u32 *counts,
                       hipStream_t s, bool octants) {
    if (n == 0)
        return hipSuccess;
    u32 blocks = (n + 255) / 256;
    if (blocks > 8192)
        blocks = 8192;
    hipLaunchKernelGGL(k_hash, dim3(blocks), dim3(256), 0, s, pos, n, (const u32 *)nullptr, 0u, d, hashes, counts,
                       octants ? 1u : 0u, 1u);
    return hipGetLastError();
}

hipError_t launch_hash_dev(const phx_float4 *pos, const u32 *d_n, u32 max_n, const PhxPhotonHashData &d, u32 *hashes,
                           u32 *counts, hipStream_t s, bool octants, u32 pos_stride) {
    u32 blocks = (max_n + 255) / 256;
    if (blocks > 8192)
        blocks = 8192;
    if (blocks == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_hash, dim3(blocks), dim3(256), 0, s, pos, 0u, d_n, max_n, d, hashes, counts,
                       octants ? 1u : 0u, pos_stride);
    return hipGetLastError();
}

/* ---------------- scatter (scatter_photons.cl semantics) ---------------- */
__global__ __launch_bounds__(256) void k_scatter_atomic(const phx_float4 *__restrict__ pos_in,
                                                        const u32 *__restrict__ pow_in, const u32 *__restrict__ dir_in,
                                                        const u32 *__restrict__ hashes, u32 *cell_end, u32 n,
                                                        phx_float4 *__restrict__ pos_out, u32 *__restrict__ pow_out,
                                                        u32 *__restrict__ dir_out) {
    u32 i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    u32 dst = atomicAdd(&cell_end[hashes[i]], 1u);
    pos_out[dst] = pos_in[i];
    pow_out[dst] = pow_in[i];
    dir_out[dst] = dir_in[i];
}

hipError_t launch_scatter_atomic(const phx_float4 *pos_in, const u32 *pow_in, const u32 *dir_in, const u32 *hashes,
                                 u32 *cell_end, u32 n, phx_float4 *pos_out, u32 *pow_out, u32 *dir_out,
                                 hipStream_t s) {
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_scatter_atomic, dim3((n + 255) / 256), dim3(256), 0, s, pos_in, pow_in, dir_in, hashes,
                       cell_end, n, pos_out, pow_out, dir_out);
    return hipGetLastError();
}

/* ---------------- stable radix sort by cell ---------------- */
constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;
constexpr int SORT_BLOCK = 256;
constexpr int SORT_ITEMS = 16;
constexpr u32 SORT_TILE = SORT_BLOCK * SORT_ITEMS; /* 4096 */
constexpr int SORT_WAVES = SORT_BLOCK / 64;

/* Item count of a sort: n, or min(*d_n, n) when the count lives on the device
 * (the grid is then sized for n and the kernels loop over the tiles that
 * exist). */
__device__ __forceinline__ u32 sort_count(u32 n, const u32 *d_n) {
    if (!d_n)
        return n;
    u32 c = *d_n;
    return c < n ? c : n;
}

__global__ __launch_bounds__(SORT_BLOCK) void k_radix_hist(const u32 *__restrict__ keys, u32 n, const u32 *d_n,
                                                           u32 shift, u32 *__restrict__ hist) {
    __shared__ u32 h[RADIX];
    u32 t = threadIdx.x;
    n = sort_count(n, d_n);
    u32 n_tiles = (u32)(((size_t)n + SORT_TILE - 1) / SORT_TILE);
    for (u32 tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        h[t] = 0u;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < SORT_ITEMS; k++) {
            u32 e = tile * SORT_TILE + (u32)k * SORT_BLOCK + t;
            if (e < n)
                atomicAdd(&h[(keys[e] >> shift) & (RADIX - 1)], 1u);
        }
        __syncthreads();
        hist[(size_t)t * n_tiles + tile] = h[t];
        __syncthreads();
    }
}

/* Element order inside a tile is (wave, item, lane): element
 * tile·4096 + wave·1024 + item·64 + lane, so ranks computed item by item
 * within a wave, then offset by the counts of earlier waves, are stable. */
__global__ __launch_bounds__(SORT_BLOCK) void k_radix_scatter(
    const u32 *__restrict__ keys_in, const u32 *__restrict__ vals_in, u32 n, const u32 *d_n, u32 shift,
    const u32 *__restrict__ offs, u32 *__restrict__ keys_out, u32 *__restrict__ vals_out, int final_pass,
    const phx_float4 *__restrict__ pos_in, const u32 *__restrict__ pow_in, const u32 *__restrict__ dir_in,
    phx_float4 *__restrict__ pos_out, u32 *__restrict__ pow_out, u32 *__restrict__ dir_out, u32 *__restrict__ perm,
    const phx_float4 *__restrict__ rec_in, phx_float4 *__restrict__ dirpow_out) {
    __shared__ u32 cnt[SORT_WAVES][RADIX];
    u32 t = threadIdx.x, lane = t & 63u, w = t >> 6;
    n = sort_count(n, d_n);
    u32 n_tiles = (u32)(((size_t)n + SORT_TILE - 1) / SORT_TILE);
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    for (u32 tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        for (u32 i = t; i < SORT_WAVES * RADIX; i += SORT_BLOCK)
            (&cnt[0][0])[i] = 0u;
        __syncthreads();
        u32 key[SORT_ITEMS], val[SORT_ITEMS], rank[SORT_ITEMS];
        u32 ebase = tile * SORT_TILE + w * (SORT_ITEMS * 64);
#pragma unroll
        for (int k = 0; k < SORT_ITEMS; k++) {
            u32 e = ebase + (u32)k * 64 + lane;
            bool valid = e < n;
            key[k] = valid ? keys_in[e] : 0u;
            val[k] = valid ? (vals_in ? vals_in[e] : e) : 0u;
            u32 d = (key[k] >> shift) & (RADIX - 1);
            unsigned long long peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < RADIX_BITS; b++) {
                bool bit = (d >> b) & 1u;
                unsigned long long m = __ballot(bit);
                peers &= bit ? m : ~m;
            }
            u32 below = (u32)__popcll(peers & lt_mask);
            u32 c = cnt[w][d];
            rank[k] = c + below;
            if (valid && below == 0)
                cnt[w][d] = c + (u32)__popcll(peers);
        }
        __syncthreads();
        {
            /* cnt[w][d] ← global offset of digit d for this tile + counts of earlier waves */
            u32 d = t;
            u32 run = offs[(size_t)d * n_tiles + tile];
#pragma unroll
            for (int ww = 0; ww < SORT_WAVES; ww++) {
                u32 c = cnt[ww][d];
                cnt[ww][d] = run;
                run += c;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < SORT_ITEMS; k++) {
            u32 e = ebase + (u32)k * 64 + lane;
            if (e >= n)
                continue;
            u32 d = (key[k] >> shift) & (RADIX - 1);
            u32 dst = cnt[w][d] + rank[k];
            if (final_pass) {
                u32 src = val[k];
                u32 pw, dr;
                if (rec_in) { /* one 32-byte record per deposit */
                    F4 p = ld(rec_in + 2 * (size_t)src), q = ld(rec_in + 2 * (size_t)src + 1);
                    st(pos_out + dst, p);
                    pw = phx_as_uint(q.x);
                    dr = phx_as_uint(q.y);
                } else {
                    pos_out[dst] = pos_in[src];
                    pw = pow_in[src];
                    dr = dir_in[src];
                }
                pow_out[dst] = pw;
                dir_out[dst] = dr;
                if (dirpow_out) {
                    F4 dv = decode_oct(dr);
                    st(dirpow_out + dst, mk(dv.x, dv.y, dv.z, phx_as_float(pw)));
                }
                if (perm)
                    perm[dst] = src;
                if (final_pass > 1) /* the sorted keys are wanted too */
                    keys_out[dst] = key[k];
            } else {
                keys_out[dst] = key[k];
                vals_out[dst] = val[k];
            }
        }
        __syncthreads();
    }
}

/* 256 digits × n_tiles histogram entries, the length of the offset scan */
__global__ void k_sort_hist_len(u32 n, const u32 *d_n, u32 *len) {
    n = sort_count(n, d_n);
    *len = (u32)(((size_t)n + SORT_TILE - 1) / SORT_TILE) * RADIX;
}

/* cell_start[h] = number of sorted keys below h (the exclusive scan of the
 * cell counts, scan_small_chunks.cl, found by binary search instead) */
__global__ __launch_bounds__(256) void k_cell_start_search(const u32 *__restrict__ keys, u32 n, const u32 *d_n,
                                                           u32 n_cells, u32 *__restrict__ cell_start, u32 key_shift) {
    n = sort_count(n, d_n);
    for (u32 h = blockIdx.x * 256 + threadIdx.x; h < n_cells; h += gridDim.x * 256) {
        u32 lo = 0, hi = n;
        u32 kh = h << key_shift;
        while (lo < hi) {
            u32 mid = lo + (hi - lo) / 2;
            if (keys[mid] < kh)
                lo = mid + 1;
            else
                hi = mid;
        }
        cell_start[h] = lo;
    }
}

/* oct_start[8h + s] = index of the first record of cell h in octant s or a
 * later octant (the cell's end after its last octant), for every cell h that
 * holds records, from the sorted octant keys (8h + s): each entry of such a
 * cell is written once, by the record that starts the octant run it belongs
 * to; entries of empty cells are left alone (the gather never reads them). */
__global__ __launch_bounds__(256) void k_octant_starts(const u32 *__restrict__ keys, u32 n, const u32 *d_n,
                                                       u32 *__restrict__ oct_start) {
    n = sort_count(n, d_n);
    for (u32 i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        u32 key = keys[i];
        u32 h = key >> 3, o = key & 7u;
        u32 prev = i > 0 ? keys[i - 1] : 0xFFFFFFFFu;
        u32 lo = (i == 0 || (prev >> 3) != h) ? 0u : (prev == key ? 8u : (prev & 7u) + 1u);
        u32 *row = oct_start + (size_t)h * 8u;
        for (u32 q = lo; q <= o && lo < 8u; q++)
            row[q] = i;
        if (i + 1 == n || (keys[i + 1] >> 3) != h)
            for (u32 q = o + 1; q < 8u; q++)
                row[q] = i + 1;
    }
}

__global__ __launch_bounds__(256) void k_cell_end(const u32 *__restrict__ cell_start, u32 *__restrict__ cell_end,
                                                  u32 n_cells, u32 n, const u32 *d_n) {
    n = sort_count(n, d_n);
    for (u32 h = blockIdx.x * 256 + threadIdx.x; h < n_cells; h += gridDim.x * 256)
        cell_end[h] = (h + 1 < n_cells) ? cell_start[h + 1] : n;
}

static u32 key_bits(u32 n_cells) {
    u32 maxkey = n_cells > 0 ? n_cells - 1 : 0;
    u32 bits = 0;
    while (bits < 32 && (maxkey >> bits) != 0)
        bits++;
    return bits < 1 ? 1 : bits;
}

size_t sort_workspace_bytes(u32 n, u32 n_cells) {
    (void)n_cells;
    size_t tiles = ((size_t)n + SORT_TILE - 1) / SORT_TILE;
    size_t b = 256;                                           /* device scan length */
    b += 4 * align_up((size_t)n * sizeof(u32), 256);        /* keys/vals ping-pong */
    b += 2 * align_up(tiles * RADIX * sizeof(u32), 256);     /* hist, offs */
    b += align_up(scan_workspace_bytes((u32)(tiles * RADIX)), 256);
    return b;
}

/* Persistent grid for a device-side count: enough blocks to fill the GPU,
 * each looping over tiles. */
constexpr u32 SORT_DEV_BLOCKS = 2048;

hipError_t launch_sort(const phx_float4 *pos_in, const u32 *pow_in, const u32 *dir_in, const u32 *hashes,
                       const u32 *cell_start, u32 *cell_end, u32 n, u32 n_cells, phx_float4 *pos_out, u32 *pow_out,
                       u32 *dir_out, void *ws, hipStream_t s, const u32 *d_n, u32 *perm_out,
                       u32 *cell_start_out, u32 *oct_start_out, const phx_float4 *rec_in, phx_float4 *dirpow_out) {
    hipError_t e;
    const u32 *sorted_keys = hashes; /* n == 0: nothing is read */
    if (n > 0) {
        u32 n_tiles = (u32)(((size_t)n + SORT_TILE - 1) / SORT_TILE);
        u32 grid = n_tiles;
        if (d_n && grid > SORT_DEV_BLOCKS)
            grid = SORT_DEV_BLOCKS;
        char *p = (char *)ws;
        u32 *hist_len = (u32 *)p;
        p += 256;
        u32 *kbuf[2], *vbuf[2];
        kbuf[0] = (u32 *)p;
        p += align_up((size_t)n * sizeof(u32), 256);
        kbuf[1] = (u32 *)p;
        p += align_up((size_t)n * sizeof(u32), 256);
        vbuf[0] = (u32 *)p;
        p += align_up((size_t)n * sizeof(u32), 256);
        vbuf[1] = (u32 *)p;
        p += align_up((size_t)n * sizeof(u32), 256);
        u32 *hist = (u32 *)p;
        p += align_up((size_t)n_tiles * RADIX * sizeof(u32), 256);
        u32 *offs = (u32 *)p;
        p += align_up((size_t)n_tiles * RADIX * sizeof(u32), 256);
        void *scan_ws = p;
        if (d_n) {
            hipLaunchKernelGGL(k_sort_hist_len, dim3(1), dim3(1), 0, s, n, d_n, hist_len);
            if ((e = hipGetLastError()) != hipSuccess)
                return e;
        }
        u32 passes = (key_bits(oct_start_out ? n_cells * 8u : n_cells) + RADIX_BITS - 1) / RADIX_BITS;
        const u32 *kin = hashes;
        const u32 *vin = nullptr;
        for (u32 pass = 0; pass < passes; pass++) {
            u32 shift = pass * RADIX_BITS;
            bool last = pass + 1 == passes;
            hipLaunchKernelGGL(k_radix_hist, dim3(grid), dim3(SORT_BLOCK), 0, s, kin, n, d_n, shift, hist);
            if ((e = hipGetLastError()) != hipSuccess)
                return e;
            if ((e = launch_scan(hist, n_tiles * RADIX, offs, nullptr, nullptr, scan_ws, s,
                                 d_n ? hist_len : nullptr)) != hipSuccess)
                return e;
            u32 *kout = kbuf[pass & 1], *vout = vbuf[pass & 1];
            hipLaunchKernelGGL(k_radix_scatter, dim3(grid), dim3(SORT_BLOCK), 0, s, kin, vin, n, d_n, shift, offs,
                               kout, vout, last ? (cell_start_out || oct_start_out ? 2 : 1) : 0, pos_in, pow_in,
                               dir_in, pos_out, pow_out, dir_out, perm_out, rec_in, dirpow_out);
            if ((e = hipGetLastError()) != hipSuccess)
                return e;
            kin = kout;
            vin = vout;
        }
        sorted_keys = kin;
    }
    if (cell_start_out && n_cells > 0) {
        u32 blocks = (n_cells + 255) / 256;
        if (blocks > 4096)
            blocks = 4096;
        hipLaunchKernelGGL(k_cell_start_search, dim3(blocks), dim3(256), 0, s, sorted_keys, n, d_n, n_cells,
                           cell_start_out, oct_start_out ? 3u : 0u);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
        cell_start = cell_start_out;
    }
    if (oct_start_out && n > 0) {
        u32 blocks = (n + 255) / 256;
        if (blocks > 8192)
            blocks = 8192;
        hipLaunchKernelGGL(k_octant_starts, dim3(blocks), dim3(256), 0, s, sorted_keys, n, d_n, oct_start_out);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    if (n_cells > 0) {
        u32 blocks = (n_cells + 255) / 256;
        if (blocks > 4096)
            blocks = 4096;
        hipLaunchKernelGGL(k_cell_end, dim3(blocks), dim3(256), 0, s, cell_start, cell_end, n_cells, n, d_n);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    return hipSuccess;
}

/* ---------------- gather (gather_flux.cl + photon_hash.h:56-103) ---------------- */
/* One wave per 8×8 pixel tile: neighbouring hitpoints query the same grid
 * cells, so the lanes of a wave read the same photon records (one memory
 * transaction serves many lanes) and run loops of similar length; tiles are
 * assigned XCD by XCD so each L2 holds the photons of one image region.  Per
 * pixel the arithmetic and the order of the sums are those of
 * photon_hash.h:65-100. */
template <bool DECODED> __global__ __launch_bounds__(256) void k_gather(GatherArgs a, u32 width, u32 tiles_x) {
    u32 tid;
    if (tiles_x == 0) { /* no image shape known: consecutive pixels */
        tid = blockIdx.x * 256 + threadIdx.x;
    } else {
        u32 wave = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
        u32 lane = threadIdx.x & 63u;
        u32 tx = wave % tiles_x, ty = wave / tiles_x;
        u32 x = tx * 8 + (lane & 7u), y = ty * 8 + (lane >> 3);
        if (x >= width)
            return;
        tid = y * width + x;
    }
    if (tid >= a.n_pixels)
        return;
    gather_pixel<DECODED>(a, tid, a.d, a.alpha, 0u);
}

/* decode_oct / decode_rgbe of every stored photon record, once per pass, for
 * the gather (the same values it would decode at every visit) */
/* dirpow[i] = (decode_oct(pdir[i]).xyz, as_float(ppow[i])): the direction
 * the gather tests decoded once per pass, the power still packed (decoded
 * only for photons that pass both tests) */
__global__ __launch_bounds__(256) void k_decode_photons(const u32 *__restrict__ ppow, const u32 *__restrict__ pdir,
                                                        const u32 *__restrict__ d_n, u32 max_n,
                                                        phx_float4 *__restrict__ dirpow) {
    u32 n = d_n ? *d_n : max_n;
    n = n < max_n ? n : max_n;
    for (u32 i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        F4 dir = decode_oct(pdir[i]);
        st(dirpow + i, mk(dir.x, dir.y, dir.z, phx_as_float(ppow[i])));
    }
}

hipError_t launch_decode_photons(const u32 *ppow, const u32 *pdir, const u32 *d_n, u32 max_n, phx_float4 *dirpow,
                                 hipStream_t s) {
    u32 blocks = (max_n + 255) / 256;
    blocks = blocks > 8192 ? 8192 : (blocks ? blocks : 1);
    hipLaunchKernelGGL(k_decode_photons, dim3(blocks), dim3(256), 0, s, ppow, pdir, d_n, max_n, dirpow);
    return hipGetLastError();
}

/* The gather with gather_pixel_flat's flattened loops: one wave per 8×8 tile
 * of the image as k_gather, or, with the image shape unknown (tiles_x = 0),
 * one wave per 64 consecutive pixels; blocks assigned XCD by XCD either way. */
template <bool DECODED>
__global__ __launch_bounds__(256) void k_gather_flat(GatherArgs a, u32 width, u32 height, u32 tiles_x, u32 tw_log2) {
    u32 wave = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    u32 lane = threadIdx.x & 63u;
    if (tiles_x == 0) {
        u32 tid = wave * 64u + lane;
        gather_pixel_flat<DECODED>(a, tid < a.n_pixels, tid, a.d, a.alpha, 0u);
        return;
    }
    u32 tw = 1u << tw_log2, th = 64u >> tw_log2; /* the wave's tile: tw × th pixels */
    u32 x = (wave % tiles_x) * tw + (lane & (tw - 1u)), y = (wave / tiles_x) * th + (lane >> tw_log2);
    gather_pixel_flat<DECODED>(a, x < width && y < height, y * width + x, a.d, a.alpha, 0u);
}

hipError_t launch_gather(const phx_float4 *hp_pos, const phx_float4 *hp_normal, const phx_float4 *hp_throughput,
                         const phx_float4 *hp_base_color, const phx_float4 *hp_direct, const phx_float4 *hp_emission,
                         u32 n_pixels, const phx_float4 *ppos, const u32 *ppow, const u32 *pdir, const u32 *cell_start,
                         const u32 *cell_end, const PhxPhotonHashData &d, f32 alpha, phx_float4 *flux,
                         f32 *photon_weight, phx_float4 *radiance, u32 width, hipStream_t s, f32 *rec_ratio,
                         f32 *rec_r2, const phx_float4 *pdirpow, const u32 *oct_start, u32 gather_tile_w) {
    if (n_pixels == 0)
        return hipSuccess;
    GatherArgs a{hp_pos,     hp_normal, hp_throughput, hp_base_color, hp_direct, hp_emission,
                 n_pixels,   ppos,      ppow,          pdir,          cell_start, cell_end,
                 d,          alpha,     flux,          photon_weight, radiance,  rec_ratio,
                 rec_r2,     pdirpow,   oct_start};
    bool dec = pdirpow != nullptr;
    if (width == 0) { /* image shape unknown (C ABI call): 64 consecutive pixels per wave */
        u32 waves = (n_pixels + 63) / 64;
        if (dec)
            hipLaunchKernelGGL(k_gather_flat<true>, dim3((waves + 3) / 4), dim3(256), 0, s, a, 0u, 0u, 0u, 0u);
        else
            hipLaunchKernelGGL(k_gather_flat<false>, dim3((waves + 3) / 4), dim3(256), 0, s, a, 0u, 0u, 0u, 0u);
        return hipGetLastError();
    }
    u32 tw_log2 = 3;
    for (u32 k = 0; k <= 6; k++)
        if (gather_tile_w == (1u << k))
            tw_log2 = k;
    u32 tw = 1u << tw_log2, th = 64u >> tw_log2;
    u32 h = (n_pixels + width - 1) / width;
    u32 tiles_x = (width + tw - 1) / tw, tiles_y = (h + th - 1) / th;
    u32 waves = tiles_x * tiles_y;
    if (dec)
        hipLaunchKernelGGL(k_gather_flat<true>, dim3((waves + 3) / 4), dim3(256), 0, s, a, width, h, tiles_x, tw_log2);
    else
        hipLaunchKernelGGL(k_gather_flat<false>, dim3((waves + 3) / 4), dim3(256), 0, s, a, width, h, tiles_x, tw_log2);
    return hipGetLastError();
}

bool kernel_info_map(PhxKernelInfo *out, u32 max, u32 &n) {
    PhxKernelInfo k;
    bool ok = true;
    PHX_INFO("k_hash", k_hash, 256)
    PHX_INFO("k_radix_scatter", k_radix_scatter, SORT_BLOCK)
    PHX_INFO("k_decode_photons", k_decode_photons, 256)
    PHX_INFO("k_gather_flat", k_gather_flat<true>, 256)
    return ok;
}

} // namespace phx
