// Arena assembly: graphs held on the device one by one (local ids) copied
// into one disjoint union (global ids), and labels moved between the two
// numberings.  Element-wise kernels; the table-driven ones cover every graph
// of an arena in one launch.
#ifndef SCD_ARENA_KERNELS_HPP
#define SCD_ARENA_KERNELS_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev.hpp"

namespace scd_dev {

// dst[i] = src[i] + off, i < count (CSR offsets and adjacency ids).
__global__ void k_arena_shift(const uint32_t *__restrict__ src, uint64_t count, uint32_t off,
                              uint32_t *__restrict__ dst) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = src[i] + off;
}

// Labels local -> global (to_global) or global -> local; INVALID stays.
__global__ void k_arena_labels(const uint32_t *__restrict__ src, uint32_t n, uint32_t off, int to_global,
                               uint32_t *__restrict__ dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t l = src[i];
    dst[i] = l == SCD_INVALID ? l : (to_global ? l + off : l - off);
}

// One graph of an arena assembled from device-resident graphs: its local
// arrays and where its vertices and adjacency entries start in the arena.
struct ArenaPart {
    const uint32_t *indices, *adj;
    const int32_t *tri, *tri_deg;
    const uint32_t *cur, *best;  // local labels (best may be null: cur), or both null
    uint32_t n, vo;              // vertices, first arena vertex
    uint32_t e2, eo;             // adjacency entries (2m), first arena entry
};

// One graph's labels leaving an arena: arena vertices [vo, vo + n) with
// labels offset by label_off, written local to cur / best (either may be
// null) from position so of the output space.
struct ArenaOut {
    uint32_t *cur, *best;
    uint32_t vo, n, label_off, so;
};

constexpr uint32_t ARENA_CHUNK = 4096;  // elements per workgroup (256 threads x 16)

// Largest g < G with key(g) <= x, for keys non-decreasing in g and key(0) <= x.
template <typename K>
__device__ __forceinline__ uint32_t arena_find(uint32_t G, uint32_t x, K key) {
    uint32_t lo = 0, hi = G;  // key(lo) <= x, answer in [lo, hi)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (key(mid) <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Arena vertices [0, n]: offsets shifted by the graph's first entry, the
// triangle arrays, the graph id and (labels set) both label arrays made
// global; indices[n] = E.  Each workgroup covers ARENA_CHUNK consecutive
// vertices and finds its first graph once.
__global__ __launch_bounds__(256) void k_arena_vertices(const ArenaPart *__restrict__ parts, uint32_t G, uint32_t n,
                                                        uint32_t E, int labels, uint32_t *__restrict__ indices,
                                                        int32_t *__restrict__ tri, int32_t *__restrict__ tri_deg,
                                                        uint32_t *__restrict__ gid, uint32_t *__restrict__ cur,
                                                        uint32_t *__restrict__ best, uint32_t *__restrict__ bad) {
    __shared__ uint32_t s_g;
    bool ok = true;
    const uint32_t base = blockIdx.x * ARENA_CHUNK;
    if (threadIdx.x == 0) s_g = base < n ? arena_find(G, base, [&](uint32_t g) { return parts[g].vo; }) : 0u;
    __syncthreads();
    uint32_t g = s_g;
    uint32_t lo = parts[g].vo, hi = lo + parts[g].n;
    for (uint32_t k = 0; k < ARENA_CHUNK; k += 256) {
        const uint32_t i = base + k + threadIdx.x;
        if (i > n) break;
        if (i == n) {
            indices[n] = E;
            break;
        }
        while (i >= hi) {
            g++;
            lo = parts[g].vo;
            hi = lo + parts[g].n;
        }
        const ArenaPart &P = parts[g];
        const uint32_t v = i - lo;
        const uint32_t b0 = P.indices[v], b1 = P.indices[v + 1];
        // the part's CSR offsets: from 0 to 2m, not decreasing
        ok &= b0 <= b1 && (v != 0 || b0 == 0) && (v + 1 != P.n || b1 == P.e2);
        indices[i] = b0 + P.eo;
        tri[i] = P.tri[v];
        tri_deg[i] = P.tri_deg[v];
        gid[i] = g;
        if (labels) {
            const uint32_t c = P.cur[v], b = P.best ? P.best[v] : c;
            ok &= (c == SCD_INVALID || c < P.n) && (b == SCD_INVALID || b < P.n);
            cur[i] = c == SCD_INVALID ? c : c + lo;
            best[i] = b == SCD_INVALID ? b : b + lo;
        }
    }
    if (__any(!ok) && (threadIdx.x & 63) == 0) bad[0] = 1u;
}

// Arena adjacency [0, E): ids shifted by the graph's first vertex.
__global__ __launch_bounds__(256) void k_arena_entries(const ArenaPart *__restrict__ parts, uint32_t G, uint32_t E,
                                                       uint32_t *__restrict__ adj, uint32_t *__restrict__ bad) {
    __shared__ uint32_t s_g;
    bool ok = true;
    const uint32_t base = blockIdx.x * ARENA_CHUNK;
    if (threadIdx.x == 0) s_g = arena_find(G, base, [&](uint32_t g) { return parts[g].eo; });
    __syncthreads();
    uint32_t g = s_g;
    uint32_t lo = parts[g].eo, hi = lo + parts[g].e2;
    for (uint32_t k = 0; k < ARENA_CHUNK; k += 256) {
        const uint32_t j = base + k + threadIdx.x;
        if (j >= E) break;
        while (j >= hi) {
            g++;
            lo = parts[g].eo;
            hi = lo + parts[g].e2;
        }
        const uint32_t x = parts[g].adj[j - lo];
        ok &= x < parts[g].n;
        adj[j] = x + parts[g].vo;
    }
    if (__any(!ok) && (threadIdx.x & 63) == 0) bad[0] = 1u;
}

// Labels of the graphs in outs[0..G) (output space [0, total), graph g at
// [so, so + n)) from the arena's global labels to local ones.
__global__ __launch_bounds__(256) void k_arena_export(const ArenaOut *__restrict__ outs, uint32_t G, uint32_t total,
                                                      const uint32_t *__restrict__ cur,
                                                      const uint32_t *__restrict__ best) {
    __shared__ uint32_t s_g;
    const uint32_t base = blockIdx.x * ARENA_CHUNK;
    if (threadIdx.x == 0) s_g = arena_find(G, base, [&](uint32_t g) { return outs[g].so; });
    __syncthreads();
    uint32_t g = s_g;
    uint32_t lo = outs[g].so, hi = lo + outs[g].n;
    for (uint32_t k = 0; k < ARENA_CHUNK; k += 256) {
        const uint32_t x = base + k + threadIdx.x;
        if (x >= total) break;
        while (x >= hi) {
            g++;
            lo = outs[g].so;
            hi = lo + outs[g].n;
        }
        const ArenaOut &O = outs[g];
        const uint32_t v = x - lo, i = O.vo + v;
        if (O.cur) {
            const uint32_t l = cur[i];
            O.cur[v] = l == SCD_INVALID ? l : l - O.label_off;
        }
        if (O.best) {
            const uint32_t l = best[i];
            O.best[v] = l == SCD_INVALID ? l : l - O.label_off;
        }
    }
}

// ---------------------------------------------------------------------------
// Input checks before any kernel indexes by an adjacency entry or a label.
// Each sets bad[0] = 1 when it finds a violation (bad cleared beforehand).
// ---------------------------------------------------------------------------
// Every adjacency entry below n.
__global__ __launch_bounds__(256) void k_check_adj(const uint32_t *__restrict__ adj, uint64_t E, uint32_t n,
                                                   uint32_t *__restrict__ bad) {
    bool ok = true;
    for (uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x; j < E; j += (uint64_t)gridDim.x * 256)
        ok &= adj[j] < n;
    if (__any(!ok) && (threadIdx.x & 63) == 0) bad[0] = 1u;
}

// Arena of G graphs uploaded as one union (graph g: vertices [vtx_off[g],
// vtx_off[g+1])): every list points into its own graph's vertices.  One
// wave per vertex, four per workgroup.
__global__ __launch_bounds__(256) void k_check_arena(const uint32_t *__restrict__ indices,
                                                     const uint32_t *__restrict__ adj,
                                                     const uint32_t *__restrict__ gid,
                                                     const uint64_t *__restrict__ vtx_off, uint32_t n,
                                                     uint32_t *__restrict__ bad) {
    const uint32_t v = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (v >= n) return;  // wave-uniform
    const uint32_t g = gid[v];
    const uint32_t lo = (uint32_t)vtx_off[g], hi = (uint32_t)vtx_off[g + 1];
    bool ok = true;
    for (uint32_t j = indices[v] + (threadIdx.x & 63); j < indices[v + 1]; j += 64) ok &= adj[j] >= lo && adj[j] < hi;
    if (__any(!ok) && (threadIdx.x & 63) == 0) bad[0] = 1u;
}

}  // namespace scd_dev

#endif
