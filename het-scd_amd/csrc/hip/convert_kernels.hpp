// The convert pipeline on the device (src/convert.cpp:466-537): the
// reference's steps with sorts and scans instead of its sequential passes.
//   1. ids by first appearance: the first position 2i (+1 for the larger
//      endpoint) at which a vertex occurs is an atomicMin; the new id is the
//      number of vertices whose first position is smaller (a scan over
//      positions);
//   2. undirected edge keys (lo << 32 | hi) sorted, loops and duplicates
//      dropped;
//   3. the simple graph renumbered by (degree desc, id asc) so that the
//      triangle kernels' orientation toward lower ids points at hubs, both
//      directions sorted by (source, target): the symmetric CSR;
//   4. triangle counts and triangle edges (the internal-triangle kernels
//      over the whole adjacency, in the context);
//   5. relabel by (triangle edges desc, id desc), triangle-free edges and
//      isolated vertices dropped, both directions sorted again.
// 64-bit keys are sorted by a stable LSD radix sort with 8-bit digits whose
// passes skip digits that are equal in every key.
#ifndef SCD_CONVERT_KERNELS_HPP
#define SCD_CONVERT_KERNELS_HPP

#include "dev.hpp"

namespace scd_dev {

constexpr uint32_t CV_TILE = 4096;  // keys per radix-sort workgroup (256 x 16)

constexpr uint32_t CV_REDUCE_BLOCKS = 2048;  // grid of the whole-array reductions

// Largest endpoint: a grid-stride pass, one atomic per workgroup.
__global__ __launch_bounds__(256) void k_cv_max(const uint32_t *__restrict__ u, const uint32_t *__restrict__ v,
                                                uint64_t count, uint32_t *__restrict__ mx) {
    __shared__ uint32_t sm[4];
    uint32_t m = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256)
        m = max(m, max(u[i], v[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor(m, o, SCD_WAVE));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = max(max(sm[0], sm[1]), max(sm[2], sm[3]));
        if (m) atomicMax(mx, m);
    }
}

// first[x] = min over occurrences of 2i (smaller endpoint of edge i) or
// 2i + 1 (larger endpoint); src/convert.cpp:141-170 numbers a before b.
__global__ void k_cv_first(const uint32_t *__restrict__ u, const uint32_t *__restrict__ v, uint64_t count,
                           unsigned long long *__restrict__ first) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t a = min(u[i], v[i]), b = max(u[i], v[i]);
    // a plain read first: hubs occur in many edges, and once an earlier
    // position is in place the atomic is skipped (a stale read only costs
    // an atomic that changes nothing)
    const unsigned long long pa = 2 * i, pb = 2 * i + 1;
    if (first[a] > pa) atomicMin(&first[a], pa);
    if (first[b] > pb) atomicMin(&first[b], pb);
}

// mark[first[x]] = 1 for every vertex that occurs.
__global__ void k_cv_mark(const unsigned long long *__restrict__ first, uint64_t nid, uint32_t *__restrict__ mark) {
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nid) return;
    const unsigned long long f = first[x];
    if (f != ~0ull) mark[f] = 1u;
}

// Undirected key (min << 32 | max) of edge i in the new ids
// (seq = exclusive scan of mark).
__global__ void k_cv_keys(const uint32_t *__restrict__ u, const uint32_t *__restrict__ v, uint64_t count,
                          const unsigned long long *__restrict__ first, const uint32_t *__restrict__ seq,
                          uint64_t *__restrict__ keys) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t x = seq[first[u[i]]], y = seq[first[v[i]]];
    keys[i] = ((uint64_t)min(x, y) << 32) | max(x, y);
}

// AND / OR of all keys (digits equal everywhere need no sort pass): a
// grid-stride pass, one atomic pair per workgroup.
__global__ __launch_bounds__(256) void k_cv_bits(const uint64_t *__restrict__ keys, uint64_t n,
                                                 unsigned long long *__restrict__ bits) {
    __shared__ unsigned long long sa[4], so[4];
    unsigned long long a = ~0ull, o = 0ull;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const unsigned long long k = keys[i];
        a &= k;
        o |= k;
    }
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        a &= __shfl_xor(a, s, SCD_WAVE);
        o |= __shfl_xor(o, s, SCD_WAVE);
    }
    if ((threadIdx.x & 63) == 0) {
        sa[threadIdx.x >> 6] = a;
        so[threadIdx.x >> 6] = o;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAnd(&bits[0], sa[0] & sa[1] & sa[2] & sa[3]);
        atomicOr(&bits[1], so[0] | so[1] | so[2] | so[3]);
    }
}

__global__ void k_cv_bits_init(unsigned long long *bits) {
    if (threadIdx.x == 0) {
        bits[0] = ~0ull;
        bits[1] = 0ull;
    }
}

// Per-tile digit histogram, digit-major: hist[d * ntiles + tile].
__global__ __launch_bounds__(256) void k_cv_hist(const uint64_t *__restrict__ keys, uint64_t n, int shift,
                                                 uint32_t ntiles, uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * CV_TILE;
    for (uint32_t k = 0; k < CV_TILE / 256; k++) {
        const uint64_t i = base + k * 256 + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of a tile: within a digit, elements keep their index order
// (wave ballots rank the lanes, per-wave counts order the waves).
__global__ __launch_bounds__(256) void k_cv_scatter(const uint64_t *__restrict__ keys, uint64_t n, int shift,
                                                    uint32_t ntiles, const uint32_t *__restrict__ off,
                                                    uint64_t *__restrict__ out) {
    __shared__ uint32_t s_base[256];
    __shared__ uint32_t s_wc[4][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    s_base[threadIdx.x] = off[(uint64_t)threadIdx.x * ntiles + blockIdx.x];
    const uint64_t base = (uint64_t)blockIdx.x * CV_TILE;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint32_t k = 0; k < CV_TILE / 256; k++) {
        for (int q = 0; q < 4; q++) s_wc[q][threadIdx.x] = 0;
        __syncthreads();
        const uint64_t i = base + k * 256 + threadIdx.x;
        const bool act = i < n;
        uint64_t key = 0;
        uint32_t d = 0;
        if (act) {
            key = keys[i];
            d = (uint32_t)(key >> shift) & 0xFFu;
        }
        uint64_t same = __ballot(act);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t m = __ballot((d >> b) & 1u);
            same &= ((d >> b) & 1u) ? m : ~m;
        }
        const uint32_t rank_in_wave = (uint32_t)__builtin_popcountll(same & lt);
        if (act && rank_in_wave == 0) s_wc[w][d] = (uint32_t)__builtin_popcountll(same);
        __syncthreads();
        if (act) {
            uint32_t before = 0;
            for (int q = 0; q < w; q++) before += s_wc[q][d];
            out[s_base[d] + before + rank_in_wave] = key;
        }
        __syncthreads();
        s_base[threadIdx.x] += s_wc[0][threadIdx.x] + s_wc[1][threadIdx.x] + s_wc[2][threadIdx.x] +
                               s_wc[3][threadIdx.x];
        __syncthreads();
    }
}

// keep[i] = key i starts a new value and is not a loop.
__global__ void k_cv_unique(const uint64_t *__restrict__ keys, uint64_t n, uint32_t *__restrict__ keep) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    keep[i] = ((uint32_t)(k >> 32) != (uint32_t)k && (i == 0 || keys[i - 1] != k)) ? 1u : 0u;
}

// Degrees of the simple graph (first-appearance ids).  Keys are sorted by
// their smaller end, so each wave adds the run of one smaller end with a
// single atomic (hubs, numbered early, are mostly smaller ends); larger ends
// are added one by one.
__global__ __launch_bounds__(256) void k_cv_deg(const uint64_t *__restrict__ keys, uint64_t n,
                                                const uint32_t *__restrict__ keep, uint32_t *__restrict__ deg) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool act = i < n;
    uint64_t k = 0;
    uint32_t kp = 0;
    if (act) {
        k = keys[i];
        kp = keep[i];
    }
    const uint32_t lo = (uint32_t)(k >> 32);
    const uint32_t prev = __shfl_up(lo, 1, SCD_WAVE);
    const bool head = act && (lane == 0 || prev != lo);
    const uint64_t heads = __ballot(head), kept = __ballot(act && kp);
    if (head) {
        const uint64_t above = heads & ~((2ull << lane) - 1ull);  // later heads (lane 63: none)
        const int end = above ? __builtin_ctzll(above) : 64;
        const uint64_t run = (end == 64 ? ~0ull : ((1ull << end) - 1ull)) & ~((1ull << lane) - 1ull);
        const uint32_t c = (uint32_t)__builtin_popcountll(kept & run);
        if (c) atomicAdd(&deg[lo], c);
    }
    if (act && kp) atomicAdd(&deg[(uint32_t)k], 1u);
}

// Sort keys of the degree order: degree desc, id asc.
__global__ void k_cv_rank_keys(const uint32_t *__restrict__ deg, uint32_t n, uint64_t *__restrict__ keys) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n) keys[x] = ((uint64_t)(~deg[x]) << 32) | x;
}

__global__ void k_cv_rank(const uint64_t *__restrict__ order, uint32_t n, uint32_t *__restrict__ rank,
                          uint32_t *__restrict__ perm) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint32_t x = (uint32_t)order[r];
    rank[x] = r;
    perm[r] = x;
}

// Both directions of every kept undirected key, in degree-order ids.
__global__ void k_cv_directed(const uint64_t *__restrict__ keys, uint64_t n, const uint32_t *__restrict__ keep,
                              const uint32_t *__restrict__ pos, const uint32_t *__restrict__ rank,
                              uint64_t *__restrict__ dk) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    const uint64_t k = keys[i];
    const uint64_t x = rank[(uint32_t)(k >> 32)], y = rank[(uint32_t)k];
    const uint64_t j = pos[i];
    dk[2 * j] = (x << 32) | y;
    dk[2 * j + 1] = (y << 32) | x;
}

// CSR of sorted directed keys: adj[i] = target, deg[source] counted.
__global__ void k_cv_csr(const uint64_t *__restrict__ dk, uint64_t E, uint32_t *__restrict__ adj,
                         uint32_t *__restrict__ deg) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E) return;
    const uint64_t k = dk[i];
    adj[i] = (uint32_t)k;
    if (i == 0 || (dk[i - 1] >> 32) != (k >> 32)) {  // run head: its length is the degree
        const uint32_t s = (uint32_t)(k >> 32);
        uint64_t lo = i + 1, hi = E;  // first index whose source differs
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if ((uint32_t)(dk[mid] >> 32) == s) lo = mid + 1;
            else hi = mid;
        }
        deg[s] = (uint32_t)(lo - i);
    }
}

// Order keys of the relabel, per degree-order vertex r: triangle edges
// desc, first-appearance id desc -> ascending keys.
__global__ void k_cv_order(const uint32_t *__restrict__ td, const uint32_t *__restrict__ perm, uint32_t n,
                           uint64_t *__restrict__ keys) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    keys[r] = ((uint64_t)(~td[r]) << 32) | (uint32_t)~perm[r];
}

__global__ void k_cv_newid(const uint64_t *__restrict__ order, uint32_t n, const uint32_t *__restrict__ rank,
                           uint32_t *__restrict__ newid) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    newid[rank[~(uint32_t)order[p]]] = p;
}

// Number of vertices with at least one triangle edge.
__global__ void k_cv_count_pos(const uint32_t *__restrict__ td, uint32_t n, uint32_t *__restrict__ cnt) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t b = __ballot(v < n && td[v] > 0);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(cnt, (uint32_t)__builtin_popcountll(b));
}

// Kept (triangle) edges: the triangle kernels flag the slot u -> w (w < u)
// of every edge in a triangle, so flag[s] marks each kept undirected edge
// exactly once.
__global__ void k_cv_flag32(const uint8_t *__restrict__ flag, uint64_t E, uint32_t *__restrict__ out) {
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < E) out[s] = flag[s];
}

// Both directions of every kept edge in the final ids; dk holds the sorted
// directed keys (source << 32 | target) of the slots.
__global__ void k_cv_final_keys(const uint64_t *__restrict__ dk, const uint8_t *__restrict__ flag, uint64_t E,
                                const uint32_t *__restrict__ newid, const uint32_t *__restrict__ pos,
                                uint64_t *__restrict__ out) {
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= E || !flag[s]) return;
    const uint64_t k = dk[s];
    const uint32_t nu = newid[(uint32_t)(k >> 32)], nv = newid[(uint32_t)k];
    const uint64_t j = pos[s];
    out[2 * j] = ((uint64_t)nu << 32) | nv;
    out[2 * j + 1] = ((uint64_t)nv << 32) | nu;
}

// tri / tri_deg in the final ids (tri = 2 x triangles at the vertex).
__global__ void k_cv_final_tri(const uint32_t *__restrict__ T, const uint32_t *__restrict__ td,
                               const uint32_t *__restrict__ newid, uint32_t n, uint32_t n_new,
                               int32_t *__restrict__ tri, int32_t *__restrict__ tri_deg) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const uint32_t r = newid[v];
    if (r >= n_new) return;
    tri[r] = (int32_t)(2u * T[v]);
    tri_deg[r] = (int32_t)td[v];
}

}  // namespace scd_dev

#endif
