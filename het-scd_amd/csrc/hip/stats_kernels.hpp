// Community statistics, internal adjacency and WCC kernels.
//
// Statistics (collect_stats, src/scd.cpp:57-111): for every vertex p with a
// valid label L, sizes[L] += 1, ext[L] += #neighbours with another label
// (INVALID neighbours included), int[L] += #neighbours with label L.  Integer
// atomics, so the result is independent of order.  The same pass writes p's
// same-label neighbours, in adjacency order, to int_adj at p's own CSR offset
// (int_adj[indices[p] .. + int_deg[p]]): the internal subgraph the WCC needs,
// built without a global scan.
//
// The WCC kernels (wcc_kernels.hpp) count triangles on this internal
// subgraph; the mean of the per-vertex terms is a fixed-shape tree sum over
// vertex order (k_reduce_*), so its bits do not depend on the degree bins or
// tile sizes.
#ifndef SCD_STATS_KERNELS_HPP
#define SCD_STATS_KERNELS_HPP

#include "dev.hpp"

namespace scd_dev {

struct StatArgs {
    const uint32_t *__restrict__ indices;
    const uint32_t *__restrict__ adj;
    const uint32_t *__restrict__ labels;  // current (new) labels
    int4 *__restrict__ stats;
    uint32_t *__restrict__ int_adj;
    uint32_t *__restrict__ int_deg;
    uint8_t *__restrict__ iflag;  // per internal entry: cleared here, set by the WCC triangle count
    unsigned long long *__restrict__ counters;  // [0] += internal entries, [1] += deg_int^2
    // batched arenas: vertices of stopped graphs get no internal list
    const uint32_t *__restrict__ gid;
    const uint32_t *__restrict__ active_g;
};

__device__ __forceinline__ bool stat_stopped(const StatArgs &a, uint32_t p) {
    return a.gid && !a.active_g[a.gid[p]];
}

// Adds (members, external, internal) counts to community L.  Size and
// external count share one 64-bit atomic (x in the low word, y in the high
// word of the 16-byte entry): a size never reaches 2^32, so no carry crosses
// into y, and the 32-bit halves wrap as the int32 sums of the reference do.
__device__ __forceinline__ void add_stats(int4 *stats, uint32_t L, uint32_t sz, uint32_t ext, uint32_t cnt) {
    if (sz > 0 || ext > 0)
        atomicAdd(reinterpret_cast<unsigned long long *>(&stats[L].x),
                  ((unsigned long long)ext << 32) | (unsigned long long)sz);
    if (cnt > 0) atomicAdd(&stats[L].z, (int)cnt);
}

// Community statistics from the per-vertex internal degrees: sizes[L],
// ext[L] = sum of (deg - int_deg), int[L] = sum of int_deg over the members
// of L.  Large communities would serialise on one address, so each block
// (1024 threads, grid-stride over the vertices) first sums its vertices per
// label in an LDS hash table and flushes one atomic per label and field;
// a vertex whose label finds no free slot within STATS_PROBES probes adds
// its counts directly.  Integer sums, so the result is order-independent.
constexpr uint32_t STATS_SLOTS = 2048;
constexpr uint32_t STATS_PROBES = 8;

__global__ __launch_bounds__(1024) void k_stats_agg(const uint32_t *__restrict__ indices,
                                                    const uint32_t *__restrict__ labels,
                                                    const uint32_t *__restrict__ int_deg, uint32_t n,
                                                    const uint32_t *__restrict__ gid,
                                                    const uint32_t *__restrict__ active_g,
                                                    int4 *__restrict__ stats) {
    __shared__ uint32_t s_key[STATS_SLOTS];
    __shared__ uint32_t s_sz[STATS_SLOTS], s_ext[STATS_SLOTS], s_int[STATS_SLOTS];
    for (uint32_t h = threadIdx.x; h < STATS_SLOTS; h += 1024) {
        s_key[h] = SCD_INVALID;
        s_sz[h] = s_ext[h] = s_int[h] = 0;
    }
    __syncthreads();
    for (uint32_t p = blockIdx.x * 1024u + threadIdx.x; p < n; p += gridDim.x * 1024u) {
        const uint32_t L = labels[p];
        if (L == SCD_INVALID || (gid && !active_g[gid[p]])) continue;
        const uint32_t ic = int_deg[p];
        const uint32_t ext = indices[p + 1] - indices[p] - ic;
        uint32_t h = (L * 0x9E3779B1u) >> (32 - 11);
        bool done = false;
        for (uint32_t k = 0; k < STATS_PROBES && !done; k++, h = (h + 1) & (STATS_SLOTS - 1)) {
            uint32_t key = s_key[h];
            if (key == SCD_INVALID) key = atomicCAS(&s_key[h], SCD_INVALID, L);
            if (key == SCD_INVALID || key == L) {
                atomicAdd(&s_sz[h], 1u);
                atomicAdd(&s_ext[h], ext);
                atomicAdd(&s_int[h], ic);
                done = true;
            }
        }
        if (!done) add_stats(stats, L, 1, ext, ic);
    }
    __syncthreads();
    for (uint32_t h = threadIdx.x; h < STATS_SLOTS; h += 1024) {
        const uint32_t L = s_key[h];
        if (L != SCD_INVALID) add_stats(stats, L, s_sz[h], s_ext[h], s_int[h]);
    }
}

__device__ __forceinline__ void count_internal(unsigned long long *counters, uint64_t ent,
                                               uint64_t sq) {
    if (!counters) return;  // counting passes only
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        ent += __shfl_xor(ent, off, SCD_WAVE);
        sq += __shfl_xor(sq, off, SCD_WAVE);
    }
    if ((threadIdx.x & 63) == 0 && counters) {
        atomicAdd(counters + 0, (unsigned long long)ent);
        atomicAdd(counters + 1, (unsigned long long)sq);
    }
}

__global__ void k_fill_int4(int4 *__restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_int4(0, 0, 0, 0);
}

__global__ void k_fill_u8(uint8_t *__restrict__ p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_fill_u64v(unsigned long long *__restrict__ p, uint64_t n, unsigned long long v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long *__restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// S lanes per vertex (deg <= S <= 64).
template <int S>
__global__ __launch_bounds__(256) void k_filter_sub(StatArgs a, const uint32_t *__restrict__ verts,
                                                    uint32_t nv) {
    const int lane = threadIdx.x & 63;
    const int lig = threadIdx.x & (S - 1);
    const int gbase = lane & ~(S - 1);
    const uint32_t slot = (blockIdx.x * 256u + threadIdx.x) / S;
    const bool active = slot < nv;
    const uint32_t p = active ? verts[slot] : 0u;
    const uint32_t beg = active ? a.indices[p] : 0u;
    const uint32_t deg = active ? a.indices[p + 1] - beg : 0u;
    const uint32_t L = (active && !stat_stopped(a, p)) ? a.labels[p] : SCD_INVALID;
    bool same = false;
    uint32_t q = 0;
    if ((uint32_t)lig < deg && L != SCD_INVALID) {
        q = a.adj[beg + lig];
        same = (L != SCD_INVALID) && a.labels[q] == L;
    }
    uint64_t gm;
    if constexpr (S == 64) gm = __ballot(same);
    else gm = (__ballot(same) >> gbase) & ((1ull << S) - 1ull);
    const uint32_t cnt = (uint32_t)__builtin_popcountll(gm);
    if (same) {
        const uint32_t at = beg + __builtin_popcountll(gm & ((1ull << lig) - 1ull));
        a.int_adj[at] = q;
        a.iflag[at] = 0;
    }
    const bool lead = active && lig == 0;
    if (lead) {
        a.int_deg[p] = cnt;
    }
    count_internal(a.counters, lead ? cnt : 0, lead ? (uint64_t)cnt * cnt : 0);
}

// One wave per vertex, deg <= 64 R: every adjacency and label load of the
// vertex is issued before the compaction.
template <int R>
__global__ __launch_bounds__(256) void k_filter_wave(StatArgs a, const uint32_t *__restrict__ verts,
                                                     uint32_t nv) {
    const int lane = threadIdx.x & 63;
    const uint32_t slot = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (slot >= nv) return;
    const uint32_t p = verts[slot];
    const uint32_t beg = a.indices[p];
    const uint32_t L = a.labels[p];
    if (L == SCD_INVALID || stat_stopped(a, p)) {  // no internal list; uniform per wave
        if (lane == 0) a.int_deg[p] = 0;
        return;
    }
    const uint32_t deg = a.indices[p + 1] - beg;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t q[R], lq[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t i = r * 64 + lane;
        q[r] = i < deg ? a.adj[beg + i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t i = r * 64 + lane;
        lq[r] = i < deg ? a.labels[q[r]] : SCD_INVALID;
    }
    uint32_t run = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const bool same = (L != SCD_INVALID) && lq[r] == L && (uint32_t)(r * 64 + lane) < deg;
        const uint64_t m = __ballot(same);
        if (same) {
            const uint32_t at = beg + run + __builtin_popcountll(m & lt);
            a.int_adj[at] = q[r];
            a.iflag[at] = 0;
        }
        run += (uint32_t)__builtin_popcountll(m);
    }
    if (lane == 0) {
        a.int_deg[p] = run;
    }
    count_internal(a.counters, lane == 0 ? run : 0, lane == 0 ? (uint64_t)run * run : 0);
}

// One workgroup per vertex (deg <= T): all loads first, then one
// block-wide order-preserving compaction over the T / 256 slices.
template <int T>
__global__ __launch_bounds__(256) void k_filter_block(StatArgs a, const uint32_t *__restrict__ verts,
                                                      uint32_t nv) {
    constexpr int U = T / 256;
    __shared__ uint32_t s_cnt[U][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t p = verts[blockIdx.x];
    const uint32_t beg = a.indices[p];
    const uint32_t L = a.labels[p];
    if (L == SCD_INVALID || stat_stopped(a, p)) {  // uniform per block, before any barrier
        if (threadIdx.x == 0) a.int_deg[p] = 0;
        return;
    }
    const uint32_t deg = a.indices[p + 1] - beg;
    uint32_t q[U];
    bool same[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const uint32_t i = u * 256 + threadIdx.x;
        q[u] = i < deg ? a.adj[beg + i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const uint32_t i = u * 256 + threadIdx.x;
        same[u] = i < deg && L != SCD_INVALID && a.labels[q[u]] == L;
    }
    uint64_t m[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        m[u] = __ballot(same[u]);
        if (lane == 0) s_cnt[u][w] = (uint32_t)__builtin_popcountll(m[u]);
    }
    __syncthreads();
    uint32_t run = 0;
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int u = 0; u < U; u++) {
        uint32_t before = run;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            before += (k < w) ? s_cnt[u][k] : 0u;
            run += s_cnt[u][k];
        }
        if (same[u]) {
            const uint32_t at = beg + before + __builtin_popcountll(m[u] & lt);
            a.int_adj[at] = q[u];
            a.iflag[at] = 0;
        }
    }
    if (threadIdx.x == 0) {
        a.int_deg[p] = run;
    }
    count_internal(a.counters, threadIdx.x == 0 ? run : 0,
                   threadIdx.x == 0 ? (uint64_t)run * run : 0);
}

// Order-preserving compaction of one slice of up to 256 entries by a block
// of 256 threads: returns the slice's count; `same` entries are written at
// out[run + rank] (and oflag[run + rank] cleared, when given).
__device__ __forceinline__ uint32_t block_compact256(bool same, uint32_t q, uint32_t *out, uint8_t *oflag,
                                                     uint32_t run, uint32_t *s_cnt) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t m = __ballot(same);
    if (lane == 0) s_cnt[w] = (uint32_t)__builtin_popcountll(m);
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        before += (k < w) ? s_cnt[k] : 0u;
        total += s_cnt[k];
    }
    if (same) {
        const uint32_t at = run + before + __builtin_popcountll(m & ((1ull << lane) - 1ull));
        out[at] = q;
        if (oflag) oflag[at] = 0;
    }
    __syncthreads();
    return total;
}

// Hubs, pass 1: each chunk of C entries gathers its neighbours' labels
// once and compacts its same-label entries, in adjacency order, into the
// hub's scratch segment at the chunk's own offset (tmp[hub_off + o ..], the
// move step's hub buffer, free during the statistics); the count goes to
// chunk_cnt.
__global__ __launch_bounds__(256) void k_hub_filter_compact(StatArgs a, const uint32_t *__restrict__ hub_v,
                                                            const uint32_t *__restrict__ hub_d,
                                                            const uint32_t *__restrict__ hub_off,
                                                            const uint32_t *__restrict__ chunk_hub,
                                                            const uint32_t *__restrict__ chunk_start, uint32_t C,
                                                            uint32_t *__restrict__ chunk_cnt,
                                                            uint32_t *__restrict__ tmp) {
    __shared__ uint32_t s_cnt[4];
    const uint32_t hb = chunk_hub[blockIdx.x];
    const uint32_t o = chunk_start[blockIdx.x];
    const uint32_t p = hub_v[hb];
    const uint32_t d = hub_d[hb];
    const uint32_t beg = a.indices[p];
    const uint32_t L = a.labels[p];
    const uint32_t end = min(o + C, d);
    uint32_t *out = tmp + hub_off[hb] + o;
    uint32_t run = 0;
    for (uint32_t b = o; b < end; b += 256) {
        const uint32_t i = b + threadIdx.x;
        bool same = false;
        uint32_t q = 0;
        if (i < end && L != SCD_INVALID) {
            q = a.adj[beg + i];
            same = a.labels[q] == L;
        }
        run += block_compact256(same, q, out, nullptr, run, s_cnt);
    }
    if (threadIdx.x == 0) chunk_cnt[blockIdx.x] = run;
}

// Hubs, pass 2: each chunk's entries go after those of the hub's earlier
// chunks in int_adj (their triangle flags cleared); the hub's first chunk
// records int_deg.
__global__ __launch_bounds__(256) void k_hub_filter_place(StatArgs a, const uint32_t *__restrict__ hub_v,
                                                          const uint32_t *__restrict__ hub_off,
                                                          const uint32_t *__restrict__ hub_chunk_begin,
                                                          const uint32_t *__restrict__ chunk_hub,
                                                          const uint32_t *__restrict__ chunk_start,
                                                          const uint32_t *__restrict__ chunk_cnt,
                                                          const uint32_t *__restrict__ tmp) {
    const uint32_t c = blockIdx.x;
    const uint32_t hb = chunk_hub[c];
    const uint32_t o = chunk_start[c];
    const uint32_t p = hub_v[hb];
    const uint32_t beg = a.indices[p];
    uint32_t run = 0;
    for (uint32_t k = hub_chunk_begin[hb]; k < c; k++) run += chunk_cnt[k];
    if (o == 0 && threadIdx.x == 0) {
        uint32_t tot = 0;
        for (uint32_t k = hub_chunk_begin[hb]; k < hub_chunk_begin[hb + 1]; k++) tot += chunk_cnt[k];
        a.int_deg[p] = tot;
        if (a.counters) {
            atomicAdd(a.counters + 0, (unsigned long long)tot);
            atomicAdd(a.counters + 1, (unsigned long long)tot * tot);
        }
    }
    const uint32_t cnt = chunk_cnt[c];
    const uint32_t *src = tmp + hub_off[hb] + o;
    for (uint32_t k = threadIdx.x; k < cnt; k += 256) {
        a.int_adj[beg + run + k] = src[k];
        a.iflag[beg + run + k] = 0;
    }
}

// Fixed-shape sum of x[0..n): block b sums x[b*4096 .. +4096) as 256 strided
// partial sums (thread t adds x[base + t + 256k], k ascending) followed by a
// pairwise tree; the shape depends on n only.
__device__ __forceinline__ double block_tree_sum256(double v, double *s) {
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + w];
        __syncthreads();
    }
    return s[0];
}

__global__ __launch_bounds__(256) void k_reduce_partial(const double *__restrict__ x, uint32_t n,
                                                        double *__restrict__ partial) {
    __shared__ double s[256];
    const uint64_t base = (uint64_t)blockIdx.x * 4096u;
    double acc = 0.0;
    for (uint32_t k = 0; k < 16; k++) {
        const uint64_t i = base + threadIdx.x + 256u * k;
        if (i < n) acc = acc + x[i];
    }
    const double r = block_tree_sum256(acc, s);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// Single block: the same shape over the partial sums, then / count.
__global__ __launch_bounds__(256) void k_reduce_final(const double *__restrict__ partial, uint32_t np,
                                                      uint32_t count, double *__restrict__ out) {
    __shared__ double s[256];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < np; i += 256) acc = acc + partial[i];
    const double r = block_tree_sum256(acc, s);
    if (threadIdx.x == 0) *out = count > 0 ? r / (double)count : 0.0;
}

// Per-graph means of an arena with the single-graph tree shape: item t is
// block b of graph g (starting at vertex gs = g's first vertex + 4096 b);
// partial[t] sums x over that block exactly like k_reduce_partial, and
// k_reduce_final_g reduces graph g's partials like k_reduce_final.
__global__ __launch_bounds__(256) void k_reduce_partial_g(const double *__restrict__ x,
                                                          const uint32_t *__restrict__ item_start,
                                                          const uint32_t *__restrict__ item_end,
                                                          double *__restrict__ partial) {
    __shared__ double s[256];
    const uint32_t base = item_start[blockIdx.x], end = item_end[blockIdx.x];
    double acc = 0.0;
    for (uint32_t k = 0; k < 16; k++) {
        const uint32_t i = base + threadIdx.x + 256u * k;
        if (i < end) acc = acc + x[i];
    }
    const double r = block_tree_sum256(acc, s);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ __launch_bounds__(256) void k_reduce_final_g(const double *__restrict__ partial,
                                                        const uint32_t *__restrict__ g_item_begin,
                                                        const uint64_t *__restrict__ vtx_off,
                                                        double *__restrict__ out) {
    __shared__ double s[256];
    const uint32_t g = blockIdx.x;
    const uint32_t b = g_item_begin[g], np = g_item_begin[g + 1] - b;
    const uint32_t count = (uint32_t)(vtx_off[g + 1] - vtx_off[g]);
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < np; i += 256) acc = acc + partial[b + i];
    const double r = block_tree_sum256(acc, s);
    if (threadIdx.x == 0) out[g] = count > 0 ? r / (double)count : 0.0;
}

// best[p] = cur[p] for the vertices of graphs flagged in store_g.
__global__ void k_store_best_g(const uint32_t *__restrict__ cur, const uint32_t *__restrict__ gid,
                               const uint32_t *__restrict__ store_g, uint32_t n, uint32_t *__restrict__ best) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n && store_g[gid[p]]) best[p] = cur[p];
}

// ---------------------------------------------------------------------------
// Exclusive scan of uint32 (wrapping) in three steps: per-tile sums, scan of
// the tile sums (recursively, host-driven), tile-local scan plus tile offset.
// Tiles are 2048 entries: 256 threads x 8 consecutive entries.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t *s_w, uint32_t &total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(inc, off, SCD_WAVE);
        if (lane >= off) inc += t;
    }
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        before += (k < w) ? s_w[k] : 0u;
        total += s_w[k];
    }
    __syncthreads();
    return before + inc - v;
}

__global__ __launch_bounds__(256) void k_scan_tile_sums(const uint32_t *__restrict__ in, uint64_t n,
                                                        uint32_t *__restrict__ sums) {
    __shared__ uint32_t s_w[4];
    const uint64_t base = (uint64_t)blockIdx.x * 2048u;
    uint32_t acc = 0;
    for (int k = 0; k < 8; k++) {
        const uint64_t i = base + threadIdx.x + 256u * k;
        if (i < n) acc += in[i];
    }
    uint32_t total;
    block_excl_scan256(acc, s_w, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_scan_tiles(const uint32_t *__restrict__ in, uint64_t n,
                                                    const uint32_t *__restrict__ tile_off,
                                                    uint32_t *__restrict__ out) {
    __shared__ uint32_t s_v[2048];
    __shared__ uint32_t s_w[4];
    const uint64_t base = (uint64_t)blockIdx.x * 2048u;
    for (int k = 0; k < 8; k++) {
        const uint64_t i = base + threadIdx.x + 256u * k;
        s_v[threadIdx.x + 256 * k] = i < n ? in[i] : 0u;
    }
    __syncthreads();
    uint32_t loc[8], acc = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        loc[k] = s_v[threadIdx.x * 8 + k];
        acc += loc[k];
    }
    uint32_t total;
    uint32_t run = block_excl_scan256(acc, s_w, total) + (tile_off ? tile_off[blockIdx.x] : 0u);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        s_v[threadIdx.x * 8 + k] = run;
        run += loc[k];
    }
    __syncthreads();
    for (int k = 0; k < 8; k++) {
        const uint64_t i = base + threadIdx.x + 256u * k;
        if (i < n) out[i] = s_v[threadIdx.x + 256 * k];
    }
}

// total[0] = exclusive-scan offset of the last tile + its sum.
// Exclusive scan of at most 2048 tile sums by one workgroup, and their total:
// the top level of scan_u32_dev in one launch.
__global__ __launch_bounds__(256) void k_scan_top(const uint32_t *__restrict__ sums, uint32_t nt,
                                                  uint32_t *__restrict__ scan, uint32_t *__restrict__ total) {
    __shared__ uint32_t s_w[4];
    uint32_t loc[8], acc = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const uint32_t i = threadIdx.x * 8u + k;
        loc[k] = i < nt ? sums[i] : 0u;
        acc += loc[k];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan256(acc, s_w, tot);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const uint32_t i = threadIdx.x * 8u + k;
        if (i < nt) scan[i] = run;
        run += loc[k];
    }
    if (threadIdx.x == 0 && total) total[0] = tot;
}

__global__ void k_scan_total(const uint32_t *__restrict__ tile_off, const uint32_t *__restrict__ tile_sums,
                             uint32_t ntiles, uint32_t *__restrict__ total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) total[0] = tile_off[ntiles - 1] + tile_sums[ntiles - 1];
}

// Export of the reduce-by-key runs: flag[i] = (count[i] > 0).
__global__ void k_run_flags(const uint32_t *__restrict__ cnt, uint64_t n, uint32_t *__restrict__ flag) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = cnt[i] > 0 ? 1u : 0u;
}

__global__ void k_run_scatter(const uint32_t *__restrict__ sorted, const uint32_t *__restrict__ cnt,
                              const uint32_t *__restrict__ pos, uint64_t n,
                              uint32_t *__restrict__ run_label, uint32_t *__restrict__ run_count) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && cnt[i] > 0) {
        run_label[pos[i]] = sorted[i];
        run_count[pos[i]] = cnt[i];
    }
}

__global__ void k_run_offsets(const uint32_t *__restrict__ indices, const uint32_t *__restrict__ pos,
                              uint32_t n, uint64_t E, uint32_t total, uint32_t *__restrict__ off) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v > n) return;
    const uint32_t e = indices[v];
    off[v] = (v == n || e >= E) ? total : pos[e];
}

}  // namespace scd_dev

#endif
