// Best-move kernels: one synchronous SCD move step (src/scd.cpp:221-341).
//
// For every vertex p the kernels gather the labels of its neighbours, sort
// them (own segmented sort: register bitonic across a wave64 or a lane
// group, register chunks + merge-path passes in LDS for a workgroup, tiles +
// merge passes in global memory for hubs), reduce
// equal labels to (label, count) runs with ballots (reduce-by-key), evaluate
// the improvement of moving p into each neighbouring community with the CPU
// path's fp64 arithmetic, take the argmax and apply the move rule.  Every
// read uses the old labels and statistics; the new label goes to a separate
// array, which is the reference's synchronous update (src/scd.cpp:229, :338).
//
// Vertices are split into degree bins on the host (engine.hip); each kernel
// below serves one bin shape:
//   k_move_sub<S>   deg <= S (8, 16, 32): S lanes per vertex, one key a lane
//   k_move_wave<R>  deg <= 64R (R = 1..16): one wave per vertex, R keys a lane
//   k_move_lds<T>   deg <= T: one workgroup per vertex, sort in LDS
//   hub path        deg > T: tiles of T sorted in LDS by separate workgroups,
//                   merged in global memory, then evaluated in chunks.
// In export mode (MoveArgs::ex_sorted != nullptr) each kernel also writes
// the sorted keys and, at every run head, the run length: the segmented
// sort + reduce-by-key output that replaces the reference GPU's
// SegSortKeysFromIndices and ReduceByKey (src/cuda.cu:205-227).
#ifndef SCD_MOVE_KERNELS_HPP
#define SCD_MOVE_KERNELS_HPP

#include "dev.hpp"

namespace scd_dev {

struct MoveArgs {
    const uint32_t *__restrict__ indices;
    const uint32_t *__restrict__ adj;
    const uint32_t *__restrict__ labels;   // old labels
    const int4 *__restrict__ stats;        // (size, ext, int, -) of the old labels
    double p_ext;                          // Graph::clustering_coef
    uint32_t *__restrict__ labels_out;     // new labels (nullptr: do not write)
    unsigned long long *__restrict__ counters;  // [0] += evaluated pairs
    uint32_t *__restrict__ ex_sorted;      // export: sorted keys at adj positions
    uint32_t *__restrict__ ex_count;       // export: run length at heads, else 0
    // batched arenas (nullptr for one graph): graph of each vertex, each
    // graph's clustering coefficient and whether it is still refining
    const uint32_t *__restrict__ gid;
    const double *__restrict__ cc_g;
    const uint32_t *__restrict__ active_g;
};

// Vertices of a batched graph that has stopped are skipped (their label
// stays).
__device__ __forceinline__ bool vtx_stopped(const MoveArgs &a, uint32_t p) {
    return a.gid && !a.active_g[a.gid[p]];
}

__device__ __forceinline__ double vtx_pext(const MoveArgs &a, uint32_t p) {
    return a.gid ? a.cc_g[a.gid[p]] : a.p_ext;
}

// A vertex of a graph that has stopped keeps its label.
__device__ __forceinline__ uint32_t vtx_result(const MoveArgs &a, uint32_t p, uint32_t moved, uint32_t my) {
    return (a.gid && !a.active_g[a.gid[p]]) ? my : moved;
}

template <int S>
__device__ __forceinline__ uint64_t group_bits(uint64_t ballot, int gbase) {
    if constexpr (S == 64) {
        return ballot;
    } else {
        return (ballot >> gbase) & ((1ull << S) - 1ull);
    }
}

__device__ __forceinline__ void count_pairs(unsigned long long *counters, uint32_t heads_in_lane) {
    // wave sum, then one atomic per wave; only in counting passes (the
    // atomics all hit one address)
    if (!counters) return;
    uint32_t s = heads_in_lane;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, SCD_WAVE);
    if ((threadIdx.x & 63) == 0 && counters) atomicAdd(counters, (unsigned long long)s);
}

// ---------------------------------------------------------------------------
// S lanes per vertex, deg <= S < 64.
// ---------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(256) void k_move_sub(MoveArgs a, const uint32_t *__restrict__ verts,
                                                  uint32_t nv) {
    const int lane = threadIdx.x & 63;
    const int lig = threadIdx.x & (S - 1);
    const int gbase = lane & ~(S - 1);
    const uint32_t slot = (blockIdx.x * 256u + threadIdx.x) / S;
    const bool active = slot < nv;
    const uint32_t p = active ? verts[slot] : 0u;
    const bool live = active && !vtx_stopped(a, p);
    const uint32_t beg = live ? a.indices[p] : 0u;
    const uint32_t deg = live ? a.indices[p + 1] - beg : 0u;
    const uint32_t my = active ? a.labels[p] : SCD_INVALID;
    const double pext = live ? vtx_pext(a, p) : 0.0;
    uint32_t v = SCD_INVALID;
    if ((uint32_t)lig < deg) v = a.labels[a.adj[beg + lig]];
    v = group_bitonic_sort<S>(v, lig);

    const uint32_t prev = __shfl(v, (lane + 63) & 63, SCD_WAVE);
    const bool boundary = (lig == 0) || (v != prev);
    const uint64_t gm = group_bits<S>(__ballot(boundary), gbase);
    const uint64_t above = (lig + 1 < S) ? (gm >> (lig + 1)) : 0ull;
    const uint32_t next = above ? (uint32_t)(lig + 1 + __builtin_ctzll(above)) : (uint32_t)S;
    const uint32_t freq = next - (uint32_t)lig;
    const bool head = boundary && v != SCD_INVALID;

    double imp = 0.0;
    if (head) imp = candidate_improvement(v, freq, deg, my, a.stats, pext);

    const uint64_t mm = group_bits<S>(__ballot(head && v == my), gbase);
    const int src = mm ? gbase + __builtin_ctzll(mm) : lane;
    const double ximp = __shfl(imp, src, SCD_WAVE);
    const double rem = mm ? -ximp : 0.0;

    double bi = 0.0;
    uint32_t bl = SCD_INVALID;
    if (head && imp > 0.0) {
        bi = imp;
        bl = v;
    }
    group_argmax<S>(bi, bl);

    if (active && lig == 0 && a.labels_out) a.labels_out[p] = vtx_result(a, p, move_rule(bi, bl, rem, my), my);
    if (a.ex_sorted && (uint32_t)lig < deg) {
        a.ex_sorted[beg + lig] = v;
        a.ex_count[beg + lig] = head ? freq : 0u;
    }
    count_pairs(a.counters, head ? 1u : 0u);
}

// ---------------------------------------------------------------------------
// One wave per vertex, deg <= 64 * R; key i in register i % R of lane i / R
// (blocked: each lane gathers R consecutive neighbours and the short sort
// distances stay in registers).
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(256) void k_move_wave(MoveArgs a, const uint32_t *__restrict__ verts,
                                                   uint32_t nv) {
    const int lane = threadIdx.x & 63;
    const uint32_t slot = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (slot >= nv) return;  // uniform per wave
    const uint32_t p = verts[slot];
    const uint32_t my = a.labels[p];
    if (vtx_stopped(a, p)) {  // uniform per wave
        if (lane == 0 && a.labels_out) a.labels_out[p] = my;
        return;
    }
    const uint32_t beg = a.indices[p];
    const uint32_t deg = a.indices[p + 1] - beg;
    const double pext = vtx_pext(a, p);
    const uint32_t i0 = (uint32_t)(lane * R);  // index of the lane's first key

    uint32_t x[R];
    {
        uint32_t q[R];
#pragma unroll
        for (int r = 0; r < R; r++) q[r] = (i0 + r < deg) ? a.adj[beg + i0 + r] : 0u;
#pragma unroll
        for (int r = 0; r < R; r++) x[r] = (i0 + r < deg) ? a.labels[q[r]] : SCD_INVALID;
    }
    wave_bitonic_sort_blocked<R>(x, lane);

    // run heads: key i differs from key i - 1 (key 0 always)
    const uint32_t prev_lane_last = __shfl_up(x[R - 1], 1, SCD_WAVE);
    bool bnd[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t prev = r ? x[r - 1] : prev_lane_last;
        bnd[r] = (i0 + r == 0) || (x[r] != prev);
    }
    // the first head of every lane (64 R: none), for runs ending in a later lane
    uint32_t first = 64u * R;
#pragma unroll
    for (int r = R - 1; r >= 0; r--)
        if (bnd[r]) first = i0 + r;
    const uint64_t has = __ballot(first != 64u * R);
    const uint64_t later = lane < 63 ? (has >> (lane + 1)) : 0ull;
    const int nl = later ? lane + 1 + __builtin_ctzll(later) : lane;
    const uint32_t nfirst = __shfl(first, nl, SCD_WAVE);
    const uint32_t next_lane_head = later ? nfirst : 64u * R;

    double bi = 0.0, rem_lane = 0.0;
    uint32_t bl = SCD_INVALID;
    bool has_rem = false;
    uint32_t heads = 0;
    uint32_t freq[R];
    {
        uint32_t next = next_lane_head;  // next head after register r, scanning down
#pragma unroll
        for (int r = R - 1; r >= 0; r--) {
            freq[r] = (bnd[r] && x[r] != SCD_INVALID) ? next - (i0 + r) : 0u;
            if (bnd[r]) next = i0 + r;
        }
    }
    // the statistics loads of SB candidates before their evaluations (all R
    // at once for R <= 8; two batches at R = 16 keep the kernel at 4 waves
    // per SIMD instead of 3)
    constexpr int SB = R < 4 ? R : 4;
#pragma unroll
    for (int r0 = 0; r0 < R; r0 += SB) {
        int4 st[SB];
#pragma unroll
        for (int q = 0; q < SB; q++) st[q] = freq[r0 + q] ? a.stats[x[r0 + q]] : make_int4(0, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < SB; q++) {
            const int r = r0 + q;
            const uint32_t i = i0 + r;
            if (freq[r]) {
                const double imp = candidate_improvement_s(st[q], x[r], freq[r], deg, my, pext);
                if (x[r] == my) {
                    rem_lane = -imp;
                    has_rem = true;
                }
                if (imp > 0.0 && better(imp, x[r], bi, bl)) {
                    bi = imp;
                    bl = x[r];
                }
                heads++;
            }
            if (a.ex_sorted && i < deg) {
                a.ex_sorted[beg + i] = x[r];
                a.ex_count[beg + i] = freq[r];
            }
        }
    }
    group_argmax<64>(bi, bl);
    const uint64_t mm = __ballot(has_rem);
    const int src = mm ? __builtin_ctzll(mm) : lane;
    const double xr = __shfl(rem_lane, src, SCD_WAVE);
    const double rem = mm ? xr : 0.0;
    if (lane == 0 && a.labels_out) a.labels_out[p] = vtx_result(a, p, move_rule(bi, bl, rem, my), my);
    count_pairs(a.counters, heads);
}

// Block-wide argmax of (bi, bl) over 256 threads; thread 0 gets the result.
__device__ __forceinline__ void block_argmax256(double &bi, uint32_t &bl, double *s_bi,
                                                uint32_t *s_bl) {
    group_argmax<64>(bi, bl);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_bi[w] = bi;
        s_bl[w] = bl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++)
            if (better(s_bi[k], s_bl[k], bi, bl)) {
                bi = s_bi[k];
                bl = s_bl[k];
            }
    }
}

// Reduce-by-key + evaluation over positions [lo, hi) of a sorted key array
// S (length n) by 256 threads, four positions per thread at a time: run
// heads and run lengths first, then every statistics load, then the
// evaluations.  The own-label improvement goes to *s_rem (one writer).
__device__ __forceinline__ void eval_sorted_batched(const MoveArgs &a, const uint32_t *S, uint32_t n, uint32_t lo,
                                                    uint32_t hi, uint32_t deg, uint32_t my, uint32_t beg, double pext,
                                                    double &bi, uint32_t &bl, uint32_t &heads, double *s_rem,
                                                    int *s_has_rem) {
    constexpr int BATCH = 4;
    for (uint32_t base = lo; base < hi; base += 256 * BATCH) {
        uint32_t v[BATCH], fr[BATCH];
        int4 st[BATCH];
#pragma unroll
        for (int k = 0; k < BATCH; k++) {
            const uint32_t i = base + k * 256 + threadIdx.x;
            v[k] = SCD_INVALID;
            fr[k] = 0;
            if (i < hi) {
                v[k] = S[i];
                if ((i == 0 || S[i - 1] != v[k]) && v[k] != SCD_INVALID) fr[k] = run_end_u32(S, n, i, v[k]) - i;
            }
        }
#pragma unroll
        for (int k = 0; k < BATCH; k++) st[k] = fr[k] ? a.stats[v[k]] : make_int4(0, 0, 0, 0);
#pragma unroll
        for (int k = 0; k < BATCH; k++) {
            const uint32_t i = base + k * 256 + threadIdx.x;
            if (fr[k]) {
                const double imp = candidate_improvement_s(st[k], v[k], fr[k], deg, my, pext);
                if (v[k] == my) {
                    *s_rem = -imp;
                    *s_has_rem = 1;
                }
                if (imp > 0.0 && better(imp, v[k], bi, bl)) {
                    bi = imp;
                    bl = v[k];
                }
                heads++;
            }
            if (a.ex_sorted && i < hi) {
                a.ex_sorted[beg + i] = v[k];
                a.ex_count[beg + i] = fr[k];
            }
        }
    }
}

__device__ __forceinline__ uint32_t next_pow2(uint32_t x) {
    return x <= 2 ? 2u : (1u << (32 - __builtin_clz(x - 1)));
}

// keys[i] = labels[adj[beg + i]] for i < deg, INVALID up to n2; all of a
// thread's loads are issued before any is used.  Ends with a barrier.
template <int T>
__device__ __forceinline__ void gather_labels(const MoveArgs &a, uint32_t beg, uint32_t deg, uint32_t n2,
                                              uint32_t *keys) {
    constexpr int U = T / 256;
    uint32_t q[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const uint32_t i = u * 256 + threadIdx.x;
        q[u] = i < deg ? a.adj[beg + i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const uint32_t i = u * 256 + threadIdx.x;
        if (i < n2) keys[i] = i < deg ? a.labels[q[u]] : SCD_INVALID;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// One workgroup (256 threads) per vertex, deg <= T: register sorts of
// 1024-key chunks and merge passes in LDS.
// ---------------------------------------------------------------------------
template <int T>
__global__ __launch_bounds__(256) void k_move_lds(MoveArgs a, const uint32_t *__restrict__ verts,
                                                  uint32_t nv) {
    __shared__ uint32_t keys[T];
    __shared__ double s_bi[4];
    __shared__ uint32_t s_bl[4];
    __shared__ double s_rem;
    __shared__ int s_has_rem;
    const int tid = threadIdx.x;
    const uint32_t p = verts[blockIdx.x];
    const uint32_t my = a.labels[p];
    if (vtx_stopped(a, p)) {  // uniform per block, before any barrier
        if (tid == 0 && a.labels_out) a.labels_out[p] = my;
        return;
    }
    const uint32_t beg = a.indices[p];
    const uint32_t deg = a.indices[p + 1] - beg;
    const uint32_t n2 = max(next_pow2(deg), 1024u);
    if (tid == 0) s_has_rem = 0;
    gather_labels<T>(a, beg, deg, n2, keys);
    block_sort_lds<T>(keys, n2, deg);

    double bi = 0.0;
    uint32_t bl = SCD_INVALID;
    uint32_t heads = 0;
    eval_sorted_batched(a, keys, n2, 0, deg, deg, my, beg, vtx_pext(a, p), bi, bl, heads, &s_rem, &s_has_rem);
    count_pairs(a.counters, heads);
    block_argmax256(bi, bl, s_bi, s_bl);
    if (tid == 0 && a.labels_out) {
        const double rem = s_has_rem ? s_rem : 0.0;
        a.labels_out[p] = vtx_result(a, p, move_rule(bi, bl, rem, my), my);
    }
}

// ---------------------------------------------------------------------------
// LDS bin without export, by hashing instead of sorting: the candidates are
// the distinct valid neighbour labels with their frequencies and the argmax
// order is total, so a (label, count) table in LDS replaces the sort (a few
// instructions per neighbour instead of the bitonic stages).  One workgroup
// per vertex; slots hold label << 32 | count (64-bit LDS atomics), at least
// 1.5 deg of them (at most 1.5 T: 48 KB for T = 4096), linear probing from a
// multiplicative hash scaled to the table size.  The occupied slots are then
// compacted in place (ballot ranks, 256 slots a round), so the fp64
// evaluations run on dense lanes.
// ---------------------------------------------------------------------------
template <int T>
__global__ __launch_bounds__(256) void k_move_lds_hash(MoveArgs a, const uint32_t *__restrict__ verts,
                                                       uint32_t nv) {
    constexpr uint32_t CAPMAX = T + T / 2;
    constexpr unsigned long long EMPTY = ~0ull;  // label SCD_INVALID is never inserted
    __shared__ unsigned long long s_tab[CAPMAX];
    __shared__ double s_bi[4];
    __shared__ uint32_t s_bl[4];
    __shared__ uint32_t s_wn[4];
    __shared__ double s_rem;
    __shared__ int s_has_rem;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t p = verts[blockIdx.x];
    const uint32_t my = a.labels[p];
    if (vtx_stopped(a, p)) {  // uniform per block, before any barrier
        if (tid == 0 && a.labels_out) a.labels_out[p] = my;
        return;
    }
    const uint32_t beg = a.indices[p];
    const uint32_t deg = a.indices[p + 1] - beg;
    const uint32_t cap = min(CAPMAX, (deg + deg / 2 + 255u) & ~255u);
    for (uint32_t i = tid; i < cap; i += 256) s_tab[i] = EMPTY;
    if (tid == 0) s_has_rem = 0;
    constexpr int U = T / 256;
    uint32_t q[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const uint32_t i = u * 256 + tid;
        q[u] = i < deg ? a.adj[beg + i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < U; u++) q[u] = (u * 256 + tid) < deg ? a.labels[q[u]] : SCD_INVALID;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; u++) {
        const uint32_t L = q[u];
        if (L == SCD_INVALID) continue;
        uint32_t h = (uint32_t)(((uint64_t)(L * 0x9E3779B1u) * cap) >> 32);
        while (true) {
            unsigned long long cur = s_tab[h];
            if (cur == EMPTY) {
                cur = atomicCAS(&s_tab[h], EMPTY, ((unsigned long long)L << 32) | 1ull);
                if (cur == EMPTY) break;  // inserted with count 1
            }
            if ((uint32_t)(cur >> 32) == L) {
                atomicAdd(&s_tab[h], 1ull);
                break;
            }
            h = h + 1 == cap ? 0u : h + 1;
        }
    }
    __syncthreads();
    // compaction: round r reads slots [256 r, 256 r + 256) and writes them
    // below, at the count of occupied slots so far (never past what was read)
    uint32_t nu = 0;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint32_t r0 = 0; r0 < cap; r0 += 256) {
        const unsigned long long e = s_tab[r0 + tid];  // cap is a multiple of 256
        const uint64_t occ = __ballot(e != EMPTY);
        if (lane == 0) s_wn[wv] = (uint32_t)__builtin_popcountll(occ);
        __syncthreads();
        uint32_t off = nu;
        for (int k = 0; k < wv; k++) off += s_wn[k];
        if (e != EMPTY) s_tab[off + (uint32_t)__builtin_popcountll(occ & lt)] = e;
        nu += s_wn[0] + s_wn[1] + s_wn[2] + s_wn[3];
        __syncthreads();
    }
    double bi = 0.0;
    uint32_t bl = SCD_INVALID;
    uint32_t heads = 0;
    const double pext = vtx_pext(a, p);
    constexpr int BATCH = 4;
    for (uint32_t base = 0; base < nu; base += 256 * BATCH) {
        uint32_t v[BATCH], fr[BATCH];
        int4 st[BATCH];
#pragma unroll
        for (int k = 0; k < BATCH; k++) {
            const uint32_t i = base + k * 256 + tid;
            const unsigned long long e = i < nu ? s_tab[i] : EMPTY;
            v[k] = (uint32_t)(e >> 32);
            fr[k] = i < nu ? (uint32_t)e : 0u;
        }
#pragma unroll
        for (int k = 0; k < BATCH; k++) st[k] = fr[k] ? a.stats[v[k]] : make_int4(0, 0, 0, 0);
#pragma unroll
        for (int k = 0; k < BATCH; k++) {
            if (!fr[k]) continue;
            const double imp = candidate_improvement_s(st[k], v[k], fr[k], deg, my, pext);
            if (v[k] == my) {
                s_rem = -imp;
                s_has_rem = 1;
            }
            if (imp > 0.0 && better(imp, v[k], bi, bl)) {
                bi = imp;
                bl = v[k];
            }
            heads++;
        }
    }
    count_pairs(a.counters, heads);
    block_argmax256(bi, bl, s_bi, s_bl);
    if (tid == 0 && a.labels_out) {
        const double rem = s_has_rem ? s_rem : 0.0;
        a.labels_out[p] = vtx_result(a, p, move_rule(bi, bl, rem, my), my);
    }
}

// ---------------------------------------------------------------------------
// Hub path (deg > T).  Hub h has vertex hub_v[h], degree hub_d[h] and a
// scratch segment at hub_off[h] in two ping-pong buffers.  Work items:
//   tiles  (tile_hub[t], tile_start[t]): T keys each, sorted in LDS;
//   chunks (chunk_hub[c], chunk_start[c]): C outputs each for merge / eval.
// ---------------------------------------------------------------------------
struct HubArgs {
    const uint32_t *__restrict__ hub_v;
    const uint32_t *__restrict__ hub_d;
    const uint32_t *__restrict__ hub_off;
    const uint32_t *__restrict__ hub_chunk_begin;  // nh + 1
    const uint32_t *__restrict__ tile_hub;
    const uint32_t *__restrict__ tile_start;
    const uint32_t *__restrict__ chunk_hub;
    const uint32_t *__restrict__ chunk_start;
    double *__restrict__ chunk_bi;
    uint32_t *__restrict__ chunk_bl;
    double *__restrict__ chunk_rem;
    uint32_t *__restrict__ chunk_has_rem;
};

template <int T>
__global__ __launch_bounds__(256) void k_hub_tile_sort(MoveArgs a, HubArgs h,
                                                       uint32_t *__restrict__ out) {
    __shared__ uint32_t keys[T];
    const int tid = threadIdx.x;
    const uint32_t hb = h.tile_hub[blockIdx.x];
    const uint32_t st = h.tile_start[blockIdx.x];
    const uint32_t p = h.hub_v[hb];
    const uint32_t d = h.hub_d[hb];
    const uint32_t beg = a.indices[p] + st;
    const uint32_t len = min((uint32_t)T, d - st);
    const uint32_t n2 = max(next_pow2(len), 1024u);
    gather_labels<T>(a, beg, len, n2, keys);
    block_sort_lds<T>(keys, n2, len);
    uint32_t *dst = out + h.hub_off[hb] + st;
    for (uint32_t i = tid; i < len; i += 256) dst[i] = keys[i];
    (void)tid;
}

// One merge pass: sorted runs of width w -> 2w inside every hub segment.
// C (outputs per workgroup) is a power of two with C <= 2w, so a chunk never
// straddles two run pairs.  The chunk goes HM_SUB outputs at a time: the
// co-ranks of every sub-chunk boundary first (one merge-path search in
// global memory each, in parallel), then per sub-chunk its two input runs
// are read into LDS, merged there four outputs per thread, and written out
// (coalesced reads and writes instead of per-thread sequential merges in
// global memory).
constexpr uint32_t HM_SUB = 1024;
constexpr uint32_t HM_SPLITS = 32768 / HM_SUB + 1;  // merge_chunk <= 2 x 16384

__global__ __launch_bounds__(256) void k_hub_merge(HubArgs h, const uint32_t *__restrict__ src,
                                                   uint32_t *__restrict__ dst, uint32_t w, uint32_t C) {
    __shared__ uint32_t s_in[HM_SUB], s_out[HM_SUB];
    __shared__ uint32_t s_split[HM_SPLITS];
    const int tid = threadIdx.x;
    const uint32_t hb = h.chunk_hub[blockIdx.x];
    const uint32_t o = h.chunk_start[blockIdx.x];
    const uint32_t d = h.hub_d[hb];
    const uint32_t base = h.hub_off[hb];
    const uint32_t ps = (o / (2 * w)) * (2 * w);
    const uint32_t la = min(w, d - ps);
    const uint32_t lb = (d - ps > w) ? min(w, d - ps - w) : 0u;
    const uint32_t *A = src + base + ps;
    const uint32_t *B = A + la;
    uint32_t *D = dst + base + ps;
    const uint32_t out_lo = o - ps;
    const uint32_t out_hi = min(o + C, ps + la + lb) - ps;
    if (out_lo >= out_hi) return;  // uniform
    const uint32_t ns = (out_hi - out_lo + HM_SUB - 1) / HM_SUB;
    for (uint32_t s = tid; s <= ns; s += 256) s_split[s] = merge_path(A, la, B, lb, min(out_lo + s * HM_SUB, out_hi));
    __syncthreads();
    for (uint32_t s = 0; s < ns; s++) {
        const uint32_t d0 = out_lo + s * HM_SUB, d1 = min(d0 + HM_SUB, out_hi);
        const uint32_t i0 = s_split[s], i1 = s_split[s + 1];
        const uint32_t na = i1 - i0, nb = (d1 - i1) - (d0 - i0), n = d1 - d0;
        for (uint32_t k = tid; k < na; k += 256) s_in[k] = A[i0 + k];
        for (uint32_t k = tid; k < nb; k += 256) s_in[na + k] = B[d0 - i0 + k];
        __syncthreads();
        const uint32_t t0 = tid * (HM_SUB / 256);
        if (t0 < n) {
            const uint32_t *a = s_in, *b = s_in + na;
            uint32_t i = merge_path(a, na, b, nb, t0);
            uint32_t j = t0 - i;
#pragma unroll
            for (uint32_t k = 0; k < HM_SUB / 256; k++) {
                if (t0 + k < n) {
                    const bool takeA = (j >= nb) || (i < na && a[i] <= b[j]);
                    s_out[t0 + k] = takeA ? a[i++] : b[j++];
                }
            }
        }
        __syncthreads();
        for (uint32_t k = tid; k < n; k += 256) D[d0 + k] = s_out[k];
        __syncthreads();  // s_in and s_out are reused
    }
}

// Reduce-by-key + evaluation over a chunk of a hub's sorted segment; the
// chunk's best candidate and (if present) the own-label improvement go to
// per-chunk slots.  The tiles of T keys were sorted into buf_a; a hub of
// degree d then took one merge pass per width T, 2T, ... below d, each
// moving its keys to the other buffer.
__global__ __launch_bounds__(256) void k_hub_eval(MoveArgs a, HubArgs h, const uint32_t *__restrict__ buf_a,
                                                  const uint32_t *__restrict__ buf_b, uint32_t T, uint32_t C) {
    __shared__ double s_bi[4];
    __shared__ uint32_t s_bl[4];
    __shared__ double s_rem;
    __shared__ int s_has_rem;
    const int tid = threadIdx.x;
    const uint32_t c = blockIdx.x;
    const uint32_t hb = h.chunk_hub[c];
    const uint32_t o = h.chunk_start[c];
    const uint32_t d = h.hub_d[hb];
    const uint32_t p = h.hub_v[hb];
    const uint32_t my = a.labels[p];
    const uint32_t beg = a.indices[p];
    uint32_t passes = 0;
    for (uint64_t w = T; w < d; w *= 2) passes++;
    const uint32_t *S = ((passes & 1u) ? buf_b : buf_a) + h.hub_off[hb];
    if (tid == 0) s_has_rem = 0;
    __syncthreads();
    double bi = 0.0;
    uint32_t bl = SCD_INVALID;
    uint32_t heads = 0;
    const uint32_t end = min(o + C, d);
    eval_sorted_batched(a, S, d, o, end, d, my, beg, vtx_pext(a, p), bi, bl, heads, &s_rem, &s_has_rem);
    count_pairs(a.counters, heads);
    block_argmax256(bi, bl, s_bi, s_bl);
    if (tid == 0) {
        h.chunk_bi[c] = bi;
        h.chunk_bl[c] = bl;
        h.chunk_has_rem[c] = (uint32_t)s_has_rem;
        h.chunk_rem[c] = s_has_rem ? s_rem : 0.0;
    }
}

// ---------------------------------------------------------------------------
// Hub path of the move step without export: the candidates of a hub are its
// distinct valid neighbour labels with their frequencies, and the argmax
// order (improvement desc, label asc) is total, so they need no global sort.
// Each tile of T neighbour labels is sorted in LDS and every run inserts
// (label, run length) into the hub's hash table in global memory (one atomic
// per run, so a frequent label is one hot slot per tile, not per entry);
// then chunks of table slots are evaluated like sorted runs, and each slot is
// reset for the next iteration.  Slot = label << 32 | count.
// ---------------------------------------------------------------------------
constexpr unsigned long long HUB_EMPTY = 0xFFFFFFFF00000000ull;

struct HubTable {
    unsigned long long *__restrict__ slots;
    const uint64_t *__restrict__ tab_off;  // per hub
    const uint32_t *__restrict__ tab_cap;  // per hub, a power of two
    // chunks of table slots: hub, first slot; results per chunk
    const uint32_t *__restrict__ tchunk_hub;
    const uint32_t *__restrict__ tchunk_start;
};

__device__ __forceinline__ uint32_t hub_hash(uint32_t L, uint32_t mask) { return (L * 0x9E3779B1u) & mask; }

__device__ __forceinline__ void hub_table_add(unsigned long long *tab, uint32_t mask, uint32_t L, uint32_t cnt) {
    const unsigned long long key = (unsigned long long)L << 32;
    for (uint32_t i = hub_hash(L, mask);; i = (i + 1) & mask) {
        unsigned long long cur = tab[i];
        if (cur == HUB_EMPTY) {
            cur = atomicCAS(&tab[i], HUB_EMPTY, key | cnt);
            if (cur == HUB_EMPTY) return;
        }
        if ((cur >> 32) == L) {
            atomicAdd(&tab[i], (unsigned long long)cnt);
            return;
        }
    }
}

template <int T>
__global__ __launch_bounds__(256) void k_hub_tile_count(MoveArgs a, HubArgs h, HubTable t) {
    __shared__ uint32_t keys[T];
    const int tid = threadIdx.x;
    const uint32_t hb = h.tile_hub[blockIdx.x];
    const uint32_t st = h.tile_start[blockIdx.x];
    const uint32_t p = h.hub_v[hb];
    if (vtx_stopped(a, p)) return;  // uniform per block
    const uint32_t d = h.hub_d[hb];
    const uint32_t beg = a.indices[p] + st;
    const uint32_t len = min((uint32_t)T, d - st);
    const uint32_t n2 = max(next_pow2(len), 1024u);
    gather_labels<T>(a, beg, len, n2, keys);
    block_sort_lds<T>(keys, n2, len);
    unsigned long long *tab = t.slots + t.tab_off[hb];
    const uint32_t mask = t.tab_cap[hb] - 1u;
    for (uint32_t i = tid; i < len; i += 256) {
        const uint32_t L = keys[i];
        if (L == SCD_INVALID || (i > 0 && keys[i - 1] == L)) continue;
        hub_table_add(tab, mask, L, run_end_u32(keys, len, i, L) - i);
    }
}

// Evaluation of a chunk of C table slots (every slot reset afterwards).
__global__ __launch_bounds__(256) void k_hub_table_eval(MoveArgs a, HubArgs h, HubTable t, uint32_t C) {
    __shared__ double s_bi[4];
    __shared__ uint32_t s_bl[4];
    __shared__ double s_rem;
    __shared__ int s_has_rem;
    const int tid = threadIdx.x;
    const uint32_t c = blockIdx.x;
    const uint32_t hb = t.tchunk_hub[c];
    const uint32_t o = t.tchunk_start[c];
    const uint32_t d = h.hub_d[hb];
    const uint32_t p = h.hub_v[hb];
    const uint32_t my = a.labels[p];
    const double pext = vtx_pext(a, p);
    unsigned long long *tab = t.slots + t.tab_off[hb];
    const uint32_t end = min(o + C, t.tab_cap[hb]);
    if (tid == 0) s_has_rem = 0;
    __syncthreads();
    double bi = 0.0;
    uint32_t bl = SCD_INVALID;
    uint32_t heads = 0;
    for (uint32_t i = o + tid; i < end; i += 256) {
        const unsigned long long e = tab[i];
        if (e == HUB_EMPTY) continue;
        tab[i] = HUB_EMPTY;
        const uint32_t L = (uint32_t)(e >> 32), f = (uint32_t)e;
        const double imp = candidate_improvement_s(a.stats[L], L, f, d, my, pext);
        if (L == my) {
            s_rem = -imp;
            s_has_rem = 1;
        }
        if (imp > 0.0 && better(imp, L, bi, bl)) {
            bi = imp;
            bl = L;
        }
        heads++;
    }
    count_pairs(a.counters, heads);
    block_argmax256(bi, bl, s_bi, s_bl);
    if (tid == 0) {
        h.chunk_bi[c] = bi;
        h.chunk_bl[c] = bl;
        h.chunk_has_rem[c] = (uint32_t)s_has_rem;
        h.chunk_rem[c] = s_has_rem ? s_rem : 0.0;
    }
}

// One thread per hub: combine the hub's chunks and apply the move rule.
// One wave per hub (the largest hubs of R-MAT s24 have hundreds of chunks):
// the chunks' winners reduce in the total (improvement, label) order, and
// the own-label improvement comes from the last chunk that has one (there
// is one: the chunk of the own label's run head).
__global__ __launch_bounds__(256) void k_hub_finalize(MoveArgs a, HubArgs h, uint32_t nh) {
    const uint32_t hb = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (hb >= nh) return;  // uniform per wave
    const int lane = threadIdx.x & 63;
    double bi = 0.0, rem = 0.0;
    uint32_t bl = SCD_INVALID;
    int rem_c = -1;  // the lane's last chunk with an own-label improvement
    for (uint32_t c = h.hub_chunk_begin[hb] + lane; c < h.hub_chunk_begin[hb + 1]; c += 64) {
        if (better(h.chunk_bi[c], h.chunk_bl[c], bi, bl)) {
            bi = h.chunk_bi[c];
            bl = h.chunk_bl[c];
        }
        if (h.chunk_has_rem[c]) {
            rem = h.chunk_rem[c];
            rem_c = (int)c;
        }
    }
    group_argmax<64>(bi, bl);
    int last = rem_c;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) last = max(last, __shfl_xor(last, off, SCD_WAVE));
    const uint64_t owner = __ballot(rem_c == last && last >= 0);
    const double r = owner ? __shfl(rem, __builtin_ctzll(owner), SCD_WAVE) : 0.0;
    if (lane == 0 && a.labels_out) {
        const uint32_t p = h.hub_v[hb];
        const uint32_t my = a.labels[p];
        a.labels_out[p] = vtx_result(a, p, move_rule(bi, bl, r, my), my);
    }
}

}  // namespace scd_dev

#endif
