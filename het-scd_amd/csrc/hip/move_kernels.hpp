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

// ---------------------------------------------------------------------------
// Hub path by label-range buckets (move without export).  A hub's neighbour
// labels are split at sampled label values into buckets of about HB_TARGET
// entries; each bucket is sorted and evaluated in LDS by one workgroup, and
// k_hub_finalize combines the buckets' best candidates per hub.  One label
// gather, one scatter and one LDS sort per entry: no tiles merged in global
// memory.
//   k_hb_split: per hub, at most HB_SAMPLES labels sampled (every sigma-th
//     neighbour) and sorted in LDS; every kstep-th sample value starts a
//     bucket, and a value met at kstep or more samples gets a bucket of its
//     own (a single label of any multiplicity is one candidate: counted,
//     never sorted).  Clears the hub's bucket counters.
//   k_hb_part: per chunk of HB_PCH entries, the labels gathered and counted
//     per bucket in LDS; one atomic per bucket reserves the chunk's range in
//     the bucket's region of HB_CAP entries (fewer only when a test asks),
//     and entries past it go to the hub's overflow area.  Single-label
//     buckets are only counted.
//   k_hb_eval: per bucket slot: one label -> one evaluation; at most HB_CAP
//     entries -> sort, reduce-by-key and evaluation in LDS; more (a sampling
//     outlier: a bucket aims at HB_TARGET entries and holds at most about
//     twice that in expectation) -> its distinct labels one at a time by block
//     reductions over the region and the overflow area, slower but exact.
// ---------------------------------------------------------------------------
constexpr uint32_t HB_SAMPLES = 4096;  // sampled labels per hub, at most
constexpr uint32_t HB_TARGET = 1024;   // entries per bucket aimed at
constexpr uint32_t HB_CAP = 4096;      // entries per bucket region
constexpr uint32_t HB_PCH = 2048;      // entries per partition workgroup
constexpr uint32_t HB_MAXB = 2048;     // buckets per hub, at most
static_assert(HB_MAXB <= 65536, "k_hb_part keeps bucket ids in 16 bits");

struct HubBuckets {
    const uint32_t *__restrict__ sigma;      // per hub: sample stride
    const uint32_t *__restrict__ kstep;      // per hub: samples per bucket start
    const uint32_t *__restrict__ slot0;      // per hub (nh + 1): first bucket slot
    const uint32_t *__restrict__ slot_hub;   // per slot: its hub
    const uint32_t *__restrict__ pch_hub;    // per partition chunk: hub, first entry
    const uint32_t *__restrict__ pch_start;
    uint32_t *__restrict__ lo;       // per slot: bucket j holds labels [lo[j], lo[j + 1])
    uint32_t *__restrict__ nb;       // per hub: buckets in use
    uint32_t *__restrict__ cnt;      // per slot: entries
    uint32_t *__restrict__ ovf_cnt;  // per hub: entries in the overflow area
    uint32_t *__restrict__ region;   // HB_CAP entries per slot
    uint32_t *__restrict__ ovf;      // overflow area of hub h at hub_off[h]
    uint32_t cap;                    // entries a region takes (HB_CAP; less only to test the overflow path)
};

// The bucket of label L: the largest j < K with lo[j] <= L (lo[0] = 0).
__device__ __forceinline__ uint32_t hb_bucket(const uint32_t *lo, uint32_t K, uint32_t L) {
    uint32_t a = 0, b = K;
    while (b - a > 1) {
        const uint32_t mid = (a + b) >> 1;
        if (lo[mid] <= L) a = mid;
        else b = mid;
    }
    return a;
}

__device__ __forceinline__ uint32_t hb_hi(const uint32_t *lo, uint32_t K, uint32_t j) {
    return j + 1 < K ? lo[j + 1] : SCD_INVALID;
}

__global__ __launch_bounds__(256) void k_hb_split(MoveArgs a, HubArgs h, HubBuckets b) {
    __shared__ uint32_t s[HB_SAMPLES];
    __shared__ uint32_t s_off[257];
    const int tid = threadIdx.x;
    const uint32_t hb = blockIdx.x;
    const uint32_t p = h.hub_v[hb], d = h.hub_d[hb];
    const uint32_t s0 = b.slot0[hb], ns = b.slot0[hb + 1] - s0;
    for (uint32_t i = tid; i < ns; i += 256) b.cnt[s0 + i] = 0;
    if (tid == 0) b.ovf_cnt[hb] = 0;
    if (vtx_stopped(a, p)) {  // uniform per block
        if (tid == 0) b.nb[hb] = 0;
        return;
    }
    const uint32_t sg = b.sigma[hb], k = b.kstep[hb];
    const uint32_t beg = a.indices[p];
    const uint32_t m = (d + sg - 1) / sg;  // <= HB_SAMPLES (host)
    const uint32_t n2 = max(next_pow2(m), 1024u);
    {
        constexpr int U = HB_SAMPLES / 256;
        uint32_t q[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t j = u * 256 + tid;
            q[u] = j < m ? a.adj[beg + j * sg] : 0u;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t j = u * 256 + tid;
            if (j < n2) s[j] = j < m ? a.labels[q[u]] : SCD_INVALID;
        }
    }
    __syncthreads();
    block_sort_lds<HB_SAMPLES>(s, n2, m);
    // valid samples: those before the first INVALID
    uint32_t mv;
    {
        uint32_t lo_ = 0, hi_ = m;
        while (lo_ < hi_) {
            const uint32_t mid = (lo_ + hi_) >> 1;
            if (s[mid] != SCD_INVALID) lo_ = mid + 1;
            else hi_ = mid;
        }
        mv = lo_;
    }
    // bucket starts: candidate j (1..J) is the sample value v_j = s[j k]; it
    // starts a bucket unless it equals v_{j-1} (or is 0, the first start);
    // a value met at k or more samples also ends its own bucket at v + 1
    // (after its last candidate, unless v + 1 starts the next one anyway)
    const uint32_t J = mv ? (mv - 1) / k : 0u;
    const uint32_t per = (J + 255) / 256;
    const uint32_t j0 = 1 + tid * per, j1 = min(J, (tid + 1) * per);  // [j0, j1]
    auto heavy = [&](uint32_t v) {
        uint32_t l0 = 0, h0 = mv;  // first index with s >= v
        while (l0 < h0) {
            const uint32_t mid = (l0 + h0) >> 1;
            if (s[mid] < v) l0 = mid + 1;
            else h0 = mid;
        }
        return run_end_u32(s, mv, l0, v) - l0 >= k;
    };
    auto emits = [&](uint32_t j, uint32_t *out) {
        const uint32_t v = s[j * k];
        uint32_t e = 0;
        const bool fresh = j == 1 ? v != 0u : v != s[(j - 1) * k];
        if (fresh) out[e++] = v;
        const bool last = j == J || s[(j + 1) * k] != v;
        if (last && v + 1u != SCD_INVALID && (j == J || s[(j + 1) * k] != v + 1u) && heavy(v)) out[e++] = v + 1u;
        return e;
    };
    uint32_t mine = 0;
    for (uint32_t j = j0; j <= j1; j++) {
        uint32_t tmp[2];
        mine += emits(j, tmp);
    }
    s_off[tid + 1] = mine;
    __syncthreads();
    if (tid == 0) {
        s_off[0] = 1;  // slot 0 holds the start 0
        for (int t = 1; t <= 256; t++) s_off[t] += s_off[t - 1];
        b.lo[s0] = 0;
        b.nb[hb] = s_off[256];
    }
    __syncthreads();
    uint32_t at = s0 + s_off[tid];
    for (uint32_t j = j0; j <= j1; j++) {
        uint32_t tmp[2];
        const uint32_t e = emits(j, tmp);
        for (uint32_t i = 0; i < e; i++) b.lo[at++] = tmp[i];
    }
}

__global__ __launch_bounds__(256) void k_hb_part(MoveArgs a, HubArgs h, HubBuckets b) {
    __shared__ uint32_t keys[HB_PCH];
    __shared__ uint16_t s_bk[HB_PCH];  // each entry's bucket, found once
    __shared__ uint32_t s_lo[HB_MAXB], s_cnt[HB_MAXB], s_base[HB_MAXB];
    const int tid = threadIdx.x;
    const uint32_t hb = b.pch_hub[blockIdx.x], st = b.pch_start[blockIdx.x];
    const uint32_t p = h.hub_v[hb];
    if (vtx_stopped(a, p)) return;  // uniform per block
    const uint32_t d = h.hub_d[hb];
    const uint32_t s0 = b.slot0[hb], K = b.nb[hb];
    for (uint32_t i = tid; i < K; i += 256) {
        s_lo[i] = b.lo[s0 + i];
        s_cnt[i] = 0;
    }
    const uint32_t len = min(HB_PCH, d - st);
    gather_labels<HB_PCH>(a, a.indices[p] + st, len, len, keys);  // ends with a barrier
    for (uint32_t i = tid; i < len; i += 256) {
        const uint32_t L = keys[i];
        if (L != SCD_INVALID) {
            const uint32_t j = hb_bucket(s_lo, K, L);
            s_bk[i] = (uint16_t)j;
            atomicAdd(&s_cnt[j], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < K; i += 256) {
        const uint32_t c = s_cnt[i];
        s_base[i] = c ? atomicAdd(&b.cnt[s0 + i], c) : 0u;
        s_cnt[i] = 0;
    }
    __syncthreads();
    uint32_t *ov = b.ovf + h.hub_off[hb];
    for (uint32_t i = tid; i < len; i += 256) {
        const uint32_t L = keys[i];
        if (L == SCD_INVALID) continue;
        const uint32_t j = s_bk[i];
        if (hb_hi(s_lo, K, j) - s_lo[j] == 1u) continue;  // one label: counted only
        const uint32_t pos = s_base[j] + atomicAdd(&s_cnt[j], 1u);
        if (pos < b.cap) b.region[(size_t)(s0 + j) * HB_CAP + pos] = L;
        else ov[atomicAdd(&b.ovf_cnt[hb], 1u)] = L;
    }
}

// Minimum / sum over the 256 threads of a block, in every thread.
__device__ __forceinline__ uint32_t hb_block_min(uint32_t x, uint32_t *s_red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = min(x, (uint32_t)__shfl_xor(x, o, SCD_WAVE));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = x;
    __syncthreads();
    return min(min(s_red[0], s_red[1]), min(s_red[2], s_red[3]));
}

__device__ __forceinline__ uint32_t hb_block_sum(uint32_t x, uint32_t *s_red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += (uint32_t)__shfl_xor(x, o, SCD_WAVE);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = x;
    __syncthreads();
    return s_red[0] + s_red[1] + s_red[2] + s_red[3];
}

__global__ __launch_bounds__(256) void k_hb_eval(MoveArgs a, HubArgs h, HubBuckets b) {
    __shared__ uint32_t keys[HB_CAP];
    __shared__ double s_bi[4];
    __shared__ uint32_t s_bl[4], s_red[4];
    __shared__ double s_rem;
    __shared__ int s_has_rem;
    const int tid = threadIdx.x;
    const uint32_t slot = blockIdx.x;
    const uint32_t hb = b.slot_hub[slot];
    const uint32_t j = slot - b.slot0[hb], K = b.nb[hb];
    const uint32_t R = j < K ? b.cnt[slot] : 0u;
    if (R == 0) {  // uniform: a slot not in use, or an empty bucket
        if (tid == 0) {
            h.chunk_bi[slot] = 0.0;
            h.chunk_bl[slot] = SCD_INVALID;
            h.chunk_has_rem[slot] = 0;
            h.chunk_rem[slot] = 0.0;
        }
        return;
    }
    const uint32_t p = h.hub_v[hb], d = h.hub_d[hb], my = a.labels[p];
    const double pext = vtx_pext(a, p);
    const uint32_t lo = b.lo[slot], hi = j + 1 < K ? b.lo[slot + 1] : SCD_INVALID;
    if (tid == 0) s_has_rem = 0;
    __syncthreads();
    double bi = 0.0;
    uint32_t bl = SCD_INVALID;
    uint32_t heads = 0;
    auto one = [&](uint32_t L, uint32_t c) {  // a candidate, thread 0
        const double imp = candidate_improvement(L, c, d, my, a.stats, pext);
        if (L == my) {
            s_rem = -imp;
            s_has_rem = 1;
        }
        if (imp > 0.0 && better(imp, L, bi, bl)) {
            bi = imp;
            bl = L;
        }
        heads++;
    };
    if (hi - lo == 1u) {
        if (tid == 0) one(lo, R);
    } else if (R <= b.cap) {
        const uint32_t n2 = max(next_pow2(R), 1024u);
        const uint32_t *src = b.region + (size_t)slot * HB_CAP;
        for (uint32_t i = tid; i < n2; i += 256) keys[i] = i < R ? src[i] : SCD_INVALID;
        __syncthreads();
        block_sort_lds<HB_CAP>(keys, n2, R);
        eval_sorted_batched(a, keys, n2, 0, R, d, my, 0, pext, bi, bl, heads, &s_rem, &s_has_rem);
    } else {
        // the region holds cap of the bucket's entries, the overflow area
        // the rest (among other buckets' overflow)
        const uint32_t *src = b.region + (size_t)slot * HB_CAP;
        const uint32_t *ov = b.ovf + h.hub_off[hb];
        const uint32_t no = b.ovf_cnt[hb];
        uint32_t cur = lo;
        while (true) {  // uniform
            uint32_t mn = SCD_INVALID;
            for (uint32_t i = tid; i < b.cap; i += 256) {
                const uint32_t v = src[i];
                if (v >= cur && v < mn) mn = v;
            }
            for (uint32_t i = tid; i < no; i += 256) {
                const uint32_t v = ov[i];
                if (v >= cur && v >= lo && v < hi && v < mn) mn = v;
            }
            mn = hb_block_min(mn, s_red);
            if (mn == SCD_INVALID) break;
            uint32_t c = 0;
            for (uint32_t i = tid; i < b.cap; i += 256) c += src[i] == mn;
            for (uint32_t i = tid; i < no; i += 256) c += ov[i] == mn;
            c = hb_block_sum(c, s_red);
            if (tid == 0) one(mn, c);
            cur = mn + 1u;
        }
    }
    count_pairs(a.counters, heads);
    block_argmax256(bi, bl, s_bi, s_bl);
    if (tid == 0) {
        h.chunk_bi[slot] = bi;
        h.chunk_bl[slot] = bl;
        h.chunk_has_rem[slot] = (uint32_t)s_has_rem;
        h.chunk_rem[slot] = s_has_rem ? s_rem : 0.0;
    }
}

}  // namespace scd_dev


#endif
