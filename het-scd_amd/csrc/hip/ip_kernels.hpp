// Clustering-coefficient-ordered initial partition on the device
// (create_initial_partition, src/scd.cpp:113-171).
//
// The reference sorts (cc = tri/(deg*(deg-1.0)), deg, id) descending and, in
// that order, lets every unlabelled vertex open the next label and claim its
// unlabelled neighbours.  A vertex is unlabelled when its turn comes exactly
// when no earlier vertex adjacent to it opened a label, so the openers
// ("seeds") are the lexicographically-first maximal independent set of the
// order, a seed's label is its rank among the seeds, and every other vertex
// takes the label of its earliest seed neighbour.  The device computes:
//   1. sort keys and a stable LSD radix sort (8-bit digits) of
//      (cc desc, deg desc) over ids listed in descending order, which breaks
//      ties by id descending -> rank[v];
//   2. seed / non-seed decisions in rounds: undecided v becomes a non-seed
//      when an earlier neighbour is a seed and a seed when every earlier
//      neighbour is a non-seed (decisions are final, so a round may read
//      decisions made in the same round);
//   3. seed ranks by an exclusive scan over the order, then labels.
#ifndef SCD_IP_KERNELS_HPP
#define SCD_IP_KERNELS_HPP

#include "dev.hpp"

namespace scd_dev {

constexpr uint32_t RS_TILE = 4096;  // elements per radix-sort block (256 x 16)

// Maps a double to a uint64 whose unsigned order is the double's order
// (non-NaN input).
__device__ __forceinline__ uint64_t orderable(double x) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// Keys for an ascending sort that yields (cc, deg) descending; values = ids
// in descending order so the stable sort breaks ties by id descending.
// flags[0] is set when a degree below 2 makes the reference's key NaN.
// bits[0..3] = AND / OR over all keys of the high and low words: digits
// that are equal in every key need no pass.
// In an arena (gid != nullptr) the graph id is the most significant key
// (kg), so each graph's vertices form one contiguous, correctly ordered run
// and seed ranks are numbered graph by graph.
constexpr uint32_t IP_KEYS_BLOCKS = 2048;  // bounded grid: few same-address atomics on bits

__global__ __launch_bounds__(256) void k_ip_keys(const uint32_t *__restrict__ indices, const int32_t *__restrict__ tri,
                                                 const uint32_t *__restrict__ gid, uint32_t n,
                                                 uint32_t *__restrict__ kg, uint64_t *__restrict__ khi,
                                                 uint32_t *__restrict__ klo, uint32_t *__restrict__ val,
                                                 uint32_t *__restrict__ flags, unsigned long long *__restrict__ bits) {
    __shared__ unsigned long long s_red[4][6];
    uint64_t h_and = ~0ull, h_or = 0, l_and = ~0ull, l_or = 0, g_and = ~0ull, g_or = 0;
    bool low_deg = false;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t p = n - 1 - i;
        const uint32_t deg = indices[p + 1] - indices[p];
        low_deg |= deg < 2;
        const double cc = (double)tri[p] / ((double)deg * ((double)deg - 1.0));
        const uint64_t h = ~orderable(cc);
        const uint32_t l = ~(uint32_t)(int)deg;  // the reference compares deg as int
        const uint32_t g = gid ? gid[p] : 0u;
        khi[i] = h;
        klo[i] = l;
        kg[i] = g;
        val[i] = p;
        h_and &= h;
        h_or |= h;
        l_and &= l;
        l_or |= l;
        g_and &= g;
        g_or |= g;
    }
    if (__ballot(low_deg) && (threadIdx.x & 63) == 0) atomicOr(flags, 1u);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        h_and &= (uint64_t)__shfl_xor((unsigned long long)h_and, off, SCD_WAVE);
        h_or |= (uint64_t)__shfl_xor((unsigned long long)h_or, off, SCD_WAVE);
        l_and &= (uint64_t)__shfl_xor((unsigned long long)l_and, off, SCD_WAVE);
        l_or |= (uint64_t)__shfl_xor((unsigned long long)l_or, off, SCD_WAVE);
        g_and &= (uint64_t)__shfl_xor((unsigned long long)g_and, off, SCD_WAVE);
        g_or |= (uint64_t)__shfl_xor((unsigned long long)g_or, off, SCD_WAVE);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_red[w][0] = h_and;
        s_red[w][1] = h_or;
        s_red[w][2] = l_and;
        s_red[w][3] = l_or;
        s_red[w][4] = g_and;
        s_red[w][5] = g_or;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        unsigned long long x = s_red[0][k];
        for (int q = 1; q < 4; q++) x = (k & 1) ? (x | s_red[q][k]) : (x & s_red[q][k]);
        if (k & 1) atomicOr(&bits[k], x);
        else atomicAnd(&bits[k], x);
    }
}

__global__ void k_ip_bits_init(unsigned long long *bits) {
    if (threadIdx.x < 6) bits[threadIdx.x] = (threadIdx.x & 1) ? 0ull : ~0ull;
}

// passes 0..3: the low word (degree), 4..11: the high word (cc), 12..15: the
// graph id; least significant first
__device__ __forceinline__ uint32_t rs_digit(uint32_t g, uint64_t hi, uint32_t lo, int pass) {
    if (pass < 4) return (lo >> (8 * pass)) & 0xFFu;
    if (pass < 12) return (uint32_t)(hi >> (8 * (pass - 4))) & 0xFFu;
    return (g >> (8 * (pass - 12))) & 0xFFu;
}

// Per-tile digit histogram, digit-major: hist[d * ntiles + tile].
__global__ __launch_bounds__(256) void k_rs_hist(const uint32_t *__restrict__ kg, const uint64_t *__restrict__ khi,
                                                 const uint32_t *__restrict__ klo, uint32_t n, int pass, uint32_t ntiles,
                                                 uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
    for (uint32_t k = 0; k < RS_TILE / 256; k++) {
        const uint64_t i = base + k * 256 + threadIdx.x;
        if (i < n) atomicAdd(&h[rs_digit(kg[i], khi[i], klo[i], pass)], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: elements of a tile are placed in index order within each
// digit.  off[d * ntiles + tile] = exclusive scan of the histogram.
__global__ __launch_bounds__(256) void k_rs_scatter(const uint32_t *__restrict__ kg, const uint64_t *__restrict__ khi,
                                                    const uint32_t *__restrict__ klo, const uint32_t *__restrict__ val,
                                                    uint32_t n, int pass, uint32_t ntiles,
                                                    const uint32_t *__restrict__ off, uint32_t *__restrict__ og,
                                                    uint64_t *__restrict__ ohi, uint32_t *__restrict__ olo,
                                                    uint32_t *__restrict__ oval) {
    __shared__ uint32_t s_base[256];   // next output position per digit
    __shared__ uint32_t s_wc[4][256];  // per-wave digit counts of the current round
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    s_base[threadIdx.x] = off[(uint64_t)threadIdx.x * ntiles + blockIdx.x];
    const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint32_t k = 0; k < RS_TILE / 256; k++) {
        for (int q = 0; q < 4; q++) s_wc[q][threadIdx.x] = 0;
        __syncthreads();
        const uint64_t i = base + k * 256 + threadIdx.x;
        const bool act = i < n;
        uint64_t hi = 0;
        uint32_t g = 0, lo = 0, v = 0, d = 0;
        if (act) {
            g = kg[i];
            hi = khi[i];
            lo = klo[i];
            v = val[i];
            d = rs_digit(g, hi, lo, pass);
        }
        // lanes of this wave holding the same digit
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
            const uint32_t pos = s_base[d] + before + rank_in_wave;
            og[pos] = g;
            ohi[pos] = hi;
            olo[pos] = lo;
            oval[pos] = v;
        }
        __syncthreads();
        s_base[threadIdx.x] += s_wc[0][threadIdx.x] + s_wc[1][threadIdx.x] + s_wc[2][threadIdx.x] +
                               s_wc[3][threadIdx.x];
        __syncthreads();
    }
}

__global__ void k_ip_rank(const uint32_t *__restrict__ val, uint32_t n, uint32_t *__restrict__ rank,
                          uint8_t *__restrict__ state) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rank[val[i]] = i;
    state[val[i]] = 0;  // undecided
}

// Group bits of lane group [gbase, gbase + S) in a wave ballot.
template <int S>
__device__ __forceinline__ uint64_t ip_group_bits(uint64_t ballot, int gbase) {
    if constexpr (S == 64) {
        return ballot;
    } else {
        return (ballot >> gbase) & ((1ull << S) - 1ull);
    }
}

// One decision round over the vertices verts[0..nv) (a degree bin), S
// lanes per vertex walking its neighbours S at a time.  state: 0
// undecided, 1 seed, 2 non-seed.  An undecided vertex becomes a non-seed
// when an earlier-ranked neighbour is a seed, a seed when no earlier-ranked
// neighbour is seed or undecided; otherwise it stays undecided and sets
// *open_flag.  Decisions are final, so reading states that other vertices
// set in the same round only speeds the rounds up.  An undecided vertex
// resumes its walk at resume[p], the first neighbour that was still an
// undecided earlier one in its last round: every neighbour before it was
// later in the order or a non-seed, which stays so.
template <int S>
__global__ __launch_bounds__(256) void k_ip_round(const uint32_t *__restrict__ indices, const uint32_t *__restrict__ adj,
                                                  const uint32_t *__restrict__ rank, uint8_t *state,
                                                  const uint32_t *__restrict__ verts, uint32_t nv,
                                                  uint32_t *__restrict__ open_flag, uint32_t *__restrict__ resume) {
    const int lane = threadIdx.x & 63;
    const int lig = lane & (S - 1), gbase = lane & ~(S - 1);
    const uint32_t slot = (blockIdx.x * 256u + threadIdx.x) / S;
    uint32_t p = 0, b = 0, e = 0, rp = 0;
    bool act = slot < nv;
    if (act) {
        p = verts[slot];
        act = __atomic_load_n(&state[p], __ATOMIC_RELAXED) == 0;
    }
    uint32_t s0 = 0;
    if (act) {
        b = indices[p];
        e = indices[p + 1];
        rp = rank[p];
        s0 = resume[p];
    }
    bool pending = false, seed_before = false;
    uint32_t first_open = 0;
    for (uint32_t s = s0;; s += S) {
        const bool more = act && !seed_before && b + s < e;
        if (!__ballot(more)) break;  // wave-uniform exit
        bool es = false, eo = false;
        if (more) {
            const uint32_t i = b + s + lig;
            if (i < e) {
                const uint32_t q = adj[i];
                if (rank[q] < rp) {
                    const uint8_t sq = __atomic_load_n(&state[q], __ATOMIC_RELAXED);
                    es = sq == 1;
                    eo = sq == 0;
                }
            }
        }
        const uint64_t bs = ip_group_bits<S>(__ballot(es), gbase), bo = ip_group_bits<S>(__ballot(eo), gbase);
        if (more) {
            if (bs) {
                seed_before = true;
            } else if (bo) {
                if (!pending) first_open = s + (uint32_t)__builtin_ctzll(bo);
                pending = true;
            }
        }
    }
    const bool still_open = act && !seed_before && pending;
    if (act && lig == 0) {
        if (seed_before) __atomic_store_n(&state[p], (uint8_t)2, __ATOMIC_RELAXED);
        else if (!pending) __atomic_store_n(&state[p], (uint8_t)1, __ATOMIC_RELAXED);
        else resume[p] = first_open;
    }
    if (__ballot(still_open) && lane == 0) open_flag[0] = 1u;
}

// The decision round of k_ip_round for vertices of 33 to wave_max
// neighbours: a wave takes 64 entries of the list, checks their states
// together (late rounds find most of them decided) and walks the undecided
// ones one after another with all of its lanes.
__global__ __launch_bounds__(256) void k_ip_round_wave(const uint32_t *__restrict__ indices,
                                                       const uint32_t *__restrict__ adj,
                                                       const uint32_t *__restrict__ rank, uint8_t *state,
                                                       const uint32_t *__restrict__ verts, uint32_t nv,
                                                       uint32_t *__restrict__ open_flag,
                                                       uint32_t *__restrict__ resume) {
    const int lane = threadIdx.x & 63;
    const uint32_t slot = (blockIdx.x * 4u + (threadIdx.x >> 6)) * 64u + lane;
    uint32_t mine = 0;
    bool und = false;
    if (slot < nv) {
        mine = verts[slot];
        und = __atomic_load_n(&state[mine], __ATOMIC_RELAXED) == 0;
    }
    uint64_t todo = __ballot(und);
    bool any_open = false;
    while (todo) {  // wave-uniform
        const int l = __builtin_ctzll(todo);
        todo &= todo - 1;
        const uint32_t p = __shfl(mine, l, SCD_WAVE);
        const uint32_t b = indices[p], e = indices[p + 1], rp = rank[p];
        bool pending = false, seed_before = false;
        uint32_t first_open = 0;
        for (uint32_t s = resume[p]; b + s < e; s += 64) {  // wave-uniform
            const uint32_t i = b + s + lane;
            bool es = false, eo = false;
            if (i < e) {
                const uint32_t q = adj[i];
                if (rank[q] < rp) {
                    const uint8_t sq = __atomic_load_n(&state[q], __ATOMIC_RELAXED);
                    es = sq == 1;
                    eo = sq == 0;
                }
            }
            const uint64_t bs = __ballot(es), bo = __ballot(eo);
            if (bs) {
                seed_before = true;
                break;
            }
            if (bo && !pending) {
                first_open = s + (uint32_t)__builtin_ctzll(bo);
                pending = true;
            }
        }
        if (lane == 0) {
            if (seed_before) __atomic_store_n(&state[p], (uint8_t)2, __ATOMIC_RELAXED);
            else if (!pending) __atomic_store_n(&state[p], (uint8_t)1, __ATOMIC_RELAXED);
            else resume[p] = first_open;
        }
        any_open |= !seed_before && pending;
    }
    if (any_open && lane == 0) open_flag[0] = 1u;
}

// The decision round of k_ip_round for vertices of more than wave_max
// neighbours: one workgroup per vertex, 256 neighbours at a time, the
// chunk's seed / undecided findings combined through LDS.
__global__ __launch_bounds__(256) void k_ip_round_block(const uint32_t *__restrict__ indices,
                                                        const uint32_t *__restrict__ adj,
                                                        const uint32_t *__restrict__ rank, uint8_t *state,
                                                        const uint32_t *__restrict__ verts, uint32_t nv,
                                                        uint32_t *__restrict__ open_flag,
                                                        uint32_t *__restrict__ resume) {
    __shared__ uint32_t s_any_seed[4], s_first[4];
    __shared__ uint32_t s_open;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t p = verts[blockIdx.x];
    if (tid == 0) s_open = __atomic_load_n(&state[p], __ATOMIC_RELAXED) == 0;
    __syncthreads();
    if (!s_open) return;  // uniform: decided vertices are done
    const uint32_t b = indices[p], e = indices[p + 1], rp = rank[p];
    bool seed_before = false, pending = false;
    uint32_t first_open = 0;
    for (uint32_t s = resume[p]; b + s < e; s += 256) {  // uniform bounds
        const uint32_t i = b + s + tid;
        bool es = false, eo = false;
        if (i < e) {
            const uint32_t q = adj[i];
            if (rank[q] < rp) {
                const uint8_t sq = __atomic_load_n(&state[q], __ATOMIC_RELAXED);
                es = sq == 1;
                eo = sq == 0;
            }
        }
        const uint64_t bs = __ballot(es), bo = __ballot(eo);
        if (lane == 0) {
            s_any_seed[w] = bs != 0;
            s_first[w] = bo ? 64u * w + (uint32_t)__builtin_ctzll(bo) : 0xFFFFFFFFu;
        }
        __syncthreads();
        const bool any_seed = s_any_seed[0] | s_any_seed[1] | s_any_seed[2] | s_any_seed[3];
        const uint32_t fo = min(min(s_first[0], s_first[1]), min(s_first[2], s_first[3]));
        __syncthreads();
        if (any_seed) {
            seed_before = true;
            break;
        }
        if (fo != 0xFFFFFFFFu && !pending) {
            first_open = s + fo;
            pending = true;
        }
    }
    if (tid == 0) {
        if (seed_before) __atomic_store_n(&state[p], (uint8_t)2, __ATOMIC_RELAXED);
        else if (!pending) __atomic_store_n(&state[p], (uint8_t)1, __ATOMIC_RELAXED);
        else {
            resume[p] = first_open;
            open_flag[0] = 1u;
        }
    }
}

// sf[rank] = 1 for seeds (by position in the order).
__global__ void k_ip_seedflags(const uint32_t *__restrict__ val, const uint8_t *__restrict__ state, uint32_t n,
                               uint32_t *__restrict__ sf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sf[i] = state[val[i]] == 1 ? 1u : 0u;
}

// Arenas: k_g[g] = seeds of graph g.  The graph id leads the sort key, so
// graph g holds the positions [vtx_off[g], vtx_off[g+1]) of the order and
// its seed count is the difference of the exclusive seed scan (seedpos,
// total seeds `total`) at the two ends.
__global__ void k_ip_graph_seeds(const uint32_t *__restrict__ seedpos, const uint64_t *__restrict__ vtx_off,
                                 uint32_t G, uint32_t n, uint32_t total, uint32_t *__restrict__ k_g) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    const uint64_t a = vtx_off[g], b = vtx_off[g + 1];
    const uint32_t sa = a < n ? seedpos[a] : total, sb = b < n ? seedpos[b] : total;
    k_g[g] = sb - sa;
}

// Labels: a seed's label is its seed rank (seedpos = exclusive scan of sf);
// any other vertex takes the label of its earliest seed neighbour.  S lanes
// per vertex of one degree bin.
template <int S>
__global__ __launch_bounds__(256) void k_ip_labels(const uint32_t *__restrict__ indices, const uint32_t *__restrict__ adj,
                                                   const uint32_t *__restrict__ rank, const uint8_t *__restrict__ state,
                                                   const uint32_t *__restrict__ seedpos,
                                                   const uint32_t *__restrict__ verts, uint32_t nv,
                                                   uint32_t *__restrict__ labels) {
    const int lane = threadIdx.x & 63;
    const int lig = lane & (S - 1);
    const uint32_t slot = (blockIdx.x * 256u + threadIdx.x) / S;
    const bool act = slot < nv;
    const uint32_t p = act ? verts[slot] : 0u;
    const bool seed = act && state[p] == 1;
    uint32_t best = 0xFFFFFFFFu;
    if (act && !seed) {
        const uint32_t b = indices[p], e = indices[p + 1];
        for (uint32_t i = b + lig; i < e; i += S) {
            const uint32_t q = adj[i];
            if (state[q] == 1) best = min(best, rank[q]);
        }
    }
#pragma unroll
    for (int off = S / 2; off > 0; off >>= 1) best = min(best, (uint32_t)__shfl_xor(best, off, SCD_WAVE));
    if (act && lig == 0) labels[p] = seed ? seedpos[rank[p]] : seedpos[best];
}

// The labels of k_ip_labels for vertices of more than wave_max neighbours:
// one workgroup per vertex, the earliest seed neighbour reduced through LDS.
__global__ __launch_bounds__(256) void k_ip_labels_block(const uint32_t *__restrict__ indices,
                                                         const uint32_t *__restrict__ adj,
                                                         const uint32_t *__restrict__ rank,
                                                         const uint8_t *__restrict__ state,
                                                         const uint32_t *__restrict__ seedpos,
                                                         const uint32_t *__restrict__ verts, uint32_t nv,
                                                         uint32_t *__restrict__ labels) {
    __shared__ uint32_t s_min[4];
    const int tid = threadIdx.x, w = tid >> 6;
    const uint32_t p = verts[blockIdx.x];
    const bool seed = state[p] == 1;
    uint32_t best = 0xFFFFFFFFu;
    if (!seed) {
        const uint32_t b = indices[p], e = indices[p + 1];
        for (uint32_t i = b + tid; i < e; i += 256) {
            const uint32_t q = adj[i];
            if (state[q] == 1) best = min(best, rank[q]);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) best = min(best, (uint32_t)__shfl_xor(best, off, SCD_WAVE));
    if ((tid & 63) == 0) s_min[w] = best;
    __syncthreads();
    if (tid == 0) {
        best = min(min(s_min[0], s_min[1]), min(s_min[2], s_min[3]));
        labels[p] = seed ? seedpos[rank[p]] : seedpos[best];
    }
}

}  // namespace scd_dev

#endif
