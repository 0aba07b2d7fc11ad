// WCC evaluation (compute_wcc, src/wcc.cpp:31-141) as triangle counting on
// the internal subgraph.
//
// For p with valid label L, sizes[L] > 2 and tri_deg[p] != 0 the reference
// sums, over same-label neighbours q, |Nc(p) ∩ N(q)|, where Nc(p) is p's
// same-label neighbour list (int_tri), and counts the q with a non-empty
// intersection (int_tri_deg).  Every common neighbour r in Nc(p) ∩ N(q) has
// label L, so each such (q, r) is a triangle p-q-r inside community L, seen
// once from q and once from r:
//   int_tri(p)     = 2 x (internal triangles at p)
//   int_tri_deg(p) = internal edges at p that lie in an internal triangle.
// The internal lists (int_adj, sorted) are oriented from the higher to the
// lower id, each internal triangle w < v < u is found once at u, per-vertex
// counts aggregate in LDS over N+_int(u), and the edges of a found triangle
// are flagged with plain stores of 1 (the filter pass cleared the flags of
// the internal entries).  All grids are fixed by the degree bins, so an
// iteration stays capturable.  The same kernels over the whole adjacency
// (int_deg == nullptr) give the graph's t / vt (remove_edges_without_
// triangles, src/convert.cpp:271-351): tri = 2 T, tri_deg = td.
#ifndef SCD_WCC_KERNELS_HPP
#define SCD_WCC_KERNELS_HPP

#include "dev.hpp"
#include "stats_kernels.hpp"

namespace scd_dev {

struct WccArgs {
    const uint32_t *__restrict__ indices;
    const uint32_t *__restrict__ labels;
    const int4 *__restrict__ stats;
    const uint32_t *__restrict__ int_adj;
    const uint32_t *__restrict__ int_deg;
    const int32_t *__restrict__ tri;
    const int32_t *__restrict__ tri_deg;
    uint32_t *__restrict__ dplus;  // internal neighbours below each vertex
    uint32_t *__restrict__ T;      // internal triangles per vertex
    uint32_t *__restrict__ td;     // internal edges in an internal triangle
    uint8_t *__restrict__ flag;    // per int_adj entry (out-slots used)
    double *__restrict__ terms;
    int32_t *__restrict__ int_tri_out;      // optional
    int32_t *__restrict__ int_tri_deg_out;  // optional
    // optional, per label: 1 if the community's member set changed since
    // the T / td of its members were counted.  Members of the other
    // communities keep T and td (the internal subgraph of a community is
    // a function of its member set).  nullptr: count everything.
    const uint8_t *__restrict__ dirty;
    // Counts for the hot_n lowest ids (the highest-degree vertices of a
    // convert-relabelled graph, which collect the most updates) go to
    // WTC_HOT_COPIES private copies, one chosen per workgroup, and are summed
    // into T / td afterwards: fewer atomics on one address.
    uint32_t *__restrict__ T_hot;   // WTC_HOT_COPIES x hot_n
    uint32_t *__restrict__ td_hot;  // WTC_HOT_COPIES x hot_n
    uint32_t hot_n;
    // Out-lists longer than vc_min (and at most WTC_VC_CAP) are probed from
    // their own side: an edge u -> v whose v has the longer out-list is
    // walked as N+(u) below v against the set of N+(v) (k_wtc_queue<.., true>)
    // instead of N+(v) against N+(u).  0: every edge from u's side.
    uint32_t vc_min;
    // optional, per label: the row offset of a community counted by the
    // dense (bitmap) kernels below, DN_NONE for the others.  Members of a
    // dense community skip the list kernels (their out-lists count as empty).
    const uint32_t *__restrict__ dn_off = nullptr;
    // optional: the two counters of k_wtc_pack's lists, zeroed by
    // k_wtc_dplus
    uint32_t *__restrict__ pack_cnt = nullptr;
};

constexpr uint32_t WTC_HOT_MAX = 8192;
constexpr uint32_t WTC_HOT_COPIES = 16;
// the hottest ids (the largest hubs collect tens of thousands of updates
// each) get more copies, after the WTC_HOT_COPIES x hot_n block
constexpr uint32_t WTC_HOT1_N = 256;
constexpr uint32_t WTC_HOT1_COPIES = 64;
constexpr uint32_t WTC_HOT_WORDS = WTC_HOT_COPIES * WTC_HOT_MAX + WTC_HOT1_COPIES * WTC_HOT1_N;
constexpr uint32_t WTC_VC_CAP = 2048;  // the work queue's LDS set capacity

// Edge u -> v (v < u, out-list lengths dpu, dpv) is counted from v's side.
__device__ __forceinline__ bool wtc_from_v(const WccArgs &a, uint32_t dpv, uint32_t dpu) {
    return a.vc_min && dpv > a.vc_min && dpv <= WTC_VC_CAP && dpv > dpu;
}

__device__ __forceinline__ uint32_t wtc_hot_slot(const WccArgs &a, uint32_t x) {
    return x < WTC_HOT1_N
               ? WTC_HOT_COPIES * a.hot_n + (blockIdx.x & (WTC_HOT1_COPIES - 1)) * WTC_HOT1_N + x
               : (blockIdx.x & (WTC_HOT_COPIES - 1)) * a.hot_n + x;
}

__device__ __forceinline__ void wtc_add_T(const WccArgs &a, uint32_t x, uint32_t c) {
    if (x < a.hot_n) atomicAdd(&a.T_hot[wtc_hot_slot(a, x)], c);
    else atomicAdd(&a.T[x], c);
}

__device__ __forceinline__ void wtc_add_td(const WccArgs &a, uint32_t x, uint32_t c) {
    if (x < a.hot_n) atomicAdd(&a.td_hot[wtc_hot_slot(a, x)], c);
    else atomicAdd(&a.td[x], c);
}

// T[x] += sum of the copies, td likewise, copies cleared for the next pass.
// Folds the private copies of vertex x (< hot_n) into T[x] / td[x] and
// clears them.
__device__ __forceinline__ void wtc_hot_fold(const WccArgs &a, uint32_t x) {
    uint32_t t = 0, d = 0;
#pragma unroll
    for (uint32_t k = 0; k < WTC_HOT_COPIES; k++) {
        t += a.T_hot[k * a.hot_n + x];
        d += a.td_hot[k * a.hot_n + x];
        a.T_hot[k * a.hot_n + x] = 0;
        a.td_hot[k * a.hot_n + x] = 0;
    }
    if (x < WTC_HOT1_N) {
        const uint32_t b = WTC_HOT_COPIES * a.hot_n + x;
        for (uint32_t k = 0; k < WTC_HOT1_COPIES; k++) {
            t += a.T_hot[b + k * WTC_HOT1_N];
            d += a.td_hot[b + k * WTC_HOT1_N];
            a.T_hot[b + k * WTC_HOT1_N] = 0;
            a.td_hot[b + k * WTC_HOT1_N] = 0;
        }
    }
    if (t) a.T[x] += t;
    if (d) a.td[x] += d;
}

__global__ void k_wtc_hot_sum(WccArgs a) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < a.hot_n) wtc_hot_fold(a, x);
}

// dirty[old label] = dirty[new label] = 1 for every vertex that moved
// (the flags were cleared before).
__global__ void k_mark_dirty(const uint32_t *__restrict__ old_labels, const uint32_t *__restrict__ new_labels,
                             uint32_t n, uint8_t *__restrict__ dirty) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const uint32_t a = old_labels[u], b = new_labels[u];
    if (a == b) return;
    if (a != SCD_INVALID) dirty[a] = 1;
    if (b != SCD_INVALID) dirty[b] = 1;
}


__global__ void k_wtc_dplus(WccArgs a, uint32_t n) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u == 0 && a.pack_cnt) a.pack_cnt[0] = a.pack_cnt[1] = 0;
    if (u >= n) return;
    if (a.dirty) {
        const uint32_t lab = a.labels[u];
        if (lab != SCD_INVALID && !a.dirty[lab]) {
            a.dplus[u] = 0;  // community unchanged: no counting, T and td kept
            return;
        }
    }
    if (a.dn_off) {
        const uint32_t lab = a.labels[u];
        if (lab != SCD_INVALID && a.dn_off[lab] != 0xFFFFFFFFu) {
            a.dplus[u] = 0;  // T and td come from the dense kernels
            return;
        }
    }
    const uint32_t *L = a.int_adj + a.indices[u];
    // int_deg == nullptr: the whole adjacency (triangle counting of the graph)
    uint32_t lo = 0, hi = a.int_deg ? a.int_deg[u] : a.indices[u + 1] - a.indices[u];
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (L[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    a.dplus[u] = lo;
    a.T[u] = 0;
    a.td[u] = 0;
}

// Out-lists of at most S entries (S = 8, 16, 32, 64): S lanes per vertex.
// The member lists N+_int(v), v in N+_int(u), are walked flattened over the
// group's lanes (WTC_UF loads in flight per lane; the owning v of an entry
// by binary search over the group's prefix of list lengths); membership in
// N+_int(u) by binary search in LDS.  Counts per out-list position collect
// in LDS and flush once per vertex.  Longer out-lists (S = 64 groups skip
// them) go to the work queue below.
constexpr int WTC_UF = 4;                // loads in flight per lane, flattened walks
constexpr uint32_t WTC_GROUP_MAX = 64;   // longer out-lists go to the work queue

// nv_dev (optional): the list's length on the device (a k_wtc_pack list);
// the grid then strides over it.
template <int S>
__global__ __launch_bounds__(256) void k_wtc_group(WccArgs a, const uint32_t *__restrict__ verts, uint32_t nv,
                                                   const uint32_t *__restrict__ nv_dev) {
    __shared__ uint32_t s_n[256], s_vb[256], s_cnt[256], s_inc[256];
    const int tid = threadIdx.x, lane = tid & 63;
    const int lig = lane & (S - 1), gb = tid & ~(S - 1);  // group base in the block's arrays
    const uint32_t count = nv_dev ? *nv_dev : nv;
    for (uint32_t b0 = blockIdx.x * (256u / S); b0 < count; b0 += gridDim.x * (256u / S)) {  // uniform
        const uint32_t slot = b0 + (uint32_t)tid / S;
        uint32_t u = 0, beg = 0, dp = 0;
        if (slot < count) {
            u = verts[slot];
            dp = a.dplus[u];
            if (dp > (uint32_t)S) dp = 0;  // work queue
            if (dp) beg = a.indices[u];
        }
        if (!__syncthreads_or(dp != 0)) continue;  // nothing to count in this pass (uniform)
        uint32_t x = SCD_INVALID, len = 0, vb = 0;
        if ((uint32_t)lig < dp) {
            x = a.int_adj[beg + lig];
            len = a.dplus[x];
            if (wtc_from_v(a, len, dp)) len = 0;  // counted from x's side
            vb = a.indices[x];
        }
        uint32_t inc = len;  // inclusive prefix over the group
#pragma unroll
        for (int o = 1; o < S; o <<= 1) {
            const uint32_t t = __shfl_up(inc, o, SCD_WAVE);
            if (lig >= o) inc += t;
        }
        s_n[tid] = x;
        s_vb[tid] = vb;
        s_cnt[tid] = 0;
        s_inc[tid] = inc;
        __syncthreads();
        const uint32_t W = s_inc[gb + S - 1];
        uint32_t tot = 0;
        uint32_t cur = 0;  // owner of the lane's entries: the lane's f only grows
        for (uint32_t f0 = lig; f0 < W; f0 += S * WTC_UF) {
            uint32_t w[WTC_UF], at[WTC_UF], own[WTC_UF];
#pragma unroll
            for (int q = 0; q < WTC_UF; q++) {
                const uint32_t f = f0 + S * q;
                w[q] = SCD_INVALID;
                if (f < W) {
                    while (s_inc[gb + cur] <= f) cur++;  // smallest k with inc[k] > f
                    own[q] = cur;
                    const uint32_t excl = cur ? s_inc[gb + cur - 1] : 0u;
                    at[q] = s_vb[gb + cur] + (f - excl);
                    w[q] = a.int_adj[at[q]];
                }
            }
#pragma unroll
            for (int q = 0; q < WTC_UF; q++) {
                if (w[q] == SCD_INVALID) continue;
                uint32_t lo = 0, hi = dp;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_n[gb + mid] < w[q]) lo = mid + 1;
                    else hi = mid;
                }
                if (lo < dp && s_n[gb + lo] == w[q]) {
                    atomicAdd(&s_cnt[gb + own[q]], 1u);
                    atomicAdd(&s_cnt[gb + lo], 1u);
                    a.flag[at[q]] = 1;
                    tot++;
                }
            }
        }
        __syncthreads();
        if ((uint32_t)lig < dp && s_cnt[tid]) {
            wtc_add_T(a, x, s_cnt[tid]);
            a.flag[beg + lig] = 1;
        }
#pragma unroll
        for (int o = S / 2; o > 0; o >>= 1) tot += __shfl_xor(tot, o, SCD_WAVE);
        if (lig == 0 && tot) atomicAdd(&a.T[u], tot);
        __syncthreads();  // the LDS arrays are refilled by the next pass
    }
}

// LDS hash set of N+_int(u): slot -> (vertex, position in the out-list).
__device__ __forceinline__ uint32_t wtc_hash(uint32_t x, int bits) { return (x * 0x9E3779B1u) >> (32 - bits); }

__device__ __forceinline__ int wtc_find(const uint32_t *hk, const uint16_t *hp, int bits, uint32_t w) {
    const uint32_t mask = (1u << bits) - 1u;
    for (uint32_t h = wtc_hash(w, bits);; h = (h + 1) & mask) {
        const uint32_t k = hk[h];
        if (k == w) return hp[h];
        if (k == SCD_INVALID) return -1;
    }
}

// Out-lists longer than 64 entries are processed as work items (u, c): chunk c holds
// entries [c VC, (c+1) VC) of N+_int(u).  One persistent kernel claims items
// from an atomic counter until the device-computed item total is reached, so
// every block exits and the grid is fixed (the iteration stays capturable).
// Items are numbered from the highest degree bin down, so the longest
// out-lists start first.  An item rebuilds the LDS hash set of the whole
// N+_int(u), walks the member lists of its chunk — one list per wave when
// they average >= 32 entries, else flattened over the threads — counts per
// out-list position in LDS and flushes one atomic per non-zero position.
constexpr uint32_t WTC_VC = 64;     // out-list entries per work item
constexpr uint32_t WTC_VCV = 128;  // upper-list entries per work item counted from the owner's side
constexpr int WTC_QT = 512;  // work-queue workgroup: 8 waves, 64 VGPRs, 4 workgroups per CU
constexpr uint32_t WTC_CLAIM = 8;  // items claimed at once by a workgroup

// nch[i] = chunks of verts[M-1-i]; also resets the item counter.  From u's
// side (VC false): chunks of the out-list when it is longer than 64 entries;
// from v's side (VC true): chunks of the upper list N-(v) (entries past the
// out-list) of every vertex whose out-list takes edges from its side.
// The vertices of a list whose internal out-list is not empty, by its
// length: up to WTC_PACK_LO entries to lo, more to hi, with the counts in
// a.pack_cnt[0] and [1].  A workgroup of 1024 threads ranks its vertices by
// wave ballots and reserves its places with one atomic per list (device-
// scope atomics on one address cost several ns each on this 8-XCD part).
// Most vertices of degree above 32 have no or a few internal
// out-neighbours, so the triangle and final-count kernels walk these lists
// (8 lanes per vertex for lo) instead of a wave per listed vertex.  Order
// within a list varies between runs; the counts are integer sums.
constexpr uint32_t WTC_PACK_LO = 8;
constexpr int WTC_PACK_NT = 1024;

__global__ __launch_bounds__(WTC_PACK_NT) void k_wtc_pack(WccArgs a, const uint32_t *__restrict__ verts,
                                                          uint32_t M, uint32_t *__restrict__ lo,
                                                          uint32_t *__restrict__ hi) {
    constexpr int NW = WTC_PACK_NT / 64;
    __shared__ uint32_t s_off[2][NW];  // per wave: rank of its first entry, then its list position
    __shared__ uint32_t s_base[2];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t i = blockIdx.x * (uint32_t)WTC_PACK_NT + threadIdx.x;
    uint32_t v = 0, dp = 0;
    if (i < M) {
        v = verts[i];
        dp = a.dplus[v];
    }
    const bool to_lo = dp != 0 && dp <= WTC_PACK_LO, to_hi = dp > WTC_PACK_LO;
    const uint64_t m_lo = __ballot(to_lo), m_hi = __ballot(to_hi);
    if (lane == 0) {
        s_off[0][w] = (uint32_t)__builtin_popcountll(m_lo);
        s_off[1][w] = (uint32_t)__builtin_popcountll(m_hi);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int k = threadIdx.x;
        uint32_t run = 0;
        for (int x = 0; x < NW; x++) {
            const uint32_t c = s_off[k][x];
            s_off[k][x] = run;
            run += c;
        }
        s_base[k] = run ? atomicAdd(&a.pack_cnt[k], run) : 0u;
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    if (to_lo) lo[s_base[0] + s_off[0][w] + __builtin_popcountll(m_lo & below)] = v;
    if (to_hi) hi[s_base[1] + s_off[1][w] + __builtin_popcountll(m_hi & below)] = v;
}

template <bool VC>
__global__ void k_wtc_count_chunks(WccArgs a, const uint32_t *__restrict__ verts, uint32_t M,
                                   uint32_t *__restrict__ nch, uint32_t *__restrict__ next) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) next[0] = 0;
    if (i >= M) return;
    const uint32_t v = verts[M - 1 - i];
    const uint32_t dp = a.dplus[v];
    if (!VC) {
        nch[i] = dp > WTC_GROUP_MAX ? (dp + WTC_VC - 1) / WTC_VC : 0u;
        return;
    }
    uint32_t up = 0;
    if (a.vc_min && dp > a.vc_min && dp <= WTC_VC_CAP) {
        const uint32_t deg = a.int_deg ? a.int_deg[v] : a.indices[v + 1] - a.indices[v];
        up = deg - dp;
    }
    nch[i] = (up + WTC_VCV - 1) / WTC_VCV;
}

// Membership set of N+_int(u) (sorted, dp entries) for one work item.  When
// the id range R of the list fits WTC_TAB words, a bitmap over [lo, lo+R)
// whose words carry the out-list position of their first entry: w is a
// member iff its bit is set, and its position is that rank plus the bits
// below it in the word — one LDS read per probe.  Wider ranges use the
// linear-probing hash (slot -> vertex, position).
constexpr uint32_t WTC_TAB = 3800;  // (bits, rank) words: id ranges up to 121600 (4 blocks per CU)

struct WtcSet {
    uint2 *tab;
    uint32_t *hk;
    uint16_t *hp;
    uint32_t lo, R;
    int bits;  // 0: bitmap mode
    __device__ __forceinline__ int find(uint32_t w) const {
        if (bits == 0) {
            const uint32_t x = w - lo;
            if (x >= R) return -1;
            const uint2 e = tab[x >> 5];
            const uint32_t b = x & 31u;
            if (!((e.x >> b) & 1u)) return -1;
            return (int)(e.y + (uint32_t)__builtin_popcount(e.x & ((1u << b) - 1u)));
        }
        return wtc_find(hk, hp, bits, w);
    }
};

// VC false: items are WTC_VC-entry chunks of N+_int(u), each entry v's list
// N+_int(v) walked against the set of N+_int(u).  VC true: items are
// WTC_VCV-entry chunks of the upper list N-_int(v) of a vertex v whose edges
// are counted from its side; each entry u's out-list is walked against the
// set of N+_int(v), the entries below v probed (the walk's length needs no
// search up front; entries past v are read and skipped), and u's count and
// the flag of u -> v collect per chunk entry.
template <int CAP, int NT, int UW, bool VC>
__global__ __launch_bounds__(NT, NT >= 512 ? 8 : 1) void k_wtc_queue(WccArgs a, const uint32_t *__restrict__ verts, uint32_t M,
                                                   const uint32_t *__restrict__ off,
                                                   const uint32_t *__restrict__ total, uint32_t *__restrict__ next) {
    static_assert(2 * CAP * 4 + 2 * CAP * 2 <= WTC_TAB * 8, "hash must fit in the bitmap storage");
    static_assert(!VC || CAP == WTC_VC_CAP, "v-side sets hold up to WTC_VC_CAP entries");
    __shared__ uint2 s_tab[WTC_TAB];  // bitmap words, or the hash (keys, positions)
    __shared__ uint32_t s_cnt[CAP];
    constexpr uint32_t CH = VC ? WTC_VCV : WTC_VC;  // entries per chunk
    __shared__ uint32_t s_vb[CH], s_pre[CH + 1];
    __shared__ uint32_t s_ucnt[VC ? CH : 1], s_upos[VC ? CH : 1];
    __shared__ uint32_t s_item;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    constexpr int NW = NT / 64;
    WtcSet set;
    set.tab = s_tab;
    set.hk = reinterpret_cast<uint32_t *>(s_tab);
    set.hp = reinterpret_cast<uint16_t *>(set.hk + 2 * CAP);
    // the bitmap is all-zero between vertices (each clears what it set)
    for (uint32_t k = tid; k < WTC_TAB; k += NT) s_tab[k] = make_uint2(0u, 0u);
    const uint32_t items = *total;
    while (true) {
        // claim WTC_CLAIM consecutive items: chunks of one vertex in a claim
        // share one membership set and one flush of the position counts
        if (tid == 0) s_item = atomicAdd(next, WTC_CLAIM);
        __syncthreads();
        uint32_t item = s_item;
        if (item >= items) break;  // uniform: every block ends here
        const uint32_t item_end = min(item + WTC_CLAIM, items);
        while (item < item_end) {  // uniform
            uint32_t lo = 0, hi = M;  // largest i with off[i] <= item
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (off[mid] <= item) lo = mid;
                else hi = mid;
            }
            const uint32_t u = verts[M - 1 - lo];
            const uint32_t beg = a.indices[u], dp = a.dplus[u];
            const uint32_t vend = lo + 1 < M ? off[lo + 1] : items;
            const uint32_t c_end = min(vend, item_end);
            const uint32_t c0 = item - off[lo], c1 = c_end - off[lo];
            const uint32_t *N = a.int_adj + beg;
            uint32_t tot = 0;
            if (dp <= (uint32_t)CAP) {
                set.lo = N[0];
                set.R = N[dp - 1] - set.lo + 1;
                set.bits = 0;
                if (set.R > 32u * WTC_TAB) {
                    set.bits = 1;
                    while ((1u << set.bits) < 2 * dp) set.bits++;
                    for (uint32_t h = tid; h < (1u << set.bits); h += NT) set.hk[h] = SCD_INVALID;
                }
                for (uint32_t k = tid; k < dp; k += NT) s_cnt[k] = 0;
                __syncthreads();
                if (set.bits == 0) {
                    for (uint32_t k = tid; k < dp; k += NT) {
                        const uint32_t x = N[k] - set.lo;
                        atomicOr(&s_tab[x >> 5].x, 1u << (x & 31u));
                        if (k == 0 || ((N[k - 1] - set.lo) >> 5) != (x >> 5)) s_tab[x >> 5].y = k;
                    }
                } else {
                    const uint32_t mask = (1u << set.bits) - 1u;
                    for (uint32_t k = tid; k < dp; k += NT) {
                        const uint32_t v = N[k];
                        uint32_t h = wtc_hash(v, set.bits);
                        while (atomicCAS(&set.hk[h], SCD_INVALID, v) != SCD_INVALID) h = (h + 1) & mask;
                        set.hp[h] = (uint16_t)k;
                    }
                }
                // chunks of the owner u's out-list (VC false) or of its upper
                // list (VC true); per chunk entry x: the list walked is x's
                // out-list (all of it when VC: entries below u are probed,
                // u's own position is recorded, entries above u are skipped)
                const uint32_t lst = VC ? (a.int_deg ? a.int_deg[u] : a.indices[u + 1] - beg) : dp;
                const uint32_t base = VC ? dp : 0u;
                for (uint32_t c = c0; c < c1; c++) {
                    const uint32_t k0 = base + c * CH, nk = min(CH, lst - k0);
                    if (tid < (int)nk) {
                        const uint32_t v = N[k0 + tid];
                        const uint32_t dv = a.dplus[v];
                        s_vb[tid] = a.indices[v];
                        if (VC) {
                            s_pre[tid + 1] = wtc_from_v(a, dp, dv) ? dv : 0u;  // else counted from v's side
                            s_ucnt[tid] = 0;
                            s_upos[tid] = SCD_INVALID;
                        } else {
                            s_pre[tid + 1] = wtc_from_v(a, dv, dp) ? 0u : dv;  // else counted from v's side
                        }
                    }
                    if (tid == 0) s_pre[0] = 0;
                    __syncthreads();
                    if (wv < (int)((CH + 63) / 64)) {  // prefix of the chunk's list lengths, 64 per wave
                        const uint32_t e = 64u * wv + lane;
                        uint32_t x = e < nk ? s_pre[e + 1] : 0u;
#pragma unroll
                        for (int o = 1; o < 64; o <<= 1) {
                            const uint32_t t = __shfl_up(x, o, SCD_WAVE);
                            if (lane >= o) x += t;
                        }
                        if (e < nk) s_pre[e + 1] = x;
                    }
                    __syncthreads();
                    if (CH > 64) {  // carry the first waves' totals
                        for (uint32_t g = 1; g < CH / 64; g++) {
                            const uint32_t e = 64u * g + tid;
                            const uint32_t carry = s_pre[64u * g];
                            __syncthreads();
                            if (tid < 64 && e < nk) s_pre[e + 1] += carry;
                            __syncthreads();
                        }
                    }
                    const uint32_t Wc = s_pre[nk];
                    if (Wc >= 32u * nk) {
                        // one list per wave; UW loads of 64 entries in flight
                        for (uint32_t kk = wv; kk < nk; kk += NW) {
                            const uint32_t vb = s_vb[kk], dv = s_pre[kk + 1] - s_pre[kk];
                            uint32_t cnt = 0;
                            for (uint32_t j0 = 0; j0 < dv; j0 += 64 * UW) {
                                uint32_t w[UW];
#pragma unroll
                                for (int q = 0; q < UW; q++) {
                                    const uint32_t j = j0 + 64 * q + lane;
                                    w[q] = j < dv ? a.int_adj[vb + j] : SCD_INVALID;
                                }
#pragma unroll
                                for (int q = 0; q < UW; q++) {
                                    if (w[q] == SCD_INVALID) continue;
                                    if (VC && w[q] >= u) {
                                        if (w[q] == u) s_upos[kk] = vb + j0 + 64 * q + lane;
                                        continue;
                                    }
                                    const int pos = set.find(w[q]);
                                    if (pos >= 0) {
                                        atomicAdd(&s_cnt[pos], 1u);
                                        a.flag[vb + j0 + 64 * q + lane] = 1;
                                        cnt++;
                                    }
                                }
                                if (VC && __ballot(w[UW - 1] >= u)) break;  // uniform: past u
                            }
#pragma unroll
                            for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, SCD_WAVE);
                            if (lane == 0 && cnt) {
                                atomicAdd(VC ? &s_ucnt[kk] : &s_cnt[k0 + kk], cnt);
                                tot += cnt;
                            }
                        }
                    } else {
                        // flattened over the threads; WTC_UF loads in flight each
                        uint32_t cur = 0;  // owner list: the thread's f only grows
                        for (uint32_t f0 = tid; f0 < Wc; f0 += NT * WTC_UF) {
                            uint32_t w[WTC_UF], at[WTC_UF], own[WTC_UF];
#pragma unroll
                            for (int q = 0; q < WTC_UF; q++) {
                                const uint32_t f = f0 + NT * q;
                                w[q] = SCD_INVALID;
                                if (f < Wc) {
                                    while (s_pre[cur + 1] <= f) cur++;  // largest kk with pre[kk] <= f
                                    own[q] = cur;
                                    at[q] = s_vb[cur] + (f - s_pre[cur]);
                                    w[q] = a.int_adj[at[q]];
                                }
                            }
#pragma unroll
                            for (int q = 0; q < WTC_UF; q++) {
                                if (w[q] == SCD_INVALID) continue;
                                if (VC && w[q] >= u) {
                                    if (w[q] == u) s_upos[own[q]] = at[q];
                                    continue;
                                }
                                const int pos = set.find(w[q]);
                                if (pos >= 0) {
                                    atomicAdd(VC ? &s_ucnt[own[q]] : &s_cnt[k0 + own[q]], 1u);
                                    atomicAdd(&s_cnt[pos], 1u);
                                    a.flag[at[q]] = 1;
                                    tot++;
                                }
                            }
                        }
                    }
                    __syncthreads();  // s_vb / s_pre reused by the next chunk
                    if (VC) {
                        // the walked vertices: their counts and the edge to the owner
                        if (tid < (int)nk && s_ucnt[tid]) {
                            wtc_add_T(a, N[k0 + tid], s_ucnt[tid]);
                            if (s_upos[tid] != SCD_INVALID) a.flag[s_upos[tid]] = 1;
                        }
                        __syncthreads();
                    }
                }
                for (uint32_t k = tid; k < dp; k += NT) {
                    const uint32_t c = s_cnt[k];
                    if (c) {
                        wtc_add_T(a, N[k], c);
                        a.flag[beg + k] = 1;
                    }
                }
                // leave the bitmap all-zero for the next vertex
                if (set.bits == 0) {
                    for (uint32_t k = tid; k < dp; k += NT) s_tab[(N[k] - set.lo) >> 5].x = 0;
                } else {
                    const uint32_t words = (2 * CAP * 4 + 2 * CAP * 2) / 8;
                    for (uint32_t k = tid; k < words; k += NT) s_tab[k] = make_uint2(0u, 0u);
                }
            } else {
                // out-list longer than the LDS set: search it in global memory
                for (uint32_t c = c0; c < c1; c++) {
                    const uint32_t k0 = c * WTC_VC, nk = min(WTC_VC, dp - k0);
                    if (tid < (int)nk) s_vb[tid] = a.indices[N[k0 + tid]];
                    __syncthreads();
                    for (uint32_t kk = 0; kk < nk; kk++) {
                        const uint32_t v = N[k0 + kk];
                        const uint32_t vb = s_vb[kk], dv = a.dplus[v];
                        for (uint32_t j = tid; j < dv; j += NT) {
                            const uint32_t w = a.int_adj[vb + j];
                            uint32_t l2 = 0, h2 = dp;
                            while (l2 < h2) {
                                const uint32_t mid = (l2 + h2) >> 1;
                                if (N[mid] < w) l2 = mid + 1;
                                else h2 = mid;
                            }
                            if (l2 < dp && N[l2] == w) {
                                wtc_add_T(a, v, 1u);
                                wtc_add_T(a, w, 1u);
                                a.flag[vb + j] = 1;
                                a.flag[beg + k0 + kk] = 1;
                                a.flag[beg + l2] = 1;
                                tot++;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o, SCD_WAVE);
            if (lane == 0 && tot) atomicAdd(&a.T[u], tot);
            __syncthreads();  // LDS reused by the next vertex
            item = c_end;
        }
    }
}

// Flagged internal out-edges count at both ends.  S lanes per vertex of
// one degree bin.  With hub_lds (the whole graph's triangles, where the
// lower ends of most out-edges are the hubs, the lowest ids) a bounded grid
// strides over the bin and each workgroup sums the counts for ids below
// WTC_FIN_HOT in LDS, flushing them once at its end: one device-scope
// atomic per hub and workgroup instead of one per edge.  The WCC's sparse
// internal subgraphs gain nothing from it (one pass per workgroup, direct
// atomics).
constexpr uint32_t WTC_FIN_HOT = 2048;
constexpr uint32_t WTC_FIN_BLOCKS_PER_CU = 4;

template <int S>
__global__ __launch_bounds__(256) void k_wtc_final(WccArgs a, const uint32_t *__restrict__ verts, uint32_t nv_host,
                                                   int hub_lds, const uint32_t *__restrict__ nv_dev) {
    __shared__ uint32_t s_td[WTC_FIN_HOT];
    const uint32_t nv = nv_dev ? *nv_dev : nv_host;
    const int lane = threadIdx.x & 63;
    const int lig = lane & (S - 1);
    constexpr uint32_t G = 256 / S;  // vertices per workgroup pass
    const uint32_t hot = hub_lds ? WTC_FIN_HOT : 0u;  // uniform
    for (uint32_t i = threadIdx.x; i < hot; i += 256) s_td[i] = 0;
    __syncthreads();
    for (uint32_t base = blockIdx.x * G; base < nv; base += gridDim.x * G) {
        const uint32_t slot = base + threadIdx.x / S;
        uint32_t u = 0, beg = 0, dp = 0;
        if (slot < nv) {
            u = verts[slot];
            dp = a.dplus[u];
            if (dp) beg = a.indices[u];
        }
        uint32_t mine = 0;
        for (uint32_t k = lig; k < dp; k += S) {
            if (a.flag[beg + k]) {
                const uint32_t w = a.int_adj[beg + k];
                if (w < hot) atomicAdd(&s_td[w], 1u);
                else wtc_add_td(a, w, 1u);
                mine++;
            }
        }
#pragma unroll
        for (int off = S / 2; off > 0; off >>= 1) mine += __shfl_xor(mine, off, SCD_WAVE);
        if (slot < nv && lig == 0 && mine) atomicAdd(&a.td[u], mine);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < hot; i += 256)
        if (s_td[i]) wtc_add_td(a, i, s_td[i]);
}

// ---------------------------------------------------------------------------
// Dense communities: internal triangles by bitmap rows.
//
// A community of s members whose internal degrees sum to at least
// s (s - 1) / 16 (and 32 <= s <= DN_MAX_S) gives each member p a local id
// and a row of s bits, bit lid(q) set for every internal neighbour q.  Then
// for every internal neighbour q of p, c(p, q) = |Nc(p) ∩ Nc(q)| is the
// popcount of row(p) AND row(q): T[p] = (sum over q of c(p, q)) / 2 and
// td[p] = #{q : c(p, q) > 0} -- the same integers the list kernels count,
// from coalesced row reads instead of per-entry probes (the rows take at
// most half the bytes of the community's internal lists).  Local ids come
// from per-community counters, so their order varies between runs; the
// counts do not depend on it.  Rows take whole 256-byte units of a buffer
// of `units` units, handed out by one counter; a community that does not
// fit stays with the list kernels.
//
// The listed vertices lie community after community, and the waves of the
// row and count kernels stride over the list, so every wave gets members
// of large and small communities alike.  (Handing out runs of the list from
// per-XCD slices by atomic counters kept each community's rows in one
// XCD's L2 but cost more than it saved once arenas listed millions of
// vertices: 4,096 planted graphs, 193 against 145 ms per iteration.)
// ---------------------------------------------------------------------------
constexpr uint32_t DN_NONE = 0xFFFFFFFFu;
constexpr uint32_t DN_MIN_S = 32;
constexpr uint32_t DN_MAX_S = 16384;  // rows of at most 256 words
constexpr uint32_t DN_UNIT_WORDS = 32;  // 256-byte allocation units
constexpr uint32_t DN_CTR = 2;  // [0] units handed out, [1] vertices listed

struct DenseArgs {
    const uint32_t *__restrict__ labels;
    const int4 *__restrict__ stats;
    const uint8_t *__restrict__ dirty;  // nullptr: every community
    uint32_t *__restrict__ off;      // per label: first unit of its rows, or DN_NONE
    uint32_t *__restrict__ vbase;    // per dense label: first list slot of its members
    uint32_t *__restrict__ cursor;   // per label: local ids handed out
    uint32_t *__restrict__ lid;      // per vertex of a dense community
    uint32_t *__restrict__ list;     // the vertices of dense communities, by community
    uint32_t *__restrict__ ctr;      // DN_CTR counters, cleared before each count
    unsigned long long *__restrict__ bits;
    uint32_t units;                  // capacity of bits in units
};

__device__ __forceinline__ uint32_t dn_words(uint32_t s) { return (s + 63u) >> 6; }

// off[L], vbase[L] and cursor[L] for every label.
__global__ void k_dn_classify(DenseArgs d, uint32_t num_comms) {
    const uint32_t L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= num_comms) return;
    d.cursor[L] = 0;
    const int4 st = d.stats[L];
    const uint64_t s = (uint32_t)st.x, z = (uint32_t)st.z;
    uint32_t off = DN_NONE;
    if (s >= DN_MIN_S && s <= DN_MAX_S && 16ull * z >= s * (s - 1) && (!d.dirty || d.dirty[L])) {
        const uint32_t need = (uint32_t)((s * dn_words((uint32_t)s) + DN_UNIT_WORDS - 1) / DN_UNIT_WORDS);
        const uint32_t at = atomicAdd(&d.ctr[0], need);
        if ((uint64_t)at + need <= d.units) {
            off = at;
            d.vbase[L] = atomicAdd(&d.ctr[1], (uint32_t)s);
        }
    }
    d.off[L] = off;
}

// Local ids and list slots of the dense-community vertices (T and td
// cleared): per wave, one counter update per distinct label.
__global__ __launch_bounds__(256) void k_dn_lid(WccArgs a, DenseArgs d, uint32_t n) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    const int lane = threadIdx.x & 63;
    uint32_t L = SCD_INVALID;
    if (p < n) L = d.labels[p];
    const bool dense = L != SCD_INVALID && d.off[L] != DN_NONE;
    const uint64_t below = (1ull << lane) - 1ull;
    uint64_t todo = __ballot(dense);
    if (!todo) return;  // uniform per wave
    uint32_t base = 0, my_leader = 0, my_rank = 0;
    while (todo) {  // uniform
        const int l = __builtin_ctzll(todo);
        const uint32_t Ll = __shfl(L, l, SCD_WAVE);
        const uint64_t m = __ballot(dense && L == Ll);
        if (lane == l) base = atomicAdd(&d.cursor[Ll], (uint32_t)__popcll(m));
        if ((m >> lane) & 1ull) {
            my_leader = (uint32_t)l;
            my_rank = (uint32_t)__popcll(m & below);
        }
        todo &= ~m;
    }
    const uint32_t b = __shfl(base, (int)my_leader, SCD_WAVE);
    if (dense) {
        const uint32_t id = b + my_rank;
        d.lid[p] = id;
        d.list[d.vbase[L] + id] = p;
        a.T[p] = 0;
        a.td[p] = 0;
    }
}

// Row of every listed vertex: a wave per vertex builds it in LDS (64-bit
// atomic ORs at the neighbours' local ids) and stores it whole.
__global__ __launch_bounds__(256) void k_dn_rows(WccArgs a, DenseArgs d) {
    __shared__ unsigned long long s_row[4][DN_MAX_S / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned long long *row = s_row[wv];
    const uint32_t nd = d.ctr[1];
    for (uint32_t slot = blockIdx.x * 4u + wv; slot < nd; slot += gridDim.x * 4u) {  // uniform per wave
        const uint32_t p = d.list[slot];
        const uint32_t L = d.labels[p];
        const uint32_t W = dn_words((uint32_t)d.stats[L].x);
        for (uint32_t k = lane; k < W; k += 64) row[k] = 0ull;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const uint32_t beg = a.indices[p], dg = a.int_deg[p];
        for (uint32_t k = lane; k < dg; k += 64) {
            const uint32_t l = d.lid[a.int_adj[beg + k]];
            atomicOr(&row[l >> 6], 1ull << (l & 63u));
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        unsigned long long *out = d.bits + (size_t)d.off[L] * DN_UNIT_WORDS + (size_t)d.lid[p] * W;
        for (uint32_t k = lane; k < W; k += 64) out[k] = row[k];
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();  // the row is reused by the wave's next vertex
    }
}

// Sum over the wave, in every lane.
__device__ __forceinline__ uint32_t dn_wave_sum(uint32_t x) {
    x += lane_xor<1>(x);
    x += lane_xor<2>(x);
    x += lane_xor<4>(x);
    x += lane_xor<8>(x);
    x += lane_xor<16>(x);
    x += lane_xor<32>(x);
    return x;
}

// T (the sum of c(p, q), halved by k_dn_halve) and td of every listed
// vertex, a wave per vertex, DN_U neighbour rows in flight (2: 4,096
// planted graphs, 164 ms for the second iteration against 176 with 4, whose
// registers cost occupancy, and 165 with 1).  Rows of W <= 64 words: the
// wave splits into 64 / Wp groups (Wp = W rounded up to a power of two),
// each ANDing one neighbour's row with its own copy of row(p) per load.
// Longer rows (communities of more than 4,096 members, which take most of
// the bytes): every lane holds up to four words of row(p), the neighbour is
// the same for the whole wave, and each internal edge is counted once, from
// its end with the lower local id, which adds c(p, q) to q's counts by
// atomics: half the row reads (4,096 planted graphs 574 -> 547 ms; the same
// for the short rows, with a group sum per neighbour, took it to 598 ms).
constexpr int DN_U = 2;

__global__ __launch_bounds__(256) void k_dn_count(WccArgs a, DenseArgs d) {
    __shared__ uint32_t s_qv[4][64], s_ql[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t nd = d.ctr[1];
    for (uint32_t slot = blockIdx.x * 4u + wv; slot < nd; slot += gridDim.x * 4u) {  // uniform per wave
        const uint32_t p = d.list[slot];
        const uint32_t L = d.labels[p];
        const uint32_t W = dn_words((uint32_t)d.stats[L].x);
        const unsigned long long *rows = d.bits + (size_t)d.off[L] * DN_UNIT_WORDS;
        const unsigned long long *rp = rows + (size_t)d.lid[p] * W;
        const uint32_t beg = a.indices[p], dg = a.int_deg[p];
        uint32_t pc = 0, td = 0;
        if (W <= 64) {
            uint32_t Wp = 1;
            while (Wp < W) Wp <<= 1;
            const uint32_t G = 64u / Wp, g = (uint32_t)lane / Wp, k = (uint32_t)lane & (Wp - 1u);
            // bit g * Wp of `firsts`: the first lane of group g
            unsigned long long firsts = 0;
            for (uint32_t gg = 0; gg < G; gg++) firsts |= 1ull << (gg * Wp);
            const unsigned long long mine = k < W ? rp[k] : 0ull;
            for (uint32_t c0 = 0; c0 < dg; c0 += 64) {
                const uint32_t nq = min(64u, dg - c0);
                const uint32_t ql = (uint32_t)lane < nq ? d.lid[a.int_adj[beg + c0 + lane]] : 0u;
                // batch b: neighbour b * G + g for group g
                for (uint32_t b0 = 0; b0 * G < nq; b0 += DN_U) {
                    unsigned long long xv[DN_U];
#pragma unroll
                    for (int u = 0; u < DN_U; u++) {
                        const uint32_t qi = (b0 + u) * G + g;
                        const uint32_t l = __shfl(ql, (int)(qi & 63u), SCD_WAVE);
                        xv[u] = (qi < nq && k < W) ? rows[(size_t)l * W + k] : 0ull;
                    }
#pragma unroll
                    for (int u = 0; u < DN_U; u++) {
                        const unsigned long long v = xv[u] & mine;
                        pc += (uint32_t)__popcll(v);
                        unsigned long long f = __ballot(v != 0ull);
                        for (uint32_t sh = 1; sh < Wp; sh <<= 1) f |= f >> sh;  // a group's OR at its first bit
                        td += (uint32_t)__popcll(f & firsts);
                    }
                }
            }
        } else {
            constexpr int NJ = DN_MAX_S / 64 / 64;
            unsigned long long mine[NJ];
#pragma unroll
            for (int j = 0; j < NJ; j++) {
                const uint32_t k = (uint32_t)lane + 64u * j;
                mine[j] = k < W ? rp[k] : 0ull;
            }
            const uint32_t nj = (W + 63u) >> 6;  // words per lane in use (2..4)
            const uint32_t lp = d.lid[p];
            uint32_t *sqv = s_qv[wv], *sql = s_ql[wv];
            for (uint32_t c0 = 0; c0 < dg; c0 += 64) {
                // the neighbours above p in local id, compacted to the first nk lanes
                uint32_t qv = 0, ql = 0;
                bool keep = false;
                if (c0 + (uint32_t)lane < dg) {
                    qv = a.int_adj[beg + c0 + lane];
                    ql = d.lid[qv];
                    keep = ql > lp;
                }
                const uint64_t km = __ballot(keep);
                const uint32_t nk = (uint32_t)__popcll(km);
                if (keep) {
                    sqv[__popcll(km & below)] = qv;
                    sql[__popcll(km & below)] = ql;
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
                const uint32_t kv = (uint32_t)lane < nk ? sqv[lane] : 0u;
                const uint32_t kl = (uint32_t)lane < nk ? sql[lane] : 0u;
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();  // the next chunk refills sqv / sql
                for (uint32_t i0 = 0; i0 < nk; i0 += DN_U) {
                    unsigned long long xv[DN_U][NJ];
#pragma unroll
                    for (int u = 0; u < DN_U; u++) {
                        const uint32_t i = min(i0 + u, nk - 1u);  // uniform
                        const uint32_t l = (uint32_t)__builtin_amdgcn_readlane((int)kl, (int)i);
                        const unsigned long long *rq = rows + (size_t)l * W;
#pragma unroll
                        for (int j = 0; j < NJ; j++) {
                            const uint32_t k = (uint32_t)lane + 64u * j;
                            xv[u][j] = ((uint32_t)j < nj && i0 + u < nk && k < W) ? rq[k] : 0ull;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < DN_U; u++) {
                        if (i0 + u >= nk) break;  // uniform
                        uint32_t local = 0;
#pragma unroll
                        for (int j = 0; j < NJ; j++) local += (uint32_t)__popcll(xv[u][j] & mine[j]);
                        const uint32_t c = dn_wave_sum(local);
                        pc += local;
                        if (c) {
                            td++;
                            if (lane == 0) {
                                const uint32_t q = (uint32_t)__builtin_amdgcn_readlane((int)kv, (int)(i0 + u));
                                atomicAdd(&a.T[q], c);
                                atomicAdd(&a.td[q], 1u);
                            }
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) pc += __shfl_xor(pc, o, SCD_WAVE);
        if (lane == 0) {
            if (pc) atomicAdd(&a.T[p], pc);
            if (td) atomicAdd(&a.td[p], td);
        }
    }
}

// T[p] = (sum over q of c(p, q)) / 2 for the listed vertices.
__global__ void k_dn_halve(WccArgs a, DenseArgs d) {
    const uint32_t nd = d.ctr[1];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nd; i += gridDim.x * blockDim.x) {
        const uint32_t p = d.list[i];
        a.T[p] >>= 1;
    }
}

// tri[u] = 2 x triangles at u (the layout of Graph::triangles).
__global__ void k_wtc_double(const uint32_t *__restrict__ T, uint32_t n, int32_t *__restrict__ tri) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < n) tri[u] = (int32_t)(2u * T[u]);
}

// Per-vertex terms (src/wcc.cpp:89-166): vertices without a valid label, in
// a community of size <= 2 or without triangles contribute 0.  The hot ids'
// private triangle copies are folded into T / td here first (the triangle
// pass leaves them to this kernel).
__global__ void k_wcc_terms(WccArgs a, uint32_t n) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (p < a.hot_n) wtc_hot_fold(a, p);
    const uint32_t L = a.labels[p];
    bool valid = false;
    int size = 0;
    if (L != SCD_INVALID) {
        size = a.stats[L].x;
        valid = size > 2 && a.tri_deg[p] != 0;
    }
    const uint32_t itri = 2u * a.T[p], itdeg = a.td[p];
    a.terms[p] = valid ? wcc_term(a.tri[p], a.tri_deg[p], size, (int)itri, (int)itdeg) : 0.0;
    if (a.int_tri_out) {
        a.int_tri_out[p] = valid ? (int)itri : 0;
        a.int_tri_deg_out[p] = valid ? (int)itdeg : 0;
    }
}

}  // namespace scd_dev

#endif
