// Per-vertex triangle statistics on the device (the convert tool's t / vt,
// src/convert.cpp:271-340): tri[x] = sum over x's edges of the number of
// triangles through the edge (= 2 x triangles at x), tri_deg[x] = number of
// x's edges that lie in a triangle.
//
// Orientation: every edge points from the higher id to the lower id, so the
// out-list N+(u) = {v in N(u) : v < u} is the prefix of u's sorted list
// (length dplus[u]).  In convert-relabelled graphs ids are degree-descending,
// so this orientation points at the high-degree end and keeps out-lists short.
// Each triangle w < v < u is found exactly once: at u, for v in N+(u), as w
// in N+(v) ∩ N+(u).  Per find:
//   T[u], T[v], T[w] += 1       (triangles per vertex; aggregated per u in
//                                LDS counters over N+(u), flushed with one
//                                atomic per out-edge)
//   flag(v,w) = 1               (edge v-w is in a triangle; a plain store of 1
//                                to the slot of w in v's list — writers only
//                                ever store 1, so no atomics are needed)
//   flag(u,v), flag(u,w) = 1    (set at the flush from the LDS counters)
// Then tri[x] = 2 T[x] and tri_deg[x] = number of flagged edges at x.
//
// Work is the sum over u of sum_{v in N+(u)} dplus[v] probes; u with short
// out-lists and little work run one per wave with N+(u) held across the
// wave's lanes (light); the rest are cut into chunks of TC_CHUNK probes, one
// workgroup per chunk, with N+(u) staged in LDS (heavy).
#ifndef SCD_TC_KERNELS_HPP
#define SCD_TC_KERNELS_HPP

#include "dev.hpp"
#include "stats_kernels.hpp"

namespace scd_dev {

constexpr uint32_t TC_LIGHT_W = 4096;  // max probes of a light vertex
constexpr uint32_t TC_CHUNK = 8192;    // probes per heavy work item
constexpr uint32_t TC_CAP = 2048;      // out-list entries staged in LDS

struct TcArgs {
    const uint32_t *__restrict__ indices;
    const uint32_t *__restrict__ adj;
    uint32_t n;
    uint32_t *__restrict__ dplus;    // n
    uint32_t *__restrict__ T;        // n, triangles per vertex
    uint32_t *__restrict__ td;       // n, flagged edges per vertex
    uint8_t *__restrict__ flag;      // E, by adjacency slot (out-slots only)
    uint32_t *__restrict__ gpre;     // E, out-list prefix sums of heavy u with dplus > TC_CAP
    uint32_t *__restrict__ light;    // light vertex list
    uint32_t *__restrict__ heavy;    // heavy vertex list
    uint32_t *__restrict__ heavy_chunks;  // chunks per heavy vertex
    uint32_t *__restrict__ counts;   // [0] light, [1] heavy
};

// dplus[u] = #neighbours below u; clears T, td.
__global__ void k_tc_dplus(TcArgs a) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= a.n) return;
    const uint32_t b = a.indices[u], e = a.indices[u + 1];
    uint32_t lo = b, hi = e;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.adj[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    a.dplus[u] = lo - b;
    a.T[u] = 0;
    a.td[u] = 0;
}

// Probe count W(u) and the light / heavy split.
__global__ void k_tc_classify(TcArgs a) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= a.n) return;
    const uint32_t b = a.indices[u], dp = a.dplus[u];
    uint64_t W = 0;
    for (uint32_t k = 0; k < dp; k++) {
        if (dp > TC_CAP) a.gpre[b + k] = (uint32_t)W;
        W += a.dplus[a.adj[b + k]];
    }
    if (W == 0) return;
    if (dp <= 64 && W <= TC_LIGHT_W) {
        a.light[atomicAdd(&a.counts[0], 1u)] = u;
    } else {
        const uint32_t i = atomicAdd(&a.counts[1], 1u);
        a.heavy[i] = u;
        a.heavy_chunks[i] = (uint32_t)((W + TC_CHUNK - 1) / TC_CHUNK);
    }
}

// Expands heavy vertices into (vertex, chunk) items at the scanned offsets.
__global__ void k_tc_items(const uint32_t *__restrict__ heavy, const uint32_t *__restrict__ chunks,
                           const uint32_t *__restrict__ off, uint32_t nheavy,
                           uint32_t *__restrict__ item_u, uint32_t *__restrict__ item_c) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nheavy) return;
    for (uint32_t c = 0; c < chunks[i]; c++) {
        item_u[off[i] + c] = heavy[i];
        item_c[off[i] + c] = c;
    }
}

// Light: one wave per u, N+(u) (<= 64 entries) across the lanes; searches
// over it use shuffles.
__global__ __launch_bounds__(256) void k_tc_light(TcArgs a) {
    __shared__ uint32_t s_cnt[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t slot = blockIdx.x * 4u + wv;
    if (slot >= a.counts[0]) return;
    const uint32_t u = a.light[slot];
    const uint32_t bu = a.indices[u], dp = a.dplus[u];
    uint32_t nk = 0xFFFFFFFFu, vb = 0, dv = 0;
    if ((uint32_t)lane < dp) {
        nk = a.adj[bu + lane];
        vb = a.indices[nk];
        dv = a.dplus[nk];
    }
    // exclusive prefix of dv across the wave
    uint32_t inc = dv;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(inc, off, SCD_WAVE);
        if (lane >= off) inc += t;
    }
    const uint32_t pre = inc - dv;
    const uint32_t W = __shfl(inc, 63, SCD_WAVE);
    s_cnt[wv][lane] = 0;
    uint32_t tot = 0;
    for (uint32_t f0 = 0; f0 < W; f0 += 64) {
        const uint32_t f = f0 + lane;
        // kv = largest k < dp with pre[k] <= f
        int kv = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
            const int cand = kv + step;
            const uint32_t pc = __shfl(pre, cand & 63, SCD_WAVE);
            if (cand < (int)dp && pc <= f) kv = cand;
        }
        const uint32_t j = f - __shfl(pre, kv, SCD_WAVE);
        const uint32_t vbase = __shfl(vb, kv, SCD_WAVE);
        const bool act = f < W;
        const uint32_t w = act ? a.adj[vbase + j] : 0xFFFFFFFFu;
        // kw = largest k < dp with nk[k] <= w
        int kw = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
            const int cand = kw + step;
            const uint32_t nc = __shfl(nk, cand & 63, SCD_WAVE);
            if (cand < (int)dp && nc <= w) kw = cand;
        }
        // every lane shuffles (a shuffle from a lane that is switched off
        // in a branch does not return that lane's value)
        const uint32_t nkw = __shfl(nk, kw, SCD_WAVE);
        const bool hit = act && nkw == w;
        if (hit) {
            atomicAdd(&s_cnt[wv][kv], 1u);
            atomicAdd(&s_cnt[wv][kw], 1u);
            a.flag[vbase + j] = 1;
            tot++;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off, SCD_WAVE);
    if ((uint32_t)lane < dp) {
        const uint32_t c = s_cnt[wv][lane];
        if (c) {
            atomicAdd(&a.T[nk], c);
            a.flag[bu + lane] = 1;
        }
    }
    if (lane == 0 && tot) atomicAdd(&a.T[u], tot);
}

// Heavy: one workgroup per (u, chunk); N+(u), the base offsets of its
// members' lists and the probe prefix in LDS (global for dplus > TC_CAP).
__global__ __launch_bounds__(256) void k_tc_heavy(TcArgs a, const uint32_t *__restrict__ item_u,
                                                  const uint32_t *__restrict__ item_c) {
    __shared__ uint32_t s_nk[TC_CAP], s_vb[TC_CAP], s_pre[TC_CAP], s_cnt[TC_CAP];
    __shared__ uint32_t s_w[4];
    const int tid = threadIdx.x;
    const uint32_t u = item_u[blockIdx.x];
    const uint32_t c = item_c[blockIdx.x];
    const uint32_t bu = a.indices[u], dp = a.dplus[u];
    const bool in_lds = dp <= TC_CAP;
    const uint32_t *nk = in_lds ? s_nk : a.adj + bu;
    const uint32_t *pre = in_lds ? s_pre : a.gpre + bu;
    uint32_t W;
    if (in_lds) {
        // stage: thread t owns entries [8t, 8t+8)
        uint32_t loc[8], sum = 0;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const uint32_t k = tid * 8 + q;
            uint32_t d = 0;
            if (k < dp) {
                const uint32_t v = a.adj[bu + k];
                s_nk[k] = v;
                s_vb[k] = a.indices[v];
                d = a.dplus[v];
                s_cnt[k] = 0;
            }
            loc[q] = d;
            sum += d;
        }
        uint32_t total;
        uint32_t run = block_excl_scan256(sum, s_w, total);
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const uint32_t k = tid * 8 + q;
            if (k < dp) s_pre[k] = run;
            run += loc[q];
        }
        W = total;
        __syncthreads();
    } else {
        W = a.gpre[bu + dp - 1] + a.dplus[a.adj[bu + dp - 1]];
    }
    const uint64_t f0 = (uint64_t)c * TC_CHUNK;
    const uint32_t f1 = (uint32_t)min<uint64_t>(f0 + TC_CHUNK, W);
    uint32_t tot = 0;
    for (uint32_t f = (uint32_t)f0 + tid; f < f1; f += 256) {
        // kv = largest k with pre[k] <= f
        uint32_t lo = 0, hi = dp;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (pre[mid] <= f) lo = mid;
            else hi = mid;
        }
        const uint32_t kv = lo;
        const uint32_t vbase = in_lds ? s_vb[kv] : a.indices[nk[kv]];
        const uint32_t j = f - pre[kv];
        const uint32_t w = a.adj[vbase + j];
        // w in N+(u)?
        uint32_t l2 = 0, h2 = dp;
        while (l2 < h2) {
            const uint32_t mid = (l2 + h2) >> 1;
            if (nk[mid] < w) l2 = mid + 1;
            else h2 = mid;
        }
        if (l2 < dp && nk[l2] == w) {
            if (in_lds) {
                atomicAdd(&s_cnt[kv], 1u);
                atomicAdd(&s_cnt[l2], 1u);
            } else {
                atomicAdd(&a.T[nk[kv]], 1u);
                atomicAdd(&a.T[w], 1u);
                a.flag[bu + kv] = 1;
                a.flag[bu + l2] = 1;
            }
            a.flag[vbase + j] = 1;
            tot++;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off, SCD_WAVE);
    if ((tid & 63) == 0 && tot) atomicAdd(&a.T[u], tot);
    if (in_lds) {
        __syncthreads();
        for (uint32_t k = tid; k < dp; k += 256) {
            const uint32_t cnt = s_cnt[k];
            if (cnt) {
                atomicAdd(&a.T[s_nk[k]], cnt);
                a.flag[bu + k] = 1;
            }
        }
    }
}

// tri_deg from the out-slot flags (each flagged out-edge counts at both
// ends); tri = 2 T.  One wave per vertex.
__global__ __launch_bounds__(256) void k_tc_final(TcArgs a, int32_t *__restrict__ tri) {
    const int lane = threadIdx.x & 63;
    const uint32_t u = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (u >= a.n) return;
    const uint32_t bu = a.indices[u], dp = a.dplus[u];
    uint32_t mine = 0;
    for (uint32_t k = lane; k < dp; k += 64) {
        if (a.flag[bu + k]) {
            atomicAdd(&a.td[a.adj[bu + k]], 1u);
            mine++;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mine += __shfl_xor(mine, off, SCD_WAVE);
    if (lane == 0) {
        if (mine) atomicAdd(&a.td[u], mine);
        tri[u] = (int32_t)(2u * a.T[u]);
    }
}

__global__ void k_fill_u8(uint8_t *__restrict__ p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace scd_dev

#endif
