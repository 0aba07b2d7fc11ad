// Degree bins on the device: every vertex goes to the bin its degree
// selects (which kernel shape handles it), and the vertex lists of all bins
// are laid out bin after bin, each in ascending vertex order (a stable
// partition), in two passes over the CSR offsets: per-workgroup bin counts,
// then a scatter to the offsets the host derives from those counts.
#ifndef SCD_BINS_KERNELS_HPP
#define SCD_BINS_KERNELS_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev.hpp"

namespace scd_dev {

enum BinKind : uint32_t { K_SUB8, K_SUB16, K_SUB32, K_W1, K_W2, K_W4, K_W8, K_W16, K_LDS, K_LDS2, K_HUB, K_NB };

// Largest degree of the K_LDS bin: with lds_tile >= 8192 the workgroup-per-
// vertex range is split at lds_tile / 2, so that the vertices up to there
// run the kernel shape sized for lds_tile / 2 keys (99 -> 66 registers at
// 8192: 4 -> 7 waves per SIMD) and only the longer lists the one sized for
// lds_tile.  Below 8192 both shapes fit 7 waves and one launch is better.
__host__ __device__ inline uint32_t lds_split(uint32_t lds_tile) { return lds_tile >= 8192 ? lds_tile / 2 : lds_tile; }

// The bin of a vertex of degree d: sub-wave groups up to 32 neighbours,
// 1-16 waves up to wave_max, an LDS tile up to lds_tile (two bins, split at
// lds_split), hub beyond.
__host__ __device__ inline uint32_t bin_of(uint32_t d, uint32_t lds_tile, uint32_t wave_max) {
    if (d <= 8) return K_SUB8;
    if (d <= 16) return K_SUB16;
    if (d <= 32) return K_SUB32;
    if (d > wave_max || d > 1024) {
        if (d <= lds_split(lds_tile)) return K_LDS;
        return d <= lds_tile ? K_LDS2 : K_HUB;
    }
    if (d <= 64) return K_W1;
    if (d <= 128) return K_W2;
    if (d <= 256) return K_W4;
    if (d <= 512) return K_W8;
    return K_W16;
}

constexpr uint32_t BIN_CHUNK = 4096;  // vertices per workgroup (256 threads x 16)

// A vertex is binned unless gid is given and its graph has stopped
// (active_g[gid[v]] == 0): a batched arena's bins can leave out the graphs
// that no longer refine.
__device__ __forceinline__ bool bin_keep(const uint32_t *gid, const uint32_t *active_g, uint32_t v) {
    return !gid || active_g[gid[v]];
}

// cnt[blk * K_NB + b] = vertices of bin b in workgroup blk's chunk;
// esum[b] += their degrees; *dmax = the largest degree.
__global__ __launch_bounds__(256) void k_bin_count(const uint32_t *__restrict__ indices, uint32_t n, uint32_t lds_tile,
                                                   uint32_t wave_max, const uint32_t *__restrict__ gid,
                                                   const uint32_t *__restrict__ active_g, uint32_t *__restrict__ cnt,
                                                   unsigned long long *__restrict__ esum, uint32_t *__restrict__ dmax) {
    __shared__ uint32_t s_cnt[K_NB];
    __shared__ unsigned long long s_e[K_NB];
    __shared__ uint32_t s_max;
    if (threadIdx.x < K_NB) {
        s_cnt[threadIdx.x] = 0;
        s_e[threadIdx.x] = 0;
    }
    if (threadIdx.x == 0) s_max = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * BIN_CHUNK;
    uint32_t mx = 0;
    for (uint32_t k = threadIdx.x; k < BIN_CHUNK; k += 256) {
        const uint32_t v = base + k;
        if (v >= n) break;
        if (!bin_keep(gid, active_g, v)) continue;
        const uint32_t d = indices[v + 1] - indices[v];
        const uint32_t b = bin_of(d, lds_tile, wave_max);
        atomicAdd(&s_cnt[b], 1u);
        atomicAdd(&s_e[b], (unsigned long long)d);
        mx = d > mx ? d : mx;
    }
    atomicMax(&s_max, mx);
    __syncthreads();
    if (threadIdx.x < K_NB) {
        cnt[blockIdx.x * K_NB + threadIdx.x] = s_cnt[threadIdx.x];
        if (s_e[threadIdx.x]) atomicAdd(&esum[threadIdx.x], s_e[threadIdx.x]);
    }
    if (threadIdx.x == 0) atomicMax(dmax, s_max);
}

// bin_verts[off[blk * K_NB + b] + rank] = v for the vertices v of bin b in
// workgroup blk's chunk, rank counting the chunk's earlier vertices of bin b
// (ascending vertex order within every bin).
__global__ __launch_bounds__(256) void k_bin_scatter(const uint32_t *__restrict__ indices, uint32_t n,
                                                     uint32_t lds_tile, uint32_t wave_max,
                                                     const uint32_t *__restrict__ gid,
                                                     const uint32_t *__restrict__ active_g,
                                                     const uint32_t *__restrict__ off,
                                                     uint32_t *__restrict__ bin_verts) {
    __shared__ uint32_t s_off[K_NB];
    __shared__ uint32_t s_wc[4][K_NB];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x < K_NB) s_off[threadIdx.x] = off[blockIdx.x * K_NB + threadIdx.x];
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;  // lanes below this one
    const uint32_t base = blockIdx.x * BIN_CHUNK;
    for (uint32_t k = 0; k < BIN_CHUNK; k += 256) {
        const uint32_t v = base + k + threadIdx.x;
        const bool act = v < n && bin_keep(gid, active_g, v);
        const uint32_t b = act ? bin_of(indices[v + 1] - indices[v], lds_tile, wave_max) : K_NB;
        uint32_t rank = 0;
        for (uint32_t j = 0; j < K_NB; j++) {
            const uint64_t m = __ballot(b == j);
            if (b == j) rank = __popcll(m & lt);
            if (lane == 0) s_wc[w][j] = __popcll(m);
        }
        __syncthreads();
        if (act) {
            uint32_t pos = s_off[b] + rank;
            for (int q = 0; q < w; q++) pos += s_wc[q][b];
            bin_verts[pos] = v;
        }
        __syncthreads();
        if (threadIdx.x < K_NB)
            s_off[threadIdx.x] += s_wc[0][threadIdx.x] + s_wc[1][threadIdx.x] + s_wc[2][threadIdx.x] +
                                  s_wc[3][threadIdx.x];
        __syncthreads();
    }
}

// deg[i] = degree of verts[i], i < count.
__global__ void k_bin_degrees(const uint32_t *__restrict__ indices, const uint32_t *__restrict__ verts,
                              uint32_t count, uint32_t *__restrict__ deg) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) deg[i] = indices[verts[i] + 1] - indices[verts[i]];
}

}  // namespace scd_dev

#endif
