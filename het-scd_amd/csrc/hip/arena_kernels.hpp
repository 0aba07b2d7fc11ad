// Arena assembly: graphs held on the device one by one (local ids) copied
// into one disjoint union (global ids), and labels moved between the two
// numberings.  Plain element-wise kernels, one thread per element.
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

}  // namespace scd_dev

#endif
