// Device helpers shared by the SCD kernels: the exact improvement formula,
// wave64 register bitonic sort, merge-path search and the (improvement,
// label) argmax order.
//
// Arithmetic rules: int32 sub-expressions wrap (computed in uint32 and cast
// back), every double operation is its own IEEE-754 binary64 step (the file
// is compiled with -ffp-contract=off and the pragma below), divisions are the
// compiler's correctly rounded fp64 division.
#ifndef SCD_DEV_HPP
#define SCD_DEV_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

#define SCD_INVALID 0xFFFFFFFFu
#define SCD_WAVE 64

namespace scd_dev {

__device__ __forceinline__ int wmul(int a, int b) { return (int)((uint32_t)a * (uint32_t)b); }
__device__ __forceinline__ int wadd(int a, int b) { return (int)((uint32_t)a + (uint32_t)b); }
__device__ __forceinline__ int wsub(int a, int b) { return (int)((uint32_t)a - (uint32_t)b); }

// calculate_improvement, src/scd.cpp:173-219, operation by operation.
// r = community size, d_in / d_out = the vertex's edges into / out of the
// community, c_out = the community's boundary edges, alpha = 1.0.
__device__ __forceinline__ double improvement(int r, int d_in, int d_out, int c_out, double p_in,
                                              double p_ext, double alpha) {
    double t;
    if (r > 0) t = (double)wsub(c_out, d_in) / (double)r;
    else t = 0.0;

    const double din_din1 = (double)wmul(d_in, wsub(d_in, 1));
    const double dout_dout1 = (double)wmul(d_out, wsub(d_out, 1));
    const double dout_din = (double)wmul(d_out, d_in);
    const double rm1 = (double)wsub(r, 1);
    const double rm1rm2 = (double)wmul(wsub(r, 1), wsub(r, 2));
    const double two_din1 = (double)wmul(2, wsub(d_in, 1));
    const double din_plus_dout = (double)wadd(d_in, d_out);

    double A = 0.0;
    double denom = (din_din1 * p_in + dout_dout1 * p_ext) + dout_din * p_ext;
    denom *= din_plus_dout + alpha * (double)wsub(wsub(r, 1), d_in);
    if (denom != 0.0) A = ((din_din1 * p_in) * din_plus_dout) / denom;

    double BMinus = 0.0;
    denom = rm1rm2 * p_in * p_in * p_in + two_din1 * p_in + t * rm1 * p_in * p_ext +
            t * (t - 1) * p_ext + (double)d_out * p_ext;
    denom *= rm1 * p_in + 1 + t + alpha * ((double)r - rm1 * p_in - 1);
    if (denom != 0.0) BMinus = (two_din1 * p_in) * (rm1 * p_in + 1 + t) / denom;

    double CMinus = 0.0;
    denom = rm1rm2 * p_in * p_in * p_in + t * (t - 1) * p_ext + t * rm1 * p_in * p_ext;
    denom *= rm1 * p_in + t + alpha * ((double)r - rm1 * p_in);
    denom *= rm1 * p_in + t + alpha * ((double)r - rm1 * p_in - 1);
    if (denom != 0.0 && (((double)r + t) > 0) && ((rm1 + t) > 0))
        CMinus = -(rm1rm2 * p_in * p_in * p_in) * (rm1 * p_in + t) * alpha / denom;

    return A + (double)d_in * BMinus + (double)wsub(r, d_in) * CMinus;
}

// Candidate set-up of src/scd.cpp:280-307 for moving a vertex of degree
// `deg` and label `my` into community L, seen `freq` times among its
// neighbours; s = stats[L] = (size, ext_edges, int_edges, unused).
__device__ __forceinline__ double candidate_improvement_s(int4 s, uint32_t L, uint32_t freq, uint32_t deg,
                                                          uint32_t my, double p_ext) {
    int size = s.x;
    int ext_edge = s.y;
    int int_edge = s.z / 2;
    if (L == my) {
        size = wsub(size, 1);
        ext_edge = wadd(ext_edge, (int)(freq - (deg - freq)));
        int_edge = wsub(int_edge, (int)freq);
    }
    const int d_in = (int)freq;
    const int d_out = wsub((int)deg, d_in);
    double p_int = 0.0;
    if (size > 1) p_int = (2.0 * (double)int_edge) / ((double)size * ((double)size - 1.0));
    return improvement(size, d_in, d_out, ext_edge, p_int, p_ext, 1.0);
}

__device__ __forceinline__ double candidate_improvement(uint32_t L, uint32_t freq, uint32_t deg,
                                                        uint32_t my, const int4 *__restrict__ stats,
                                                        double p_ext) {
    return candidate_improvement_s(stats[L], L, freq, deg, my, p_ext);
}

// The best-move order of src/scd.cpp:309-320: the sequential scan over
// ascending labels keeps the first strictly larger improvement, starting
// from (0.0, label 0).  Equivalently: among candidates with improvement
// > 0.0, the largest improvement, ties to the smallest label.  Candidates
// are carried as (imp, label) with non-candidates (0.0, INVALID); this
// order is total, so any reduction tree gives the same winner.
__device__ __forceinline__ bool better(double a, uint32_t la, double b, uint32_t lb) {
    return a > b || (a == b && la < lb);
}

// Move rule, src/scd.cpp:321-330.
__device__ __forceinline__ uint32_t move_rule(double best_imp, uint32_t best_label, double rem,
                                              uint32_t my) {
    const uint32_t insert_label = (best_label == SCD_INVALID) ? 0u : best_label;
    const double ins = best_imp + rem;
    if (ins > rem && ins > 0.0) return insert_label;
    if (rem > 0) return SCD_INVALID;
    return my;
}

// The value of lane (lane ^ J), J < 64, without the LDS crossbar that
// shuffles use: DPP permutations inside rows of 16 lanes for J = 1, 2 (one
// quad_perm) and 4, 8 (a row or half-row mirror, then a second one), and
// the gfx950 row swaps v_permlane16_swap / v_permlane32_swap for J = 16, 32
// (called with the value as both operands, each half of the result holds
// the partner rows of one row parity).
template <int J>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v) {
    if constexpr (J == 1) {
        return __builtin_amdgcn_update_dpp(0u, v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    } else if constexpr (J == 2) {
        return __builtin_amdgcn_update_dpp(0u, v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    } else if constexpr (J == 4) {
        const uint32_t m = __builtin_amdgcn_update_dpp(0u, v, 0x141, 0xF, 0xF, false);  // row_half_mirror: ^7
        return __builtin_amdgcn_update_dpp(0u, m, 0x1B, 0xF, 0xF, false);             // quad_perm [3,2,1,0]: ^3
    } else if constexpr (J == 8) {
        const uint32_t m = __builtin_amdgcn_update_dpp(0u, v, 0x140, 0xF, 0xF, false);  // row_mirror: ^15
        return __builtin_amdgcn_update_dpp(0u, m, 0x141, 0xF, 0xF, false);            // row_half_mirror: ^7
    } else if constexpr (J == 16) {
        const auto sw = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (__lane_id() & 16) ? sw[0] : sw[1];
    } else {
        static_assert(J == 32, "lane_xor: J is a power of two below 64");
        const auto sw = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return (__lane_id() & 32) ? sw[0] : sw[1];
    }
}

// lane_xor for a distance that is a constant after unrolling.
__device__ __forceinline__ uint32_t lane_xor_c(uint32_t v, int j) {
    switch (j) {
        case 1: return lane_xor<1>(v);
        case 2: return lane_xor<2>(v);
        case 4: return lane_xor<4>(v);
        case 8: return lane_xor<8>(v);
        case 16: return lane_xor<16>(v);
        default: return lane_xor<32>(v);
    }
}

// Argmax over the lanes of aligned groups of S lanes (S a power of two
// <= 64); every lane of a group ends with the group's winner.
template <int S>
__device__ __forceinline__ void group_argmax(double &bi, uint32_t &bl) {
#pragma unroll
    for (int off = 1; off < S; off <<= 1) {
        const double oi = __shfl_xor(bi, off, SCD_WAVE);
        const uint32_t ol = __shfl_xor(bl, off, SCD_WAVE);
        if (better(oi, ol, bi, bl)) {
            bi = oi;
            bl = ol;
        }
    }
}

// Bitonic sort (ascending) of one key per lane across aligned groups of S
// lanes; lig = lane index inside the group.
template <int S>
__device__ __forceinline__ uint32_t group_bitonic_sort(uint32_t v, int lig) {
#pragma unroll
    for (int k = 2; k <= S; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint32_t o = lane_xor_c(v, j);
            const bool up = (lig & k) == 0;
            const bool lower = (lig & j) == 0;
            v = (lower == up) ? min(v, o) : max(v, o);
        }
    }
    return v;
}

// Bitonic sort (ascending) of 64*R keys held by one wave in the blocked
// layout: key i lives in register r = i % R of lane i / R.  Partner
// distances below R swap registers inside a lane, longer ones exchange
// across lanes (lane_xor): log2(R) of the network's distance levels stay in
// registers, against none below 64 in the strided layout.
template <int R>
__device__ __forceinline__ void wave_bitonic_sort_blocked(uint32_t (&x)[R], int lane) {
    constexpr int N = 64 * R;
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j < R) {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int rr = r ^ j;
                    if (rr > r) {
                        const bool up = ((lane * R + r) & k) == 0;
                        const uint32_t a = x[r], b = x[rr];
                        const bool sw = (a > b) == up;
                        x[r] = sw ? b : a;
                        x[rr] = sw ? a : b;
                    }
                }
            } else {
                const int lj = j / R;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const uint32_t o = lane_xor_c(x[r], lj);
                    const bool up = ((lane * R + r) & k) == 0;
                    const bool lower = (lane & lj) == 0;
                    x[r] = (lower == up) ? min(x[r], o) : max(x[r], o);
                }
            }
        }
    }
}


// First index in sorted s[0..n) whose key is greater than x.
__device__ __forceinline__ uint32_t upper_bound_u32(const uint32_t *s, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// upper_bound_u32 for a key known to sit at s[i]: a galloping search from
// i (runs of equal neighbour labels are mostly short, so most calls read
// s[i + 1] only), then a binary search within the last step.
__device__ __forceinline__ uint32_t run_end_u32(const uint32_t *s, uint32_t n, uint32_t i, uint32_t x) {
    uint32_t lo = i + 1, hi = i + 1, step = 1;
    while (hi < n && s[hi] <= x) {  // everything below hi + 1 is <= x
        lo = hi + 1;
        hi += step;
        step <<= 1;
    }
    if (hi > n) hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Merge path: the number of elements taken from A among the first `diag`
// outputs of merge(A[0..la), B[0..lb)) (A first on ties).
__device__ __forceinline__ uint32_t merge_path(const uint32_t *A, uint32_t la, const uint32_t *B,
                                               uint32_t lb, uint32_t diag) {
    uint32_t lo = diag > lb ? diag - lb : 0;
    uint32_t hi = diag < la ? diag : la;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (A[mid] <= B[diag - 1 - mid]) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Chunks of 64 R keys at s[c * 64 R], c = w, w + 4, ..., sorted in
// registers by wave w; chunks past the n real keys are padding only.
template <int R>
__device__ __forceinline__ void wave_sort_chunks(uint32_t *s, uint32_t n2, uint32_t n, int lane, int w) {
    constexpr uint32_t C = 64u * R;
    for (uint32_t c = w; c * C < n2; c += 4) {
        if (c * C >= n) break;  // uniform per wave
        uint32_t x[R];
#pragma unroll
        for (int r = 0; r < R; r++) x[r] = s[c * C + lane * R + r];
        wave_bitonic_sort_blocked<R>(x, lane);
#pragma unroll
        for (int r = 0; r < R; r++) s[c * C + lane * R + r] = x[r];
    }
}

// Sort of n2 keys (a power of two, 1024 <= n2 <= T; n real keys, then
// padding) held in LDS by a block of 256 threads: the waves sort chunks of
// n2 / 4 (at most 1024) keys in registers, then merge-path passes in LDS
// double the sorted run width; each thread computes its n2/256 outputs of a
// pass in registers and writes them back after a barrier.  Ends with a
// barrier.
template <int T>
__device__ __forceinline__ void block_sort_lds(uint32_t *s, uint32_t n2, uint32_t n) {
    // positions n .. n2-1 hold padding (SCD_INVALID, the largest key): a
    // chunk of padding only is sorted already, and merging a run with a
    // run of padding only leaves both in place.  Positions below n hold the
    // real keys before and after every pass, so a thread whose outputs all
    // lie at n or above would write padding over padding: it skips.
    // Chunks of n2 / 4 keys (at most 1024), so all four waves sort in
    // registers
    constexpr int PER_MAX = T / 256;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t chunk = min(n2 / 4, 1024u);
    if (chunk <= 256) wave_sort_chunks<4>(s, n2, n, lane, w);
    else if (chunk == 512) wave_sort_chunks<8>(s, n2, n, lane, w);
    else wave_sort_chunks<16>(s, n2, n, lane, w);
    __syncthreads();
    const uint32_t per = n2 / 256;
    const uint32_t o0 = threadIdx.x * per;
    for (uint32_t w2 = chunk; w2 < n2; w2 *= 2) {
        const uint32_t ps = (o0 / (2 * w2)) * (2 * w2);
        const bool skip = ps + w2 >= n || o0 >= n;  // B is padding only, or so are the outputs
        const uint32_t *A = s + ps;
        const uint32_t *B = A + w2;
        uint32_t out[PER_MAX];
        if (!skip) {
            uint32_t i = merge_path(A, w2, B, w2, o0 - ps);
            uint32_t j = o0 - ps - i;
#pragma unroll
            for (int k = 0; k < PER_MAX; k++) {
                if ((uint32_t)k < per) {
                    const bool takeA = (j >= w2) || (i < w2 && A[i] <= B[j]);
                    out[k] = takeA ? A[i++] : B[j++];
                }
            }
        }
        __syncthreads();
        if (!skip) {
#pragma unroll
            for (int k = 0; k < PER_MAX; k++)
                if ((uint32_t)k < per) s[o0 + k] = out[k];
        }
        __syncthreads();
    }
}

// |A ∩ B| of two ascending, duplicate-free lists: walks the shorter one and
// finds each key in the longer by a lower-bound search whose start only
// moves forward.
__device__ __forceinline__ uint32_t count_common(const uint32_t *A, uint32_t a, const uint32_t *B,
                                                 uint32_t b) {
    if (a > b) {
        const uint32_t *T = A;
        A = B;
        B = T;
        const uint32_t t = a;
        a = b;
        b = t;
    }
    uint32_t c = 0, lo = 0;
    for (uint32_t i = 0; i < a && lo < b; i++) {
        const uint32_t x = A[i];
        uint32_t l = lo, h = b;
        while (l < h) {
            const uint32_t mid = (l + h) >> 1;
            if (B[mid] < x) l = mid + 1;
            else h = mid;
        }
        if (l < b && B[l] == x) {
            c++;
            l++;
        }
        lo = l;
    }
    return c;
}

// WCC term of one vertex, src/wcc.cpp:100-166 (alpha = 1.0):
// (int_tri * tri_deg) / (tri * (tri_deg + (size - 1 - int_tri_deg))).
__device__ __forceinline__ double wcc_term(int tri, int tri_deg, int size, int int_tri,
                                           int int_tri_deg) {
    const double nom = (double)wmul(int_tri, tri_deg);
    const double denom =
        (double)tri * ((double)tri_deg + 1.0 * (double)wsub(wsub(size, 1), int_tri_deg));
    return denom != 0.0 ? nom / denom : 0.0;
}

}  // namespace scd_dev

#endif
