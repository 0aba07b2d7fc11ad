// Shared helpers for the host library: thread-local error text and the
// deterministic random number generator used by every graph generator.
#ifndef SCD_HOST_COMMON_HPP
#define SCD_HOST_COMMON_HPP

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

namespace scd {

// Message of the last failing call on this thread (scd_host_last_error).
void set_host_error(const char *fmt, ...);

// splitmix64 (Steele, Lea, Flood: "Fast splittable pseudorandom number
// generators", OOPSLA 2014).  Chosen because its output is fully specified,
// unlike std::*_distribution, so generated graphs are identical everywhere.
struct SplitMix64 {
    uint64_t state;
    explicit SplitMix64(uint64_t seed) : state(seed) {}
    uint64_t next() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // Uniform double in [0, 1) from the top 53 bits.
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

// The splitmix64 output function alone, used to derive independent streams.
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace scd

#endif
