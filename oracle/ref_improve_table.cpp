// ORACLE — test infrastructure only.
//
// Tables the reference's own calculate_improvement (an `inline static`
// function of src/scd.cpp:173-219, so it is reached by compiling that file
// into this translation unit from where it lies; oracle/Makefile passes the
// reference source directory as an include path).
// stdin:  lines "r d_in d_out c_out p_in_hex p_ext_hex"  (hex = u64 bits)
// stdout: lines "<result u64 bits as hex>"
#include <cinttypes>
#include <cstdio>
#include <cstring>

#include "scd.cpp"

int main() {
    int r, d_in, d_out, c_out;
    uint64_t pin_bits, pext_bits;
    while (scanf("%d %d %d %d %" SCNx64 " %" SCNx64, &r, &d_in, &d_out, &c_out, &pin_bits,
                 &pext_bits) == 6) {
        double p_in, p_ext;
        memcpy(&p_in, &pin_bits, 8);
        memcpy(&p_ext, &pext_bits, 8);
        const double v = calculate_improvement(r, d_in, d_out, c_out, p_in, p_ext, 1.0);
        uint64_t bits;
        memcpy(&bits, &v, 8);
        printf("%016" PRIx64 "\n", bits);
    }
    return 0;
}
