// ORACLE — test infrastructure only.
//
// `oracle/main <graph>`: the CPU restatement driven like the reference's
// `./main` (src/main.cpp:29-94): same stdout lines, plus a machine-readable
// "ORACLE {json}" line with the timings the CPU baseline uses.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <omp.h>
#include <vector>

struct OrcRun {
    uint32_t n, iterations, num_communities;
    double cc, initial_wcc, final_wcc;
    double *wcc;
    uint32_t *labels, *best_labels;
    double refine_ms, init_ms;
};

extern "C" int orc_run(uint32_t n, const uint32_t *indices, const uint32_t *adj,
                       const int32_t *tri, const int32_t *tri_deg, uint32_t max_iters,
                       OrcRun *out);
extern "C" void orc_run_free(OrcRun *r);

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <filename>\n", argv[0]);
        return EXIT_FAILURE;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) {
        fprintf(stderr, "failed to open %s\n", argv[1]);
        return EXIT_FAILURE;
    }
    uint32_t hdr[2];
    if (fread(hdr, 4, 2, f) != 2) {
        fprintf(stderr, "error while reading %s\n", argv[1]);
        return EXIT_FAILURE;
    }
    const uint32_t n = hdr[0], m = hdr[1];
    std::vector<uint32_t> indices((size_t)n + 1), adj((size_t)2 * m);
    std::vector<int32_t> tri(n), tri_deg(n);
    bool ok = fread(indices.data(), 4, indices.size(), f) == indices.size();
    ok = ok && fread(adj.data(), 4, adj.size(), f) == adj.size();
    ok = ok && fread(tri.data(), 4, n, f) == n;
    ok = ok && fread(tri_deg.data(), 4, n, f) == n;
    ok = ok && fgetc(f) == EOF;
    fclose(f);
    if (!ok) {
        fprintf(stderr, "error while reading %s\n", argv[1]);
        return EXIT_FAILURE;
    }
    if (n == 0) {
        // the reference loops forever here (src/main.cpp:77 with wcc = 0)
        fprintf(stderr, "graph has no vertices\n");
        return EXIT_FAILURE;
    }
    OrcRun r;
    if (orc_run(n, indices.data(), adj.data(), tri.data(), tri_deg.data(), 1000, &r) != 0) {
        fprintf(stderr, "out of memory\n");
        return EXIT_FAILURE;
    }
    printf("\nnodes: %u\nedges: %u\navg. cc: %g\ninitial WCC: %g\nnum. threads: %d\n\n", n, m, r.cc,
           r.initial_wcc, omp_get_max_threads());
    for (uint32_t i = 0; i < r.iterations; i++) printf("WCC: %g\n\n", r.wcc[i]);
    printf("----------------------------------------\n");
    printf("num. communities: %u\nWCC: %g\niterations: %u\ntime per iteration: %g ms\n",
           r.num_communities, r.final_wcc, r.iterations, r.refine_ms / r.iterations);
    printf("----------------------------------------\n");
    printf("ORACLE {\"n\": %u, \"m\": %u, \"iterations\": %u, \"refine_ms\": %.6f, \"init_ms\": %.6f, "
           "\"threads\": %d, \"final_wcc\": %.17g}\n",
           n, m, r.iterations, r.refine_ms, r.init_ms, omp_get_max_threads(), r.final_wcc);
    orc_run_free(&r);
    return EXIT_SUCCESS;
}
