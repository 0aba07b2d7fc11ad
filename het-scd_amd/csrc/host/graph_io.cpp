// Binary graph format: reader, writer and the clustering coefficient.
//
// Layout (reference src/graph.cpp:86-134, src/convert.cpp:417-463):
//   [u32 n][u32 m][u32 indices n+1][u32 adj 2m][i32 tri n][i32 tri_deg n]
// The reader accepts a file only when it ends exactly after tri_deg, which is
// the reference's EOF rule (src/graph.cpp:118-131).
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <sys/stat.h>

#include "host_common.hpp"
#include "scd_host.h"

namespace scd {

static thread_local char g_err[512];

void set_host_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

}  // namespace scd

using scd::set_host_error;

extern "C" const char *scd_host_last_error(void) { return scd::g_err; }

static scd_graph *graph_alloc(uint32_t n, uint32_t m) {
    scd_graph *g = (scd_graph *)calloc(1, sizeof(scd_graph));
    if (!g) return nullptr;
    g->n = n;
    g->m = m;
    g->indices = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    g->adj = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)2 * m + 1));
    g->tri = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
    g->tri_deg = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
    if (!g->indices || !g->adj || !g->tri || !g->tri_deg) {
        scd_graph_free(g);
        return nullptr;
    }
    return g;
}

extern "C" void scd_graph_free(scd_graph *g) {
    if (!g) return;
    free(g->indices);
    free(g->adj);
    free(g->tri);
    free(g->tri_deg);
    free(g);
}

extern "C" int scd_graph_from_arrays(uint32_t n, uint32_t m, const uint32_t *indices,
                                     const uint32_t *adj, const int32_t *tri,
                                     const int32_t *tri_deg, scd_graph **out) {
    if (!out || !indices || (m && !adj) || (n && (!tri || !tri_deg))) {
        set_host_error("scd_graph_from_arrays: null argument");
        return SCD_EINVAL;
    }
    if (m >= (1u << 31)) {
        set_host_error("scd_graph_from_arrays: m=%u needs 2m < 2^32", m);
        return SCD_ELIMIT;
    }
    scd_graph *g = graph_alloc(n, m);
    if (!g) {
        set_host_error("scd_graph_from_arrays: out of host memory");
        return SCD_ENOMEM;
    }
    memcpy(g->indices, indices, sizeof(uint32_t) * ((size_t)n + 1));
    if (m) memcpy(g->adj, adj, sizeof(uint32_t) * (size_t)2 * m);
    if (n) {
        memcpy(g->tri, tri, sizeof(int32_t) * n);
        memcpy(g->tri_deg, tri_deg, sizeof(int32_t) * n);
    }
    *out = g;
    return SCD_OK;
}

extern "C" int scd_graph_read(const char *path, scd_graph **out) {
    if (!path || !out) {
        set_host_error("scd_graph_read: null argument");
        return SCD_EINVAL;
    }
    FILE *f = fopen(path, "rb");
    if (!f) {
        set_host_error("failed to open %s", path);
        return SCD_EIO;
    }
    uint32_t hdr[2];
    if (fread(hdr, sizeof(uint32_t), 2, f) != 2) {
        fclose(f);
        set_host_error("error while reading %s", path);
        return SCD_EIO;
    }
    const uint32_t n = hdr[0], m = hdr[1];
    if (m >= (1u << 31) || n == UINT32_MAX) {
        fclose(f);
        set_host_error("error while reading %s: n=%u m=%u exceeds the format", path, n, m);
        return SCD_ELIMIT;
    }
    struct stat st;
    const uint64_t expect = 8ull + 4ull * ((uint64_t)n + 1) + 8ull * m + 8ull * n;
    if (fstat(fileno(f), &st) != 0 || (uint64_t)st.st_size != expect) {
        fclose(f);
        set_host_error("error while reading %s: size %lld, layout needs %llu bytes", path,
                       (long long)st.st_size, (unsigned long long)expect);
        return SCD_EIO;
    }
    scd_graph *g = graph_alloc(n, m);
    if (!g) {
        fclose(f);
        set_host_error("error while reading %s: out of host memory", path);
        return SCD_ENOMEM;
    }
    bool ok = fread(g->indices, sizeof(uint32_t), (size_t)n + 1, f) == (size_t)n + 1;
    ok = ok && fread(g->adj, sizeof(uint32_t), (size_t)2 * m, f) == (size_t)2 * m;
    ok = ok && fread(g->tri, sizeof(int32_t), n, f) == n;
    ok = ok && fread(g->tri_deg, sizeof(int32_t), n, f) == n;
    fclose(f);
    if (!ok) {
        scd_graph_free(g);
        set_host_error("error while reading %s", path);
        return SCD_EIO;
    }
    *out = g;
    return SCD_OK;
}

extern "C" int scd_graph_write(const scd_graph *g, const char *path) {
    if (!g || !path) {
        set_host_error("scd_graph_write: null argument");
        return SCD_EINVAL;
    }
    FILE *f = fopen(path, "wb");
    if (!f) {
        set_host_error("failed to open %s for writing", path);
        return SCD_EIO;
    }
    uint32_t hdr[2] = {g->n, g->m};
    bool ok = fwrite(hdr, sizeof(uint32_t), 2, f) == 2;
    ok = ok && fwrite(g->indices, sizeof(uint32_t), (size_t)g->n + 1, f) == (size_t)g->n + 1;
    ok = ok && fwrite(g->adj, sizeof(uint32_t), (size_t)2 * g->m, f) == (size_t)2 * g->m;
    ok = ok && fwrite(g->tri, sizeof(int32_t), g->n, f) == g->n;
    ok = ok && fwrite(g->tri_deg, sizeof(int32_t), g->n, f) == g->n;
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        set_host_error("an error occured while writing %s", path);
        return SCD_EIO;
    }
    return SCD_OK;
}

// Mean of tri[p] / (deg * (deg - 1.0)) over vertices of degree > 1, divided
// by n (reference src/graph.cpp:122-133).  The reference reduces this sum
// with OpenMP, whose bits depend on the thread count; the engine fixes the
// order to ascending vertex id, which equals the one-thread result.
extern "C" double scd_graph_clustering_coef(const scd_graph *g) {
    if (!g) return 0.0;
    double cc = 0.0;  // n == 0 gives 0.0 / 0 = NaN, as in the reference
    for (uint32_t p = 0; p < g->n; p++) {
        const uint32_t deg = g->indices[p + 1] - g->indices[p];
        if (deg > 1) cc += g->tri[p] / (deg * (deg - 1.0));
    }
    return cc / g->n;
}
