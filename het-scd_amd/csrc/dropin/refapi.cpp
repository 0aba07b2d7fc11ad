// libscd_refapi.so: the reference's CPU call surface (Graph::read, the
// scd.hpp and wcc.hpp functions) over the engine's C ABI, computed on the
// device.  One process-wide context holds the last graph seen.
#include <sys/time.h>

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <unordered_map>

#include "scd_hip.h"
#include "scd_host.h"
#include "scd_refapi.hpp"

namespace {

scd_ctx *g_ctx = nullptr;
// the graph resident in g_ctx: object, arrays and shape
const Graph *g_graph = nullptr;
const unsigned int *g_indices = nullptr, *g_adj = nullptr;
unsigned int g_n = 0, g_m = 0;

[[noreturn]] void die(const char *what, const char *msg) {
    std::cerr << what << ": " << msg << std::endl;
    exit(EXIT_FAILURE);
}

void must(int rc, const char *what) {
    if (rc != SCD_OK) die(what, scd_last_error());
}

double now_ms() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec * 1000.0 + tv.tv_usec / 1000.0;
}

// The reference functions' timer lines ("name: X ms", src/common.cpp:38-52).
struct Timer {
    const char *name;
    double t0;
    explicit Timer(const char *n) : name(n), t0(now_ms()) {}
    ~Timer() { std::cout << name << ": " << (now_ms() - t0) << " ms" << std::endl; }
};

scd_ctx *ctx_for(const Graph *g, const char *fn) {
    if (!g) die(fn, "null graph");
    if (!g_ctx) {
        const char *dev = getenv("SCD_DEVICE");
        must(scd_init(dev ? atoi(dev) : 0, 256, 0, &g_ctx), fn);
    }
    if (g != g_graph || g->indices != g_indices || g->adj != g_adj || g->n != g_n || g->m != g_m) {
        must(scd_upload(g_ctx, g->n, g->m, g->indices, g->adj, g->triangles, g->triangle_deg, g->clustering_coef),
             fn);
        g_graph = g;
        g_indices = g->indices;
        g_adj = g->adj;
        g_n = g->n;
        g_m = g->m;
    }
    return g_ctx;
}

void supported(const char *fn, double alpha, const bool *mask) {
    if (mask) die(fn, "a vertex mask (the partitioned path of src/dist.cu) is not supported");
    if (alpha != 1.0) die(fn, "alpha must be 1.0 (the value every reference driver passes)");
}

}  // namespace

const Graph *Graph::read(const char *filename) {
    scd_graph *hg = nullptr;
    if (scd_graph_read(filename, &hg) != SCD_OK) {
        std::cerr << scd_host_last_error() << std::endl;
        return nullptr;
    }
    // the arrays stay allocated for the life of the process, as the
    // reference's graph does
    return new Graph(hg->n, hg->m, scd_graph_clustering_coef(hg), hg->tri_deg, hg->tri, hg->indices, hg->adj);
}

Partition *create_initial_partition(const Graph *g) {
    Timer t("create initial partition");
    scd_ctx *c = ctx_for(g, "create_initial_partition");
    Partition *p = new Partition;
    p->labels = new unsigned int[g->n];
    p->sizes = new int[g->n];
    p->ext_edges = new int[g->n];
    p->int_edges = new int[g->n];
    uint32_t k = 0;
    must(scd_initial_partition(c, p->labels, &k), "create_initial_partition");
    p->num_communities = k;
    {
        // the statistics the reference collects at the end (src/scd.cpp:168)
        Timer ts("collect community statistics");
        must(scd_get_stats(c, nullptr, p->sizes, p->ext_edges, p->int_edges), "create_initial_partition");
    }
    return p;
}

void collect_stats(const Graph *g, Partition *p, const bool *mask) {
    supported("collect_stats", 1.0, mask);
    Timer t("collect community statistics");
    scd_ctx *c = ctx_for(g, "collect_stats");
    must(scd_set_partition(c, p->labels, p->num_communities, nullptr, nullptr, nullptr), "collect_stats");
    must(scd_get_stats(c, nullptr, p->sizes, p->ext_edges, p->int_edges), "collect_stats");
}

void improve_partition(const Graph *g, Partition *p, double alpha, const bool *mask) {
    supported("improve_partition", alpha, mask);
    Timer t("improving partition");
    scd_ctx *c = ctx_for(g, "improve_partition");
    must(scd_set_partition(c, p->labels, p->num_communities, p->sizes, p->ext_edges, p->int_edges),
         "improve_partition");
    unsigned int *next = new unsigned int[g->n];
    must(scd_move_step(c, next), "improve_partition");
    delete[] p->labels;
    p->labels = next;
}

double compute_wcc(const Graph *g, double alpha, unsigned int *labels, int *sizes, const bool *mask) {
    supported("compute_wcc", alpha, mask);
    Timer t("compute WCC");
    scd_ctx *c = ctx_for(g, "compute_wcc");
    // the communities the labels name: sizes[] covers each of them
    unsigned int num = 0;
    for (unsigned int v = 0; v < g->n; v++)
        if (labels[v] != INVALID_LABEL && labels[v] + 1 > num) num = labels[v] + 1;
    double wcc = 0.0;
    must(scd_compute_wcc_of(c, labels, num, sizes, &wcc), "compute_wcc");
    return wcc;
}

// src/scd.cpp:32-55 as it behaves: a label seen for the first time maps to
// 0 (the map's default value, inserted by the lookup) and every later
// occurrence of it, like every INVALID label, takes the next number; the
// count is the number of numbers handed out.
void compress_labels(const Graph *g, unsigned int *labels, unsigned int *num_communities) {
    Timer t("compress labels");
    std::unordered_map<unsigned int, unsigned int> seen;
    seen.reserve(g->n);
    unsigned int next = 0;
    for (unsigned int v = 0; v < g->n; v++) {
        const unsigned int l = labels[v];
        if (l == INVALID_LABEL) {
            labels[v] = next++;
            continue;
        }
        auto it = seen.find(l);
        if (it == seen.end()) {
            seen.emplace(l, 0u);
            labels[v] = 0;
        } else {
            labels[v] = it->second = next++;
        }
    }
    *num_communities = next;
}
