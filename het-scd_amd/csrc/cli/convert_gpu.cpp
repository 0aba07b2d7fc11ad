// convert-gpu <input file> <output file> [device]
//
// The reference's convert tool (src/convert.cpp:466-537) with its work on
// the device: the same arguments (plus an optional device ordinal, default
// 0), the same stdout lines in the same order and the same output file
// bytes.  The text is parsed on the host with the reference reader's rules
// (scd_edges_read_text); reindexing, sorting, deduplication, triangle-free
// edge removal and the degree relabel run in one scd_convert_edges call,
// whose per-step device times fill the reference's step timer lines.
// Errors go to stderr with the reference's messages and exit code -1.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stack>
#include <string>
#include <sys/time.h>
#include <utility>
#include <vector>

#include "scd_hip.h"
#include "scd_host.h"

static double now_ms() {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec * 1000.0 + tv.tv_usec / 1000.0;
}

// "name: X ms" lines as src/common.cpp:35-52 prints them.
static std::stack<std::pair<std::string, double>> g_timers;
static void timer_start(const char *name) { g_timers.push({name, now_ms()}); }
static void timer_end() {
    if (g_timers.empty()) return;
    auto t = g_timers.top();
    g_timers.pop();
    std::cout << t.first << ": " << (now_ms() - t.second) << " ms" << std::endl;
}
static void timer_line(const char *name, double ms) { std::cout << name << ": " << ms << " ms" << std::endl; }

int main(int argc, char *argv[]) {
    if (argc < 3) {
        std::cerr << "usage: " << argv[0] << " <input file> <output file>" << std::endl;
        return -1;
    }
    const char *input_file = argv[1];
    const char *output_file = argv[2];
    const int device = argc > 3 ? atoi(argv[3]) : 0;
    scd_ctx *ctx = nullptr;
    if (scd_init(device, 256, 0, &ctx) != SCD_OK) {
        std::cerr << scd_last_error() << std::endl;
        return -1;
    }

    timer_start("read input");
    std::cout << "reading graph from file " << input_file << std::endl;
    scd_edges *e = nullptr;
    if (scd_edges_read_text(input_file, &e) != SCD_OK) {
        std::cerr << scd_host_last_error() << std::endl;
        exit(-1);
    }
    // num_nodes = max over lines of max(a, b) + 1, in unsigned arithmetic
    // (src/convert.cpp:123)
    unsigned int num_nodes = 0;
    for (uint64_t i = 0; i < e->count; i++) {
        const unsigned int a = e->u[i], b = e->v[i];
        num_nodes = std::max((a > b ? a : b) + 1u, num_nodes);
    }
    for (uint64_t k = 1000000; k <= e->count; k += 1000000) std::cout << k << " lines read so far" << std::endl;
    std::cout << "found " << num_nodes << " nodes and " << e->count << " edges" << std::endl;
    timer_end();

    timer_start("process graph");
    uint32_t n = 0, m = 0;
    double cc = 0.0;
    if (scd_convert_edges(ctx, e->count, e->u, e->v, 0, &n, &m, &cc) != SCD_OK) {
        std::cerr << scd_last_error() << std::endl;
        exit(-1);
    }
    scd_convert_info info;
    scd_convert_last_info(ctx, &info);
    const size_t old_nodes = num_nodes;
    timer_line("reindexing vertices", info.ms_reindex);
    std::cout << "removed " << (old_nodes - info.vertices) << " nodes" << std::endl;
    timer_line("sorting edges", info.ms_sort);
    timer_line("removing duplicate edges and loops", info.ms_dedup);
    std::cout << "removed " << (e->count - info.edges_simple) << " edges " << std::endl;
    timer_line("removing edges without triangles", info.ms_triangles);
    std::cout << "removed " << (info.edges_simple - info.edges_final) << " edges " << std::endl;
    timer_line("sort nodes according to degree", info.ms_relabel);
    std::cout << "removed " << ((size_t)info.vertices - info.vertices_final) << " nodes" << std::endl;
    timer_end();
    scd_edges_free(e);

    std::cout << "final graph has " << n << " nodes and " << m << " edges" << std::endl;
    std::cout << "writing final graph to " << std::string(output_file) << std::endl;
    timer_start("write output");
    std::vector<uint32_t> indices((size_t)n + 1, 0), adj(2ull * m);
    std::vector<int32_t> tri(n), tri_deg(n);
    if (n && scd_download_graph(ctx, indices.data(), adj.data(), tri.data(), tri_deg.data()) != SCD_OK) {
        std::cerr << scd_last_error() << std::endl;
        exit(-1);
    }
    scd_graph g{n, m, indices.data(), adj.data(), tri.data(), tri_deg.data()};
    if (scd_graph_write(&g, output_file) != SCD_OK) {
        std::cerr << "an error occured while writing output file" << std::endl;
        exit(-1);
    }
    timer_end();
    scd_destroy(ctx);
    return 0;
}
