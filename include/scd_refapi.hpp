/*
 * scd_refapi.hpp — the reference's CPU call surface with its names, C++
 * signatures and data structures, computed on the MI355X:
 *   Graph::read                         src/graph.hpp:41, src/graph.cpp:86-134
 *   Partition, INVALID_LABEL            src/scd.hpp:23-31
 *   create_initial_partition            src/scd.hpp:33, src/scd.cpp:113-171
 *   compress_labels                     src/scd.hpp:34, src/scd.cpp:32-55
 *   collect_stats                       src/scd.hpp:35, src/scd.cpp:57-111
 *   improve_partition                   src/scd.hpp:36, src/scd.cpp:221-341
 *   compute_wcc                         src/wcc.hpp:24-30, src/wcc.cpp:31-141
 * implemented by libscd_refapi.so over the C ABI of scd_hip.h.  A program
 * written against the reference's scd.hpp / wcc.hpp / graph.hpp (its
 * src/main.cpp, or src/main-gpu.cpp's initial partition and initial WCC)
 * links against the library unchanged; results equal the reference's CPU
 * functions' (labels and statistics bit for bit, WCC within 1e-9).
 *
 * Behaviour kept from the reference: the Partition's arrays are new[]-ed
 * with n entries; improve_partition replaces p->labels by a new array and
 * deletes the old one (src/scd.cpp:338-339); compress_labels renumbers the
 * labels in place exactly as src/scd.cpp:32-55 does (the printed count is
 * n minus the number of distinct labels); each function prints its timer
 * line ("name: X ms") as the reference's do.
 *
 * Differences: the computations run on device SCD_DEVICE (environment,
 * default 0) in one process-wide context, which keeps the last graph it
 * saw resident and uploads a graph when the Graph object, its arrays or its
 * size change; a vertex mask (the reference's partitioned path,
 * src/dist.cu) and alpha other than 1.0 (the reference always passes 1.0)
 * end the process with a message, as a device error does.
 */
#ifndef SCD_REFAPI_HPP
#define SCD_REFAPI_HPP

#include <cstddef>

#include "scd_graph_layout.hpp"

#ifndef INVALID_LABEL
#define INVALID_LABEL (~((unsigned int)0))
#endif

typedef struct Partition {
    unsigned int *labels;
    unsigned int num_communities;
    int *sizes;
    int *ext_edges;
    int *int_edges;
} Partition;

Partition *create_initial_partition(const Graph *g);
void compress_labels(const Graph *g, unsigned int *labels, unsigned int *num_communities);
void collect_stats(const Graph *g, Partition *p, const bool *mask = NULL);
void improve_partition(const Graph *g, Partition *p, double alpha, const bool *mask = NULL);

inline static void compress_labels(const Graph *g, Partition *p) { compress_labels(g, p->labels, &(p->num_communities)); }

double compute_wcc(const Graph *g, double alpha, unsigned int *labels, int *sizes, const bool *mask = NULL);

inline static double compute_wcc(const Graph *g, Partition *p, double alpha, const bool *mask = NULL) {
    return compute_wcc(g, alpha, p->labels, p->sizes, mask);
}

#endif
