/*
 * scd_serve.h — a serving loop over the batched engine (no counterpart in
 * the reference, which runs one graph per process: src/main-gpu.cpp:30-113).
 *
 * Requests (one preprocessed graph each, in the layout of class Graph,
 * src/graph.hpp:23-33) arrive over time.  The server
 *   - admits them first-come first-served while the HBM footprint of the
 *     graphs it holds stays within a budget,
 *   - keeps every admitted graph resident on the device (CSR, current and
 *     best labels),
 *   - runs the initial partition and initial WCC of newly admitted graphs
 *     together in a preparation arena,
 *   - refines all admitted graphs together in one arena, one iteration at a
 *     time (continuous batching: graphs join between iterations, and leave
 *     as soon as their own stopping rule, src/main-gpu.cpp:71-96, ends them),
 *   - returns each graph's best labels, best WCC and iteration count, equal
 *     to what the single-graph path gives for that graph alone.
 *
 * Single-threaded: one call at a time per server.  Status codes as in
 * scd_hip.h; the message of the last failure in scd_server_last_error().
 */
#ifndef SCD_SERVE_H
#define SCD_SERVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct scd_server scd_server;

typedef struct scd_server_opts {
    int device;
    /* device bytes the resident graphs and arenas may use; 0: 60% of the
     * device memory free at creation */
    uint64_t hbm_budget;
    /* most graphs refined together (0: no limit) */
    uint32_t max_graphs;
    /* join waiting graphs once the oldest has waited this long, even while
     * the arena is busy (it always joins when the arena empties or half of
     * its graphs have finished) */
    double join_wait_ms;
    /* iteration cap per graph (0: 1000) */
    uint32_t max_iters;
    /* capture iterations as HIP graphs (1) or launch eagerly (0) */
    int graph_mode;
    /* an intake batch (graphs copied and prepared together before they can
     * join) holds at most this many graphs and footprint bytes (0: 64
     * graphs, 2 GiB); smaller batches start refining sooner, each batch costs
     * one preparation and one arena assembly */
    uint32_t intake_graphs;
    uint64_t intake_bytes;
} scd_server_opts;

typedef struct scd_request {
    uint32_t n, m;
    const uint32_t *indices; /* host arrays, valid until scd_server_run returns */
    const uint32_t *adj;
    const int32_t *tri;
    const int32_t *tri_deg;
    double clustering_coef;
    double arrival_ms; /* arrival time, ms after the start of scd_server_run */
} scd_request;

typedef struct scd_result {
    int status; /* SCD_OK, or the error that failed the request */
    uint32_t iterations;
    double initial_wcc, best_wcc;
    /* ms after the start of scd_server_run */
    double arrival_ms, admit_ms, start_ms, done_ms;
    uint32_t joined_arena_graphs; /* size of the arena it first refined in */
} scd_result;

typedef struct scd_server_stats {
    uint64_t iterations;   /* arena iterations */
    uint64_t joins;        /* arena assemblies */
    double busy_ms;        /* time spent in device work and assembly */
    uint64_t peak_bytes;   /* highest admitted footprint */
    uint32_t peak_graphs;  /* most graphs in one arena */
    /* where busy_ms went: uploads at admission, initial partitions and
     * WCC of joining graphs, arena assembly, iterations */
    double admit_ms, prep_ms, assemble_ms, step_ms;
} scd_server_stats;

const char *scd_server_last_error(void);
int scd_server_create(const scd_server_opts *opts, scd_server **out);
void scd_server_destroy(scd_server *srv);
/* Queues a request (its arrays are read when it is admitted). */
int scd_server_submit(scd_server *srv, const scd_request *req, uint64_t *id_out);
/* Serves every queued request to completion, replaying arrival times in
 * wall-clock time (sleeping while nothing has arrived). */
int scd_server_run(scd_server *srv);
/* Result of request id; labels_out (n entries, may be NULL): best labels. */
int scd_server_result(scd_server *srv, uint64_t id, uint32_t *labels_out, scd_result *out);
int scd_server_get_stats(const scd_server *srv, scd_server_stats *out);
/* Footprint (bytes) the server charges for a graph of n vertices, m edges. */
uint64_t scd_server_footprint(uint32_t n, uint32_t m);

#ifdef __cplusplus
}
#endif

#endif
