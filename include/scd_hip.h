/*
 * scd_hip.h — C ABI of the MI355X SCD refinement engine (libscd_hip.so).
 *
 * This is the drop-in boundary for the reference's single-GPU path.  The
 * reference binds its GPU runtime through six C++ functions with file-static
 * device state (src/cuda.hpp:23-28, implemented in src/cuda.cu:55-369) and
 * calls them from src/main-gpu.cpp:40-103.  Here the same steps take an
 * explicit context handle so one process can hold several graphs (batched and
 * served modes), and every call returns an int status instead of exiting:
 *   SCD_OK (0) on success, a negative SCD_E* code (scd_host.h) on failure,
 *   with a message for the calling thread in scd_last_error().
 * The reference's exact C++ names and exit-on-error behaviour are provided by
 * a thin layer over this ABI (scd_dropin.hpp / libscd_dropin.so).
 *
 * Types are plain C: uint32 vertex ids and labels, int32 triangle and
 * community counters, fp64 WCC.  INVALID label = 0xFFFFFFFF (src/scd.hpp:23).
 * A context owns all device memory it allocates and works on one HIP stream;
 * every call below returns after the device work it issued has completed.
 * Calls on one context must not overlap (the reference is single-threaded
 * over global state, src/cuda.cu:29-52); different contexts are independent.
 */
#ifndef SCD_HIP_H
#define SCD_HIP_H

#include <stdint.h>

#include "scd_host.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef SCD_INVALID_LABEL
#define SCD_INVALID_LABEL 0xFFFFFFFFu
#endif

typedef struct scd_ctx scd_ctx;

/* Message of the last failing call on this thread. */
const char *scd_last_error(void);

/* Number of visible HIP devices (0 when none), or a negative code. */
int scd_device_count(void);

/*
 * Replaces cuda_init(device_id, work_group_size, profile_flag)
 * (src/cuda.hpp:24, src/cuda.cu:55-82).  An ordinal outside [0, count)
 * fails with SCD_EDEVICE, as the reference returns false after listing the
 * devices.  work_group_size is the reference's "group size" argument
 * (src/main-gpu.cpp:38), the block size of its kernels
 * (src/cuda-common.cuh:38-44); any value in [1, 1024] is accepted, as a CUDA
 * launch would.  Here it sets the workgroup of the element-wise per-vertex
 * kernels (WCC terms, internal out-degrees, moved-community marks); the
 * sort, reduce-by-key and intersection kernels keep the workgroup shapes
 * their tiling is built for (64-lane waves, 256- and 512-thread groups).
 * profile != 0 records a HIP event pair around every kernel launch
 * (scd_prof_*) and launches iterations eagerly on one stream.
 */
int scd_init(int device_id, int work_group_size, int profile, scd_ctx **out);
void scd_destroy(scd_ctx *ctx);

/*
 * Replaces cuda_transfer_graph(g, alpha) (src/cuda.hpp:25, src/cuda.cu:84-148):
 * copies the CSR (indices n+1, adj 2m, tri n, tri_deg n) to the device and
 * sizes every buffer of the refinement loop.  clustering_coef is the graph's
 * mean local clustering coefficient (Graph::clustering_coef,
 * src/graph.cpp:57-68; scd_graph_clustering_coef computes it).  Adjacency
 * lists must be sorted ascending without self loops or duplicates, which the
 * binary format guarantees.  Replaces any graph uploaded before.  n == 0 is
 * rejected (the reference would loop forever, src/main.cpp:77), and so are
 * offsets that do not run from 0 to 2m without decreasing and adjacency
 * entries that are not vertex ids (>= n): SCD_EINVAL, checked before any
 * kernel indexes by them.
 */
int scd_upload(scd_ctx *ctx, uint32_t n, uint32_t m, const uint32_t *indices,
               const uint32_t *adj, const int32_t *tri, const int32_t *tri_deg,
               double clustering_coef);

/*
 * Per-vertex triangle statistics of the uploaded CSR on the device, with the
 * convert tool's definition (src/convert.cpp:271-340): t(e) = number of
 * triangles through edge e, tri[v] = sum of t over v's edges (2 x triangles
 * at v), tri_deg[v] = number of v's edges with t > 0.  Either output may be
 * NULL.  With replace != 0 the device copies of tri/tri_deg used by the WCC
 * are overwritten with the computed ones.
 */
int scd_count_triangles(scd_ctx *ctx, int32_t *tri_out, int32_t *tri_deg_out, int replace);

/*
 * scd_count_triangles without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  tri_out and
 * tri_deg_out (n entries each, either may be NULL) are DEVICE pointers of
 * the context's device; the pass's temporaries (out-degrees, triangle
 * counts, one flag per adjacency entry) live in a static device array of
 * 64 MiB that each GPU the library runs on holds until the process ends.
 * Graphs needing more than 8n + 2m (+ 4n without tri_deg_out) bytes of it
 * are rejected with SCD_ELIMIT; profiling must be off (SCD_ESTATE).  Every
 * call on one GPU shares that GPU's array, so calls that overlap on one GPU
 * -- from any context -- race.  Results equal scd_count_triangles'.
 */
int scd_count_triangles_noalloc(scd_ctx *ctx, int32_t *tri_out, int32_t *tri_deg_out);

/*
 * Work split of the triangle kernels (no counterpart in the reference; the
 * results never depend on it).  Edges are oriented toward the lower id and
 * a triangle w < v < u is found on the edge u -> v, by default by walking
 * v's out-list against the set of u's.  With a threshold h != 0, an edge
 * whose v has the longer out-list, of more than h and at most 2048 entries,
 * is walked the other way: u's out-list below v against the set of v's, so
 * each edge costs the shorter walk.  tc_heavy applies to scd_count_triangles
 * and scd_convert_edges (default 128), wcc_heavy to the WCC's internal
 * triangles (default 0: off).  Each is 0 or in [64, 2048).
 */
int scd_set_triangle_split(scd_ctx *ctx, uint32_t tc_heavy, uint32_t wcc_heavy);

/*
 * The convert tool (src/convert.cpp:466-537; scd_convert in scd_host.h is
 * its host restatement) on the device: an edge list (u[i], v[i]) in file
 * order becomes the binary graph layout -- ids by first appearance, loops
 * and duplicate edges dropped, triangles counted, vertices renumbered by
 * (triangle edges, id) descending, triangle-free edges and vertices without
 * a triangle edge dropped -- byte-identical to scd_convert.  The result
 * replaces the context's graph (as scd_upload would), with its clustering
 * coefficient.  u and v hold count entries each, in host memory
 * (edges_on_device 0) or device memory of the context's device (1).
 * n_out, m_out, cc_out (each may be NULL) receive the result's shape and
 * clustering coefficient.  A result without vertices leaves no graph in the
 * context and returns SCD_OK with n = m = 0 (cc NaN).  count must stay
 * below 2^31 - 2^16 (SCD_ELIMIT).
 */
int scd_convert_edges(scd_ctx *ctx, uint64_t count, const uint32_t *u, const uint32_t *v, int edges_on_device,
                      uint32_t *n_out, uint32_t *m_out, double *cc_out);

/* Counts and stage times of the context's last scd_convert_edges, in the
 * terms of the reference convert's stdout (src/convert.cpp:478-530). */
typedef struct scd_convert_info {
    uint64_t edges_read;     /* edges given ("found ... edges") */
    uint32_t id_bound;       /* largest id + 1 in uint32 ("found ... nodes") */
    uint32_t vertices;       /* vertices after reindexing by first appearance */
    uint64_t edges_simple;   /* edges after loops and duplicates are dropped */
    uint32_t edges_final;    /* edges in a triangle: the result's m */
    uint32_t vertices_final; /* vertices with a triangle edge: the result's n */
    /* wall time of each step on the device (host syncs at the step ends);
     * reindex includes the copy of host edges */
    double ms_reindex, ms_sort, ms_dedup, ms_triangles, ms_relabel;
} scd_convert_info;
int scd_convert_last_info(const scd_ctx *ctx, scd_convert_info *out);

/* Copies the context's graph to host arrays of n + 1, 2m, n and n entries
 * (each may be NULL): the layout of class Graph, src/graph.hpp:23-33. */
int scd_download_graph(scd_ctx *ctx, uint32_t *indices, uint32_t *adj, int32_t *tri, int32_t *tri_deg);

/* Shape and clustering coefficient of the context's graph (0 without one). */
int scd_graph_shape(const scd_ctx *ctx, uint32_t *n, uint32_t *m, double *clustering_coef);

/*
 * Clustering-coefficient-ordered initial partition on the device
 * (create_initial_partition, src/scd.cpp:113-171): vertices ordered by
 * (tri/(deg*(deg-1.0)), deg, id) descending; in that order an unlabelled
 * vertex opens the next label and claims its unlabelled neighbours.  The
 * labels become the device's current and best labels and the community
 * statistics are collected, like create_initial_partition followed by
 * cuda_labels_to_device.  Either output may be NULL.  Requires every degree
 * >= 2 (otherwise the reference's sort key is NaN and its order undefined).
 */
int scd_initial_partition(scd_ctx *ctx, uint32_t *labels_out, uint32_t *num_labels_out);

/*
 * Replaces cuda_labels_to_device(labels) (src/cuda.hpp:26, src/cuda.cu:150-174):
 * n labels, none INVALID; num_comms = max + 1; sets current and best labels
 * and collects the community statistics.
 */
int scd_labels_to_device(scd_ctx *ctx, const uint32_t *labels);

/*
 * Replaces cuda_improve_labels(&wcc) (src/cuda.hpp:27, src/cuda.cu:176-361):
 * one synchronous refinement iteration — best move for every vertex
 * (src/scd.cpp:221-341), community statistics (src/scd.cpp:57-111) and the
 * mean WCC (src/wcc.cpp:31-141) — with the CPU path's arithmetic.  Writes
 * the WCC to *wcc; labels_out (n entries, may be NULL) receives the labels
 * after the move.
 */
int scd_improve(scd_ctx *ctx, double *wcc, uint32_t *labels_out);

/*
 * A caller's partition, for the reference's CPU call surface (src/scd.hpp:
 * 25-40, src/wcc.hpp:24-30), where labels may be INVALID and the caller owns
 * the statistics:
 *
 * scd_set_partition: n labels (INVALID or below num_comms <= n) become the
 * current and best labels; with sizes, ext_edges and int_edges (num_comms
 * entries each, all three or none) those are the community statistics the
 * next move step reads, as improve_partition reads the Partition's arrays
 * (src/scd.cpp:280-287); without them they are collected (collect_stats,
 * src/scd.cpp:57-111; read them with scd_get_stats).
 *
 * scd_move_step: one synchronous best-move step from the current labels and
 * statistics (improve_partition, src/scd.cpp:221-341); the new labels become
 * current and go to labels_out (n entries, may be NULL).  The statistics
 * are not updated (the reference's collect_stats follows as its own call).
 *
 * scd_compute_wcc_of: compute_wcc(g, alpha = 1.0, labels, sizes)
 * (src/wcc.cpp:31-141) for n labels with num_comms communities; sizes
 * (num_comms entries, may be NULL: the sizes of the labels) is read for each
 * vertex's community size as the reference reads it.  The labels stay
 * current.
 */
int scd_set_partition(scd_ctx *ctx, const uint32_t *labels, uint32_t num_comms, const int32_t *sizes,
                      const int32_t *ext_edges, const int32_t *int_edges);
int scd_move_step(scd_ctx *ctx, uint32_t *labels_out);
int scd_compute_wcc_of(scd_ctx *ctx, const uint32_t *labels, uint32_t num_comms, const int32_t *sizes,
                       double *wcc);

/* Iteration execution mode: on (default) captures each iteration's launch
 * sequence once into a hipGraph (one per label-buffer parity) and replays
 * it; off launches every kernel from the host.  Results are identical.
 * While profiling is on, iterations launch eagerly (per-kernel HIP events
 * need stream-recorded events). */
int scd_set_graph_mode(scd_ctx *ctx, int on);
/* 1 if the next scd_improve / scd_batch_step of this context captures its
 * iteration into a HIP graph (graph mode, not captured yet for the current
 * label buffers and WCC mode), else 0.  No counterpart in the reference: for
 * callers that run other device work on other threads and must keep it out
 * of a capture. */
int scd_next_step_captures(const scd_ctx *ctx);

/* Captures (and uploads to the device) the hipGraphs of the next
 * iterations of the current labels -- the next one, then both label-buffer
 * parities with the incremental WCC -- so that the first scd_improve calls
 * only launch them.  No counterpart in the reference; the drop-in library
 * calls it in cuda_labels_to_device, before the driver's timed loop.  No
 * effect in eager mode or while profiling. */
int scd_capture_iterations(scd_ctx *ctx);

/* Turns per-kernel HIP event timing (scd_prof_*) on or off. */
int scd_set_profiling(scd_ctx *ctx, int on);

/*
 * Phase timing of each scd_improve, as cuda_improve_labels prints it
 * (src/cuda.cu:347-360): with on != 0, one-thread kernels write the device
 * wall clock at the phase boundaries of every iteration -- inside captured
 * iterations too, so iterations keep their graph replay and concurrent
 * lanes -- and scd_last_phase_ms gives the last iteration's phases in the
 * reference's order: best move, community statistics, removal of external
 * edges (the internal adjacency), WCC.  Milliseconds of wall time between
 * boundaries; the four add up to the iteration's device time.
 */
int scd_set_phase_timing(scd_ctx *ctx, int on);
int scd_last_phase_ms(const scd_ctx *ctx, double ms[4]);

/* Replaces cuda_store_labels() (src/cuda.cu:363-365): best := current. */
int scd_store_best(scd_ctx *ctx);

/* Replaces cuda_labels_from_device(labels) (src/cuda.cu:367-369): copies the
 * best labels (n entries) to the host. */
int scd_labels_from_device(scd_ctx *ctx, uint32_t *labels);

/* The current (not best) labels, n entries. */
int scd_current_labels(scd_ctx *ctx, uint32_t *labels);

/*
 * The refinement loop of src/main-gpu.cpp:71-96 (== src/main.cpp:57-84) run
 * inside the library: improve; store best when wcc > best; stop when
 * wcc < best * (1 + 0.01).  initial_wcc is the WCC of the current labels
 * (scd_compute_wcc).  Outputs (any may be NULL): the final best WCC (the
 * value the reference prints last), the iteration count, and the WCC of each
 * iteration (wcc_hist holds max_iters entries).
 */
int scd_refine(scd_ctx *ctx, double initial_wcc, uint32_t max_iters, double *best_wcc,
               uint32_t *iterations, double *wcc_hist);

/*
 * WCC of the current labels and statistics without moving any vertex
 * (compute_wcc, src/wcc.cpp:31-141).  terms_out (n entries, may be NULL)
 * receives each vertex's term; int_tri_out / int_tri_deg_out (may be NULL)
 * its internal triangle count and internal triangle degree.
 */
int scd_compute_wcc(scd_ctx *ctx, double *wcc, double *terms_out, int32_t *int_tri_out,
                    int32_t *int_tri_deg_out);

/* Community statistics of the current labels (Partition::sizes, ext_edges,
 * int_edges, src/scd.hpp:25-31), num_comms entries each; *num_comms may be
 * NULL. Arrays must hold the graph's n entries. */
int scd_get_stats(scd_ctx *ctx, uint32_t *num_comms, int32_t *sizes, int32_t *ext_edges,
                  int32_t *int_edges);

/*
 * The segmented sort + reduce-by-key of the move step on its own
 * (src/scd.cpp:255-278; the reference GPU's set_adj_labels,
 * SegSortKeysFromIndices and ReduceByKey, src/cuda.cu:199-227): for the
 * current labels, each vertex's neighbour labels sorted ascending (unsigned;
 * INVALID last) into sorted_out (2m entries, may be NULL), and the runs of
 * valid labels as (label, count) pairs, vertex-major, ascending label, with
 * run_offsets (n + 1 entries).  run_label/run_count must hold 2m entries.
 * *num_pairs receives the total number of runs.
 */
int scd_neighbor_label_pairs(scd_ctx *ctx, uint32_t *sorted_out, uint32_t *run_offsets,
                             uint32_t *run_label, uint32_t *run_count, uint64_t *num_pairs);

/*
 * Kernel shape knobs (degree-bin thresholds and tile sizes).  Results never
 * depend on them; scd_autotune picks them per graph-size class.
 *   lds_tile:    keys sorted in LDS per workgroup (2048, 4096, 8192 or
 *                16384); vertices with larger degree take the multi-tile
 *                hub path.
 *   merge_chunk: outputs per workgroup in the hub merge passes (power of
 *                two, 256..2*lds_tile).
 *   wave_max:    largest degree sorted by one wavefront in registers (256,
 *                512 or 1024; 0 means 1024); larger degrees up to lds_tile
 *                take one workgroup each.
 *   hub_table:   0: hub candidates by merge passes over sorted tiles; 1: by
 *                per-hub hash tables of (label, count) filled from sorted
 *                tiles (no merge passes); 2: by label-range buckets of about
 *                1,024 entries cut at sampled label values, each sorted and
 *                evaluated in LDS (no merge passes; a region of 16 KiB per
 *                bucket slot, about 32 bytes per hub adjacency entry).
 *   lds_hash:    0: the lds bin sorts each vertex's neighbour labels; 1: it
 *                counts them in an LDS hash table (lds_tile 2048 or 4096).
 *   Export (scd_neighbor_label_pairs) always sorts.
 *   lanes:       1: the per-bin kernels of a phase run concurrently on four
 *                streams (graphs of at least 2^20 adjacency entries); 0: one
 *                stream.  The environment variable SCD_LANES=0 (read by
 *                scd_init) keeps a context on one stream whatever the tuning.
 *   dense_wcc:   1: on graphs of at least 2^20 adjacency entries
 *                the WCC counts the internal triangles of dense communities
 *                (32..16384 members, internal degrees summing to at least
 *                s (s - 1) / 16) by AND + popcount of per-member bitmap rows,
 *                the others by the sorted-list kernels (below 2^20 entries
 *                the extra launches cost more than they save); 2: the same
 *                on every graph; 0 (default): every community by the list
 *                kernels.  It pays on arenas of planted-partition graphs
 *                (communities of thousands of members at densities near
 *                0.15); the autotuner tries it on its fastest shape.
 *                Rows take a buffer of 2 bytes per adjacency entry (1 MiB ..
 *                4 GiB); communities beyond it use the list kernels.
 *   reserved:    0.
 */
typedef struct scd_tuning {
    uint32_t lds_tile;
    uint32_t merge_chunk;
    uint32_t wave_max;
    uint32_t hub_table;
    uint32_t lds_hash;
    uint32_t lanes;
    uint32_t dense_wcc;
    uint32_t reserved;
} scd_tuning;

int scd_get_tuning(const scd_ctx *ctx, scd_tuning *out);
int scd_set_tuning(scd_ctx *ctx, const scd_tuning *t);

/*
 * Autotuner (no counterpart in the reference): times one refinement
 * iteration from the current labels under each candidate tuning (the
 * candidates that change the shape for this graph's degrees, hub modes
 * included when hubs exist; reps timed
 * runs each, median), keeps the fastest and leaves the labels, statistics
 * and snapshot as they were.  The choice is cached per graph-size class
 * (device architecture and CU count, log2 of n, of the maximum and of the
 * mean degree) in memory and, when cache_path is given, in that text file
 * (one line per class), which later calls read instead of measuring.
 * chosen / iteration_ms (may be NULL): the tuning set and its time.
 */
int scd_autotune(scd_ctx *ctx, const char *cache_path, int reps, scd_tuning *chosen, double *iteration_ms);

/* Turns the side counters below on or off (off by default: they cost one
 * same-address atomic per wave). */
int scd_set_counting(scd_ctx *ctx, int on);

/*
 * Side counters of the last scd_improve with counting on (for
 * algorithmic-byte accounting):
 * [0] P, distinct (vertex, valid neighbour label) pairs evaluated;
 * [1] internal adjacency entries (same-label directed edges);
 * [2] sum over vertices of internal degree squared;
 * [3] vertices with degree above lds_tile (hub path).
 */
int scd_last_counters(scd_ctx *ctx, uint64_t out[4]);

/*
 * The dense-community WCC of the last internal triangle count (tuning
 * dense_wcc; no counterpart in the reference): out[0] = 256-byte row units
 * requested by the communities classified dense (those past the buffer's
 * capacity were counted by the list kernels), out[1] = vertices whose
 * counts came from bitmap rows.  Both 0 when dense_wcc is off.  Waits for
 * the context's stream.
 */
int scd_dense_wcc_info(scd_ctx *ctx, uint32_t out[2]);

/*
 * Per degree bin of the last move step: bin names, vertices, adjacency
 * entries and evaluated (vertex, label) pairs; returns the number of bins
 * (at most max_bins are written; any array may be NULL).
 */
int scd_bin_info(scd_ctx *ctx, int max_bins, char (*names)[16], uint64_t *vertices,
                 uint64_t *entries, uint64_t *pairs);

/* Saves the current labels and statistics on the device; scd_restore puts
 * them back as current and best labels (device-to-device copies only). */
int scd_snapshot(scd_ctx *ctx);
int scd_restore(scd_ctx *ctx);

/* Measurement helper: times `reps` launches of a plain streaming copy
 * kernel over `bytes` bytes (read + write), for calibrating hardware
 * traffic counters against a kernel of known traffic. */
int scd_calibration_copy(scd_ctx *ctx, uint64_t bytes, int reps, double *ms_per_copy);

/*
 * Profiling (scd_init with profile != 0): per kernel name, the number of
 * launches and their summed duration from HIP events recorded on the
 * context's stream around each launch.
 */
int scd_prof_count(scd_ctx *ctx);
int scd_prof_get(scd_ctx *ctx, int index, const char **name, uint64_t *launches,
                 double *total_ms);
int scd_prof_reset(scd_ctx *ctx);

/* Device bytes currently allocated by the context. */
uint64_t scd_device_bytes(const scd_ctx *ctx);

/*
 * The move step's arithmetic on caller rows, evaluated on the device by the
 * same device functions the move kernels inline (for bit-level tests of the
 * fp64 formula; none of them needs a graph).
 *
 * scd_eval_improvement: out[i] = calculate_improvement(r[i], d_in[i],
 * d_out[i], c_out[i], p_in[i], p_ext[i], alpha = 1.0) (src/scd.cpp:173-219):
 * the int32 products wrap and every fp64 operation rounds on its own.
 *
 * scd_eval_candidate: out[i] = the improvement of moving a vertex of degree
 * degree[i] and label my_label[i] into community label[i], seen freq[i]
 * times among its neighbours, where that community has size[i] members,
 * ext_edges[i] boundary and int_edges[i] internal (directed) edge ends, and
 * p_ext = clustering_coef: the candidate set-up of src/scd.cpp:280-307
 * (own community without the vertex, int_edges halved, p_in from the size)
 * followed by calculate_improvement.
 */
int scd_eval_improvement(scd_ctx *ctx, uint32_t count, const int32_t *r, const int32_t *d_in,
                         const int32_t *d_out, const int32_t *c_out, const double *p_in,
                         const double *p_ext, double *out);
int scd_eval_candidate(scd_ctx *ctx, uint32_t count, const int32_t *size, const int32_t *ext_edges,
                       const int32_t *int_edges, const uint32_t *label, const uint32_t *freq,
                       const uint32_t *degree, const uint32_t *my_label, double clustering_coef,
                       double *out);

/*
 * Batched arenas (no counterpart in the reference, which refines one graph
 * per process): G independent graphs refined together, every kernel launch
 * covering all of them.  The arena is the disjoint union of the graphs
 * (scd_graph_concat / scd_gen_planted_batch, scd_host.h): graph g owns
 * vertices [vtx_off[g], vtx_off[g+1]), its adjacency stays inside that range,
 * cc[g] is its clustering coefficient.  Each graph follows the loop of
 * src/main-gpu.cpp:71-96 with its own stopping rule; its labels, statistics,
 * per-iteration WCC (same summation order as a single-graph context) and
 * best labels equal the single-graph path's.  Labels at this API are local
 * (0-based per graph).  A list that leaves its graph's vertex range is
 * rejected with SCD_EINVAL.
 */
int scd_batch_upload(scd_ctx *ctx, uint32_t G, const uint64_t *vtx_off, const uint32_t *indices,
                     const uint32_t *adj, const int32_t *tri, const int32_t *tri_deg,
                     const double *cc);
/* Initial partition of every graph; k_out (G entries) the label counts. */
int scd_batch_initial_partition(scd_ctx *ctx, uint32_t *labels_out, uint32_t *k_out);
/* WCC of every graph's current labels (G entries). */
int scd_batch_compute_wcc(scd_ctx *ctx, double *wcc_out);
/*
 * Refinement loops of all graphs; stops when every graph has stopped or
 * after max_iters iterations.  Outputs (G entries each, may be NULL): best
 * WCC, iterations, milliseconds from the call's start to the end of the
 * graph's last iteration; wcc_hist (G x max_iters, may be NULL) the WCC of
 * each graph's iterations.
 */
int scd_batch_refine(scd_ctx *ctx, const double *initial_wcc, uint32_t max_iters, double *best_out,
                     uint32_t *iters_out, double *done_ms, double *wcc_hist);
/* Best labels (local) of every vertex of the arena. */
int scd_batch_labels(scd_ctx *ctx, uint32_t *labels_out);
/* Current labels (local). */
int scd_batch_current_labels(scd_ctx *ctx, uint32_t *labels_out);

/*
 * Arenas from graphs already on the device (no counterpart in the
 * reference; the building blocks of continuous batching, where graphs join
 * and leave an arena between iterations).  A part is one graph in the
 * reference's CSR layout (src/graph.hpp:23-33) with local ids, every array
 * a device pointer of the context's device.
 */
typedef struct scd_part {
    uint32_t n, m;
    const uint32_t *indices; /* n + 1 */
    const uint32_t *adj;     /* 2m, may be NULL when m == 0 */
    const int32_t *tri;      /* n */
    const int32_t *tri_deg;  /* n */
    double clustering_coef;
} scd_part;
/*
 * Replaces the context's graph by the disjoint union of parts[0..G).  With
 * cur_labels (G device pointers to local labels, n entries each) the arena
 * starts from those labels (best_labels likewise, NULL entries or a NULL
 * array: best = current), statistics recomputed, ready for scd_batch_step;
 * without, an initial partition must follow.  The assembly checks each part
 * as it copies it -- CSR offsets from 0 to 2m without decreasing, adjacency
 * entries below the part's n, labels below n or INVALID -- and fails with
 * SCD_EINVAL (no graph left in the context) when one is out of range.
 */
int scd_batch_upload_parts(scd_ctx *ctx, uint32_t G, const scd_part *parts,
                           const uint32_t *const *cur_labels, const uint32_t *const *best_labels);
/* Current / best local labels of each graph into device buffers (NULL
 * arrays or entries are skipped). */
int scd_batch_export_labels(scd_ctx *ctx, uint32_t *const *cur_out, uint32_t *const *best_out);
/* Best (best != 0) or current local labels of graph g into host memory. */
int scd_batch_graph_labels(scd_ctx *ctx, uint32_t g, int best, uint32_t *labels_out);
/*
 * Best (best != 0) or current local labels of graphs[0..count) into the host
 * arrays labels_out[i] (n of graphs[i] entries each), with one launch and
 * one device-to-host copy for all of them (a serving loop hands back every
 * graph that finished in an iteration this way).  No counterpart in the
 * reference, which refines one graph per process.
 */
int scd_batch_graphs_labels(scd_ctx *ctx, uint32_t count, const uint32_t *graphs, int best,
                            uint32_t *const *labels_out);
/*
 * One iteration (cuda_improve_labels, src/cuda.cu:176-351) of the graphs
 * whose active[g] != 0 (host array, G entries); the others keep their labels.
 * wcc_out (G entries, may be NULL): the WCC of each active graph.
 */
int scd_batch_step(scd_ctx *ctx, const uint32_t *active, double *wcc_out);
/* cuda_store_labels (src/cuda.cu:353-357) for the graphs with store[g] != 0. */
int scd_batch_store_best(scd_ctx *ctx, const uint32_t *store);

#ifdef __cplusplus
}
#endif

#endif
