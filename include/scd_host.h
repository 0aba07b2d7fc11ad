/*
 * scd_host.h — host-side C ABI of the MI355X SCD engine: the binary graph
 * format, deterministic graph generators and the convert pipeline that turns
 * an edge list into the preprocessed binary CSR the refinement path consumes.
 *
 * Built into libscd_host.so.  Plain C types only, so it binds from C, C++,
 * ctypes or any other FFI.  Every call returns 0 on success or a negative
 * SCD_E* code; scd_host_last_error() holds a message for the last failure of
 * the calling thread.
 */
#ifndef SCD_HOST_H
#define SCD_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes shared by scd_host.h and scd_hip.h. */
#define SCD_OK 0
#define SCD_EINVAL (-1)   /* bad argument or malformed input                 */
#define SCD_EIO (-2)      /* file could not be opened, read or written       */
#define SCD_ENOMEM (-3)   /* host or device allocation failed                */
#define SCD_EDEVICE (-4)  /* HIP runtime error or no usable device           */
#define SCD_ELIMIT (-5)   /* size exceeds a documented limit                 */
#define SCD_ESTATE (-6)   /* call made in the wrong state (e.g. no graph)    */
#define SCD_EBUSY (-7)    /* admission refused: not enough device memory     */

const char *scd_host_last_error(void);

/*
 * Preprocessed undirected graph in the reference's binary layout
 * (Graph::read, src/graph.cpp:86-134):
 *   [u32 n][u32 m][u32 indices n+1][u32 adj 2m][i32 tri n][i32 tri_deg n]
 * adjacency lists are sorted ascending; tri[v] = 2 * #triangles at v and
 * tri_deg[v] = #incident edges that close a triangle.
 */
typedef struct scd_graph {
    uint32_t n;
    uint32_t m;          /* undirected edges; adj holds 2m entries */
    uint32_t *indices;   /* n + 1 */
    uint32_t *adj;       /* 2m */
    int32_t *tri;        /* n */
    int32_t *tri_deg;    /* n */
} scd_graph;

/* Edge list as produced by a generator: pairs (u[i], v[i]) in file order. */
typedef struct scd_edges {
    uint64_t count;
    uint32_t *u;
    uint32_t *v;
} scd_edges;

/* Reads a binary graph; rejects short files and trailing bytes like
 * Graph::read (src/graph.cpp:98-131). */
int scd_graph_read(const char *path, scd_graph **out);
/* Writes the binary layout (src/convert.cpp:417-463). */
int scd_graph_write(const scd_graph *g, const char *path);
/* Wraps caller arrays (copied) into a new graph object. */
int scd_graph_from_arrays(uint32_t n, uint32_t m, const uint32_t *indices,
                          const uint32_t *adj, const int32_t *tri,
                          const int32_t *tri_deg, scd_graph **out);
void scd_graph_free(scd_graph *g);
/* Mean local clustering coefficient, summed in vertex order
 * (Graph::Graph, src/graph.cpp:48-69; the sequential order fixes the bits). */
double scd_graph_clustering_coef(const scd_graph *g);

/* Planted partition: communities of the given sizes laid out contiguously;
 * every vertex pair is an edge with probability p_in inside a community and
 * p_out across.  Deterministic for a seed (splitmix64 + geometric skipping). */
int scd_gen_planted_partition(uint32_t k, const uint32_t *sizes, double p_in,
                              double p_out, uint64_t seed, scd_edges **out);
/* Planted partition with n log-uniform in [n_min, n_max], k uniform in
 * [k_min, k_max] and community sizes from a symmetric Dirichlet(1); writes
 * the drawn n and k when the pointers are non-NULL. */
int scd_gen_planted_random(uint64_t seed, uint32_t n_min, uint32_t n_max,
                           uint32_t k_min, uint32_t k_max, double p_in,
                           double p_out, uint32_t *n_out, uint32_t *k_out,
                           scd_edges **out);
/* R-MAT with 2^scale vertices and edge_factor * 2^scale edges. */
int scd_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b,
                 double c, uint64_t seed, scd_edges **out);
int scd_edges_from_arrays(uint64_t count, const uint32_t *u, const uint32_t *v,
                          scd_edges **out);
void scd_edges_free(scd_edges *e);
/* Writes "u v" lines, the text input format of the reference convert. */
int scd_edges_write_text(const scd_edges *e, const char *path);
/*
 * Reads a text edge list with the rules of the reference convert's reader
 * (src/convert.cpp:74-136): blank lines and lines starting with '#' (after
 * white space) are skipped, every other line starts with two decimal ids
 * separated by white space (the rest of the line is ignored).  Errors, the
 * first in line order: SCD_EINVAL "line <i> is corrupt" (1-based, every line
 * counted), SCD_EIO for a file that cannot be read or a line of 1024 or more
 * characters.  Edges keep file order; (a, b) as written.
 */
int scd_edges_read_text(const char *path, scd_edges **out);

/*
 * Convert pipeline with the reference's semantics (src/convert.cpp:466-537):
 * reindex by first appearance, drop loops and duplicates, count triangles per
 * edge, drop triangle-free edges, relabel by degree descending (ties: higher
 * id first), drop vertices left isolated.  Output is byte-identical to the
 * reference convert's file for the same edge sequence.
 */
int scd_convert(const scd_edges *e, scd_graph **out);

/*
 * Arena of several graphs as one disjoint union: graph i's vertices are
 * [vtx_off[i], vtx_off[i+1]) and its adjacency ids are offset by vtx_off[i]
 * (vtx_off holds count + 1 entries).  The batched engine entry points take
 * this layout.
 */
int scd_graph_concat(uint32_t count, const scd_graph *const *graphs, scd_graph **out,
                     uint64_t *vtx_off);

/*
 * `count` planted-partition graphs (scd_gen_planted_random with seeds
 * seed_base + i), each through scd_convert, generated in parallel and
 * concatenated into an arena.  A graph that convert empties is drawn again
 * with seed + count.  cc (count entries, may be NULL) receives each graph's
 * clustering coefficient, seeds (may be NULL) the seed used.
 */
int scd_gen_planted_batch(uint64_t seed_base, uint32_t count, uint32_t n_min, uint32_t n_max,
                          uint32_t k_min, uint32_t k_max, double p_in, double p_out,
                          scd_graph **out, uint64_t *vtx_off, double *cc, uint64_t *seeds);

/*
 * Community labels (no counterpart in the reference, which writes none).
 * scd_labels_printed_count: the count the reference prints after
 * compress_labels (src/scd.cpp:32-55, whose inverted map test makes it n
 * minus the number of distinct valid labels).
 * scd_labels_compact: the compaction compress_labels intends: distinct valid
 * labels numbered 0.. in order of first appearance, every INVALID vertex a
 * community of its own; *num_communities the number of communities.
 * Labels file: [u32 n][u32 num_communities][u32 compact labels n],
 * little-endian; scd_labels_write compacts first, scd_labels_read checks
 * the size and the range (free with scd_labels_free).
 */
#ifndef SCD_INVALID_LABEL
#define SCD_INVALID_LABEL 0xFFFFFFFFu /* a vertex without a community (src/scd.hpp:23) */
#endif
uint32_t scd_labels_printed_count(uint32_t n, const uint32_t *labels);
int scd_labels_compact(uint32_t n, const uint32_t *labels, uint32_t *out, uint32_t *num_communities);
int scd_labels_write(const char *path, uint32_t n, const uint32_t *labels);
int scd_labels_read(const char *path, uint32_t *n_out, uint32_t *k_out, uint32_t **labels_out);
void scd_labels_free(uint32_t *labels);

#ifdef __cplusplus
}
#endif

#endif
