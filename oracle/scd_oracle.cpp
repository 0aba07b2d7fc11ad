// ORACLE — test infrastructure only.
//
// A CPU restatement of the reference's OpenMP path (`./main`) for the SCD
// refinement pipeline.  It exists to check the HIP engine: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load or run it,
// and never as the thing measured or shipped.
//
// Each function cites the reference file:line it restates.  Arithmetic is
// kept in the reference's types and evaluation order: int32 products wrap
// (built with -fwrapv and written with explicit wrapping helpers), doubles
// are combined left to right with no contraction (-ffp-contract=off).
//
// Pinning (what checks each function besides the engine it is compared to):
//   orc_clustering_coef      golden cc hex from the reference build; K_n = 1
//   orc_initial_partition    golden labels; bowtie closed form; brute force
//   orc_collect_stats        golden stats; sum invariants
//   orc_improvement          golden table from the reference's own
//                            calculate_improvement, incl. int32-overflow rows
//   orc_candidate_improvement the candidate step of orc_improve: golden
//                            per-iteration labels; brute-force move rule
//   orc_improve              golden per-iteration labels; exhaustive T2 check
//   orc_compute_wcc          golden WCC (1e-9); bowtie = 0.5; brute-force
//                            internal triangle counts
//   orc_neighbor_label_pairs brute-force multiset counts
//   orc_triangles            the graph file's own tri/tri_deg; K_n closed form
//   orc_convert              golden bytes of the reference convert output
//   orc_run                  golden per-iteration labels and WCC
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <omp.h>
#include <sys/time.h>
#include <utility>
#include <vector>

#define INVALID_LABEL 0xFFFFFFFFu

static inline int wmul(int a, int b) { return (int)((uint32_t)a * (uint32_t)b); }
static inline int wadd(int a, int b) { return (int)((uint32_t)a + (uint32_t)b); }
static inline int wsub(int a, int b) { return (int)((uint32_t)a - (uint32_t)b); }

extern "C" {

// src/graph.cpp:122-133 — summed in vertex order (the one-thread result).
double orc_clustering_coef(uint32_t n, const uint32_t *indices, const int32_t *tri) {
    double cc = 0.0;
    for (uint32_t p = 0; p < n; p++) {
        unsigned int deg = indices[p + 1] - indices[p];
        if (deg > 1) cc += tri[p] / (deg * (deg - 1.0));
    }
    return cc / n;
}

// src/scd.cpp:57-111.  Only the first num_comms entries are cleared (:73-78).
void orc_collect_stats(uint32_t n, const uint32_t *indices, const uint32_t *adj,
                       const uint32_t *labels, uint32_t num_comms, int32_t *sizes,
                       int32_t *ext_edges, int32_t *int_edges) {
    for (uint32_t i = 0; i < num_comms; i++) {
        sizes[i] = 0;
        ext_edges[i] = 0;
        int_edges[i] = 0;
    }
#pragma omp parallel for schedule(dynamic, 512)
    for (int64_t pi = 0; pi < (int64_t)n; pi++) {
        const uint32_t p = (uint32_t)pi;
        const uint32_t label_p = labels[p];
        if (label_p == INVALID_LABEL) continue;
        int ext_count = 0, int_count = 0;
        for (uint32_t i = indices[p]; i != indices[p + 1]; i++) {
            if (label_p != labels[adj[i]]) ext_count++;
            else int_count++;
        }
        __sync_fetch_and_add(sizes + label_p, 1);
        if (ext_count > 0) __sync_fetch_and_add(ext_edges + label_p, ext_count);
        if (int_count > 0) __sync_fetch_and_add(int_edges + label_p, int_count);
    }
}

// src/scd.cpp:113-171 (without the trailing collect_stats, which the caller
// runs).  Returns the number of labels.
uint32_t orc_initial_partition(uint32_t n, const uint32_t *indices, const uint32_t *adj,
                               const int32_t *tri, uint32_t *labels) {
    std::fill(labels, labels + n, INVALID_LABEL);
    std::vector<std::pair<std::pair<double, int>, unsigned int>> order(n);
    for (uint32_t p = 0; p < n; p++) {
        unsigned int deg = indices[p + 1] - indices[p];
        double cc = tri[p] / (deg * (deg - 1.0));
        order[p] = std::make_pair(std::make_pair(cc, (int)deg), p);
    }
    std::sort(order.begin(), order.end(),
              std::greater<std::pair<std::pair<double, int>, unsigned int>>());
    uint32_t num_labels = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t p = order[i].second;
        if (labels[p] != INVALID_LABEL) continue;
        const uint32_t l = num_labels;
        for (uint32_t j = indices[p]; j != indices[p + 1]; j++)
            if (labels[adj[j]] == INVALID_LABEL) labels[adj[j]] = l;
        labels[p] = l;
        num_labels++;
    }
    return num_labels;
}

// src/scd.cpp:173-219, operation by operation.  Integer sub-expressions are
// 32-bit and wrap; every double operation is a separate IEEE step.
double orc_improvement(int r, int d_in, int d_out, int c_out, double p_in, double p_ext,
                       double alpha) {
    double t;
    if (r > 0) t = (double)wsub(c_out, d_in) / (double)r;
    else t = 0.0;

    const double din_din1 = (double)wmul(d_in, wsub(d_in, 1));
    const double dout_dout1 = (double)wmul(d_out, wsub(d_out, 1));
    const double dout_din = (double)wmul(d_out, d_in);
    const double rm1 = (double)wsub(r, 1);
    const double rm1rm2 = (double)wmul(wsub(r, 1), wsub(r, 2));
    const double two_din1 = (double)wmul(2, wsub(d_in, 1));

    double A = 0.0;
    double denom = (din_din1 * p_in + dout_dout1 * p_ext) + dout_din * p_ext;
    denom *= (double)wadd(d_in, d_out) + alpha * (double)wsub(wsub(r, 1), d_in);
    if (denom != 0.0) A = ((din_din1 * p_in) * (double)wadd(d_in, d_out)) / denom;

    double BMinus = 0.0;
    denom = rm1rm2 * p_in * p_in * p_in + two_din1 * p_in + t * rm1 * p_in * p_ext +
            t * (t - 1) * p_ext + (double)d_out * p_ext;
    denom *= rm1 * p_in + 1 + t + alpha * ((double)r - rm1 * p_in - 1);
    if (denom != 0.0) BMinus = (two_din1 * p_in) * (rm1 * p_in + 1 + t) / denom;

    double CMinus = 0.0;
    denom = rm1rm2 * p_in * p_in * p_in + t * (t - 1) * p_ext + t * rm1 * (p_in)*p_ext;
    denom *= rm1 * p_in + t + alpha * ((double)r - rm1 * p_in);
    denom *= rm1 * p_in + t + alpha * ((double)r - rm1 * p_in - 1);
    if (denom != 0.0 && (((double)r + t) > 0) && ((rm1 + t) > 0))
        CMinus = -(rm1rm2 * p_in * p_in * p_in) * (rm1 * p_in + t) * alpha / denom;

    return A + (double)d_in * BMinus + (double)wsub(r, d_in) * CMinus;
}

// src/scd.cpp:280-307: the improvement of moving a vertex of degree deg and
// label my_label into community label (size / ext / int edges of that
// community), seen freq times among its neighbours.
double orc_candidate_improvement(int size, int ext_edge, int int_edges, uint32_t label, uint32_t freq,
                                 uint32_t deg, uint32_t my_label, double clustering_coef, double alpha) {
    int int_edge = int_edges / 2;
    if (label == my_label) {
        size--;
        ext_edge = wadd(ext_edge, (int)(freq - (deg - freq)));
        int_edge = wsub(int_edge, (int)freq);
    }
    const int d_in = (int)freq;
    const int d_out = wsub((int)deg, d_in);
    const int c_out = ext_edge;
    double p_int = 0.0;
    if (size > 1) p_int = (2.0 * int_edge) / (size * (size - 1.0));
    return orc_improvement(size, d_in, d_out, c_out, p_int, clustering_coef, alpha);
}

// src/scd.cpp:221-341: synchronous best-move step, old labels -> new labels.
void orc_improve(uint32_t n, const uint32_t *indices, const uint32_t *adj,
                 const uint32_t *old_labels, const int32_t *sizes, const int32_t *ext_edges,
                 const int32_t *int_edges, double clustering_coef, double alpha,
                 uint32_t *new_labels) {
#pragma omp parallel
    {
        std::vector<uint32_t> buffer(4096);
#pragma omp for schedule(dynamic, 512)
        for (int64_t pi = 0; pi < (int64_t)n; pi++) {
            const uint32_t p = (uint32_t)pi;
            const uint32_t deg = indices[p + 1] - indices[p];
            const uint32_t my_label = old_labels[p];
            if (deg > buffer.size()) buffer.resize(deg);
            for (uint32_t i = 0; i < deg; i++) buffer[i] = old_labels[adj[indices[p] + i]];
            std::sort(buffer.begin(), buffer.begin() + deg);

            double remove_improvement = 0.0, insert_improvement = 0.0;
            uint32_t insert_label = 0;
            uint32_t index = 0;
            while (index < deg) {
                const uint32_t other_label = buffer[index];
                uint32_t freq = 1;
                if (other_label == INVALID_LABEL) {
                    index++;
                    continue;
                }
                while (index + freq < deg && buffer[index + freq] == other_label) freq++;

                const double improvement = orc_candidate_improvement(
                    sizes[other_label], ext_edges[other_label], int_edges[other_label], other_label, freq,
                    deg, my_label, clustering_coef, alpha);
                if (other_label == my_label) remove_improvement = -improvement;
                if (improvement > insert_improvement) {
                    insert_improvement = improvement;
                    insert_label = other_label;
                }
                index += freq;
            }
            uint32_t new_label;
            insert_improvement += remove_improvement;
            if (insert_improvement > remove_improvement && insert_improvement > 0.0)
                new_label = insert_label;
            else if (remove_improvement > 0)
                new_label = INVALID_LABEL;
            else
                new_label = my_label;
            new_labels[p] = new_label;
        }
    }
}

// src/wcc.cpp:31-141.  Per-vertex terms are kept (terms_out, optional) and
// summed in vertex order; the reference's OpenMP reduction sums the same
// terms in a thread-dependent order, so totals agree to rounding.
double orc_compute_wcc(uint32_t n, const uint32_t *indices, const uint32_t *adj,
                       const int32_t *triangles, const int32_t *triangle_deg,
                       const uint32_t *labels, const int32_t *sizes, double alpha,
                       double *terms_out, int32_t *int_tri_out, int32_t *int_tri_deg_out) {
    std::vector<double> terms(n, 0.0);
#pragma omp parallel
    {
        std::vector<uint32_t> neighbors(1024);
#pragma omp for schedule(dynamic, 32)
        for (int64_t pi = 0; pi < (int64_t)n; pi++) {
            const uint32_t p = (uint32_t)pi;
            if (int_tri_out) int_tri_out[p] = 0;
            if (int_tri_deg_out) int_tri_deg_out[p] = 0;
            const uint32_t label = labels[p];
            if (label == INVALID_LABEL) continue;
            const int tri = triangles[p], tri_deg = triangle_deg[p];
            int int_tri = 0, int_tri_deg = 0;
            const int size = sizes[label];
            if (size <= 2 || tri_deg == 0) continue;
            const uint32_t deg = indices[p + 1] - indices[p];
            if (deg > neighbors.size()) neighbors.resize(deg);
            uint32_t ns = 0;
            for (uint32_t i = indices[p]; i != indices[p + 1]; i++)
                if (labels[adj[i]] == label) neighbors[ns++] = adj[i];
            for (uint32_t i = 0; i < ns; i++) {
                const uint32_t q = neighbors[i];
                uint32_t a = 0, b = indices[q];
                const uint32_t a_end = ns, b_end = indices[q + 1];
                bool found = false;
                while (a != a_end && b != b_end) {
                    const uint32_t x = neighbors[a], y = adj[b];
                    if (x == y) {
                        found = true;
                        int_tri = wadd(int_tri, 1);
                    }
                    if (x <= y) a++;
                    if (x >= y) b++;
                }
                if (found) int_tri_deg++;
            }
            if (int_tri_out) int_tri_out[p] = int_tri;
            if (int_tri_deg_out) int_tri_deg_out[p] = int_tri_deg;
            const double nom = (double)wmul(int_tri, tri_deg);
            const double denom =
                tri * (tri_deg + alpha * (double)wsub(wsub(size, 1), int_tri_deg));
            if (denom != 0.0) terms[p] = nom / denom;
        }
    }
    double wcc = 0.0;
    for (uint32_t p = 0; p < n; p++) wcc += terms[p];
    if (terms_out) memcpy(terms_out, terms.data(), sizeof(double) * n);
    const int count = (int)n;  // every vertex counts (src/wcc.cpp:56), mask-free
    return count > 0 ? wcc / count : 0;
}

// The gather + sort + run-length step of src/scd.cpp:255-278 as explicit
// output: for each vertex, its neighbour labels sorted ascending (unsigned,
// INVALID last) into sorted_out[E], and the runs (label, count) of valid
// labels into run_label/run_count, with run_offsets[n+1] per vertex.
// Returns the total number of runs; run arrays must hold E entries.
uint64_t orc_neighbor_label_pairs(uint32_t n, const uint32_t *indices, const uint32_t *adj,
                                  const uint32_t *labels, uint32_t *sorted_out,
                                  uint32_t *run_offsets, uint32_t *run_label,
                                  uint32_t *run_count) {
    uint64_t total = 0;
    for (uint32_t p = 0; p < n; p++) {
        const uint32_t b = indices[p], e = indices[p + 1];
        for (uint32_t i = b; i < e; i++) sorted_out[i] = labels[adj[i]];
        std::sort(sorted_out + b, sorted_out + e);
        run_offsets[p] = (uint32_t)total;
        uint32_t i = b;
        while (i < e) {
            const uint32_t l = sorted_out[i];
            uint32_t f = 1;
            while (i + f < e && sorted_out[i + f] == l) f++;
            if (l != INVALID_LABEL) {
                run_label[total] = l;
                run_count[total] = f;
                total++;
            }
            i += f;
        }
    }
    run_offsets[n] = (uint32_t)total;
    return total;
}

// Per-vertex triangle statistics with the convert tool's semantics
// (src/convert.cpp:271-340): for every edge p<q, t = number of common
// neighbours found by merging the forward (higher-id) lists — each triangle
// is seen once at its lowest edge and credited to all three edges — then
// tri[v] = sum of t over incident edges, tri_deg[v] = #incident edges t > 0.
void orc_triangles(uint32_t n, const uint32_t *indices, const uint32_t *adj, int32_t *tri,
                   int32_t *tri_deg) {
    // forward list of p = entries of adj[p] greater than p (lists are sorted)
    std::vector<uint32_t> fwd(n + 1);
    for (uint32_t p = 0; p < n; p++)
        fwd[p] = (uint32_t)(std::upper_bound(adj + indices[p], adj + indices[p + 1], p) - adj);
    const uint64_t E = indices[n];
    std::vector<int> t(E, 0);  // indexed by the slot of the forward entry
#pragma omp parallel for schedule(guided)
    for (int64_t pi = 0; pi < (int64_t)n; pi++) {
        const uint32_t p = (uint32_t)pi;
        for (uint32_t i = fwd[p]; i < indices[p + 1]; i++) {
            const uint32_t q = adj[i];
            uint32_t a = fwd[p], b = fwd[q];
            const uint32_t a_end = indices[p + 1], b_end = indices[q + 1];
            while (a < a_end && b < b_end) {
                const uint32_t x = adj[a], y = adj[b];
                if (x == y) {
                    __sync_fetch_and_add(&t[i], 1);
                    __sync_fetch_and_add(&t[a], 1);
                    __sync_fetch_and_add(&t[b], 1);
                }
                if (x <= y) a++;
                if (x >= y) b++;
            }
        }
    }
    for (uint32_t v = 0; v < n; v++) {
        tri[v] = 0;
        tri_deg[v] = 0;
    }
    for (uint32_t p = 0; p < n; p++) {
        for (uint32_t i = fwd[p]; i < indices[p + 1]; i++) {
            const int c = t[i];
            if (c > 0) {
                const uint32_t q = adj[i];
                tri_deg[p]++;
                tri_deg[q]++;
                tri[p] = wadd(tri[p], c);
                tri[q] = wadd(tri[q], c);
            }
        }
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Convert restatement (src/convert.cpp:74-463), step by step, on an edge
// sequence instead of a text file.  Writes the binary image into a malloc'd
// buffer (*out, *out_len) so the caller can compare bytes.
// ---------------------------------------------------------------------------
extern "C" int orc_convert(uint64_t count, const uint32_t *eu, const uint32_t *ev,
                           uint8_t **out, uint64_t *out_len) {
    typedef std::pair<unsigned int, unsigned int> E;
    std::vector<E> edges(count);
    unsigned int num_nodes = 0;
    // read_edges (:120-125)
    for (uint64_t i = 0; i < count; i++) {
        const unsigned int a = eu[i], b = ev[i];
        edges[i] = std::make_pair(std::min(a, b), std::max(a, b));
        num_nodes = std::max(std::max(a, b) + 1, num_nodes);
    }
    // reindex_vertices (:141-170)
    {
        std::vector<unsigned int> convert(num_nodes, 0xFFFFFFFFu);
        unsigned int nn = 0;
        for (auto &e : edges) {
            const unsigned int a = e.first, b = e.second;
            if (convert[a] == 0xFFFFFFFFu) convert[a] = nn++;
            if (convert[b] == 0xFFFFFFFFu) convert[b] = nn++;
            e = std::make_pair(std::min(convert[a], convert[b]), std::max(convert[a], convert[b]));
        }
        num_nodes = nn;
    }
    // sort (:501) and remove_duplicates_and_loops (:252-268)
    std::sort(edges.begin(), edges.end());
    {
        size_t index = 0;
        for (size_t i = 0; i < edges.size(); i++)
            if (edges[i].first != edges[i].second && (i == 0 || edges[i] != edges[i - 1]))
                edges[index++] = edges[i];
        edges.resize(index);
    }
    // remove_edges_without_triangles (:271-351)
    std::vector<int> tri, tri_deg;
    {
        std::vector<unsigned int> indices(num_nodes + 1);
        size_t index = 0;
        indices[0] = 0;
        for (size_t i = 0; i < num_nodes; i++) {
            while (index < edges.size() && edges[index].first == i) index++;
            indices[i + 1] = (unsigned int)index;
        }
        std::vector<int> t(edges.size(), 0);
#pragma omp parallel for schedule(guided)
        for (int64_t ii = 0; ii < (int64_t)edges.size(); ii++) {
            const size_t i = (size_t)ii;
            const unsigned int p = edges[i].first, q = edges[i].second;
            unsigned int a = indices[p], b = indices[q];
            const unsigned int a_end = indices[p + 1], b_end = indices[q + 1];
            while (a < a_end && b < b_end) {
                const unsigned int x = edges[a].second, y = edges[b].second;
                if (x == y) {
                    __sync_fetch_and_add(&t[i], 1);
                    __sync_fetch_and_add(&t[a], 1);
                    __sync_fetch_and_add(&t[b], 1);
                }
                if (x <= y) a++;
                if (x >= y) b++;
            }
        }
        tri.assign(num_nodes, 0);
        tri_deg.assign(num_nodes, 0);
        for (size_t i = 0; i < edges.size(); i++) {
            const int c = t[i];
            if (c > 0) {
                tri_deg[edges[i].first]++;
                tri_deg[edges[i].second]++;
                tri[edges[i].first] = wadd(tri[edges[i].first], c);
                tri[edges[i].second] = wadd(tri[edges[i].second], c);
            }
        }
        index = 0;
        for (size_t i = 0; i < edges.size(); i++)
            if (t[i] > 0) edges[index++] = edges[i];
        edges.resize(index);
    }
    // sort_on_degree (:354-414)
    {
        std::vector<int> deg(num_nodes, 0);
        for (auto &e : edges) {
            deg[e.first]++;
            deg[e.second]++;
        }
        std::vector<std::pair<unsigned int, unsigned int>> degree_ids(num_nodes);
        for (unsigned int i = 0; i < num_nodes; i++) degree_ids[i] = std::make_pair((unsigned)deg[i], i);
        std::sort(degree_ids.begin(), degree_ids.end(),
                  std::greater<std::pair<unsigned int, unsigned int>>());
        std::vector<unsigned int> map(num_nodes);
        std::vector<int> ntri(num_nodes), ntd(num_nodes);
        for (unsigned int i = 0; i < num_nodes; i++) {
            map[degree_ids[i].second] = i;
            ntri[i] = tri[degree_ids[i].second];
            ntd[i] = tri_deg[degree_ids[i].second];
        }
        unsigned int new_num_nodes = 0;
        for (auto &e : edges) {
            e.first = map[e.first];
            e.second = map[e.second];
            new_num_nodes = std::max(new_num_nodes, e.first + 1);
            new_num_nodes = std::max(new_num_nodes, e.second + 1);
        }
        tri.swap(ntri);
        tri_deg.swap(ntd);
        num_nodes = new_num_nodes;
    }
    // write_edges (:417-463)
    const size_t num_edges = edges.size();
    edges.resize(2 * num_edges);
    for (size_t i = 0; i < num_edges; i++)
        edges[num_edges + i] = std::make_pair(edges[i].second, edges[i].first);
    std::sort(edges.begin(), edges.end());
    std::vector<uint32_t> words;
    words.push_back(num_nodes);
    words.push_back((uint32_t)num_edges);
    unsigned int curr_node = 0;
    words.push_back(curr_node);
    for (unsigned int i = 0; i < 2 * num_edges; i++) {
        while (edges[i].first != curr_node) {
            words.push_back(i);
            curr_node++;
        }
    }
    while (curr_node != num_nodes) {
        words.push_back((uint32_t)(2 * num_edges));
        curr_node++;
    }
    for (size_t i = 0; i < 2 * num_edges; i++) words.push_back(edges[i].second);
    for (unsigned int i = 0; i < num_nodes; i++) words.push_back((uint32_t)tri[i]);
    for (unsigned int i = 0; i < num_nodes; i++) words.push_back((uint32_t)tri_deg[i]);
    *out_len = words.size() * 4;
    *out = (uint8_t *)malloc(*out_len);
    if (!*out) return -3;
    memcpy(*out, words.data(), *out_len);
    return 0;
}

extern "C" void orc_free(void *p) { free(p); }

// ---------------------------------------------------------------------------
// Driver: src/main.cpp:35-93.  Records the labels after the initial partition
// and after every improve step, and the WCC of every iteration.
// ---------------------------------------------------------------------------
struct OrcRun {
    uint32_t n;
    uint32_t iterations;
    uint32_t num_communities;  // printed "num. communities" (src/scd.cpp:32-55)
    double cc;
    double initial_wcc;
    double final_wcc;          // the printed final WCC: best_wcc (src/main.cpp:90)
    double *wcc;               // per iteration
    uint32_t *labels;          // (iterations + 1) x n
    uint32_t *best_labels;     // n, uncompressed
    double refine_ms;          // loop wall time (src/main.cpp:57-84)
    double init_ms;            // initial partition + stats + initial WCC
};

static double now_ms() {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec * 1000.0 + tv.tv_usec / 1000.0;
}

// src/scd.cpp:32-55 (with its inverted count() test): returns the number the
// reference prints, n minus the number of distinct valid labels.
extern "C" uint32_t orc_compress_count(uint32_t n, const uint32_t *labels_in) {
    std::vector<uint32_t> labels(labels_in, labels_in + n);
    unsigned int num_labels = 0;
    std::map<unsigned int, unsigned int> label_map;
    for (uint32_t p = 0; p < n; p++) {
        unsigned int l = labels[p];
        if (l == INVALID_LABEL) l = num_labels++;
        else if (label_map.count(l)) l = label_map[l] = num_labels++;
        else l = label_map[l];
        labels[p] = l;
    }
    return num_labels;
}

extern "C" int orc_run(uint32_t n, const uint32_t *indices, const uint32_t *adj,
                       const int32_t *tri, const int32_t *tri_deg, uint32_t max_iters,
                       OrcRun *out) {
    const double alpha = 1.0, threshold = 0.01;
    memset(out, 0, sizeof(*out));
    out->n = n;
    out->cc = orc_clustering_coef(n, indices, tri);
    std::vector<uint32_t> labels(n), next(n);
    std::vector<int32_t> sizes(n + 1), ext_e(n + 1), int_e(n + 1);
    const double t0 = now_ms();
    const uint32_t num_comms = orc_initial_partition(n, indices, adj, tri, labels.data());
    orc_collect_stats(n, indices, adj, labels.data(), num_comms, sizes.data(), ext_e.data(),
                      int_e.data());
    double best_wcc = orc_compute_wcc(n, indices, adj, tri, tri_deg, labels.data(), sizes.data(),
                                      alpha, nullptr, nullptr, nullptr);
    out->init_ms = now_ms() - t0;
    out->initial_wcc = best_wcc;
    std::vector<uint32_t> best(labels);
    std::vector<uint32_t> history(labels);
    std::vector<double> wccs;
    const double before = now_ms();
    uint32_t iters = 0;
    while (iters < max_iters) {
        iters++;
        orc_improve(n, indices, adj, labels.data(), sizes.data(), ext_e.data(), int_e.data(),
                    out->cc, alpha, next.data());
        labels.swap(next);
        orc_collect_stats(n, indices, adj, labels.data(), num_comms, sizes.data(), ext_e.data(),
                          int_e.data());
        const double wcc = orc_compute_wcc(n, indices, adj, tri, tri_deg, labels.data(),
                                           sizes.data(), alpha, nullptr, nullptr, nullptr);
        if (wcc > best_wcc) best = labels;
        history.insert(history.end(), labels.begin(), labels.end());
        wccs.push_back(wcc);
        if (wcc < best_wcc * (1 + threshold)) break;
        best_wcc = wcc;
    }
    out->refine_ms = now_ms() - before;
    out->iterations = iters;
    out->final_wcc = best_wcc;
    out->num_communities = orc_compress_count(n, best.data());
    out->wcc = (double *)malloc(sizeof(double) * (iters ? iters : 1));
    out->labels = (uint32_t *)malloc(sizeof(uint32_t) * history.size() + 4);
    out->best_labels = (uint32_t *)malloc(sizeof(uint32_t) * n + 4);
    if (!out->wcc || !out->labels || !out->best_labels) return -3;
    std::copy(wccs.begin(), wccs.end(), out->wcc);
    std::copy(history.begin(), history.end(), out->labels);
    std::copy(best.begin(), best.end(), out->best_labels);
    return 0;
}

extern "C" void orc_run_free(OrcRun *r) {
    free(r->wcc);
    free(r->labels);
    free(r->best_labels);
    r->wcc = nullptr;
    r->labels = nullptr;
    r->best_labels = nullptr;
}
