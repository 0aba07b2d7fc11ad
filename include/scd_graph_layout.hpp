/*
 * scd_graph_layout.hpp — a declaration of the reference's Graph class with
 * the same public data members in the same order (src/graph.hpp:23-33), so
 * that code built without the reference's graph.hpp (the drop-in library and
 * this project's main-gpu) shares its object layout.  Only the members, the
 * inline CSR accessors and Graph::read are declared; construction is left to
 * the caller.
 */
#ifndef SCD_GRAPH_LAYOUT_HPP
#define SCD_GRAPH_LAYOUT_HPP

class Graph {
  public:
    const unsigned int n;
    const unsigned int m;
    const double clustering_coef;
    const int *const triangle_deg;
    const int *const triangles;
    const unsigned int *const indices;
    const unsigned int *const adj;

    Graph(unsigned int n_, unsigned int m_, double cc, const int *tri_deg, const int *tri,
          const unsigned int *idx, const unsigned int *adjacency)
        : n(n_), m(m_), clustering_coef(cc), triangle_deg(tri_deg), triangles(tri), indices(idx),
          adj(adjacency) {}

    unsigned int degree(unsigned int i) const { return indices[i + 1] - indices[i]; }

    /* Graph::read (src/graph.hpp:41); defined by libscd_refapi.so. */
    static const Graph *read(const char *filename);
};

#endif
