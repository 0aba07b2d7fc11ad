// ORACLE — test infrastructure only.
//
// Tables the reference's own compress_labels (src/scd.cpp:32-55), compiled
// from where it lies together with src/common.cpp (its timer).  The function
// reads only g->n of its Graph argument; the harness passes a record with
// the layout of Graph's leading members (src/graph.hpp:25-27).
// stdin:  "n" then n labels (decimal, INVALID = 4294967295), repeated
// stdout: per array, the count, then the n relabelled entries
#include <cstdio>
#include <vector>

#include "scd.hpp"

struct GraphHead {
    unsigned int n, m;
    double clustering_coef;
    const void *p[4];
};

int main() {
    unsigned int n;
    while (scanf("%u", &n) == 1) {
        std::vector<unsigned int> labels(n);
        for (unsigned int i = 0; i < n; i++)
            if (scanf("%u", &labels[i]) != 1) return 1;
        GraphHead head = {n, 0, 0.0, {nullptr, nullptr, nullptr, nullptr}};
        unsigned int count = 0;
        compress_labels(reinterpret_cast<const Graph *>(&head), labels.data(), &count);
        printf("%u", count);
        for (unsigned int i = 0; i < n; i++) printf(" %u", labels[i]);
        printf("\n");
    }
    return 0;
}
