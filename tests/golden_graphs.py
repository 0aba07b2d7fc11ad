"""Named test graphs: a recipe (generator, parameters, seed) per graph, so
the same edge sequence is produced here, in the fixtures and on any host.

Sizes:  tiny (bowtie, K_n, n <= 64 random), small (planted partitions of a
few hundred vertices, R-MAT scale 10-14), C1 (planted partition 100 x 100,
p_in 0.1, p_out 0.0005), R-MAT scale 16/18 and C2 (R-MAT scale 20, edge
factor 16, a=0.57 b=0.19 c=0.19); C3 (R-MAT scale 24, same parameters), whose
fixture keeps hashes only and whose GPU tests convert it on the device.
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "het-scd_amd"))
import scd_amd  # noqa: E402

GRAPHS = {
    "bowtie": {"kind": "pairs", "pairs": [[0, 1], [1, 2], [2, 0], [2, 3], [3, 4], [4, 2]]},
    "k5": {"kind": "complete", "n": 5},
    "k9": {"kind": "complete", "n": 9},
    "er24_s1": {"kind": "planted", "sizes": [24], "p_in": 0.35, "p_out": 0.0, "seed": 1},
    "er40_s2": {"kind": "planted", "sizes": [40], "p_in": 0.25, "p_out": 0.0, "seed": 2},
    "er64_s3": {"kind": "planted", "sizes": [64], "p_in": 0.2, "p_out": 0.0, "seed": 3},
    "pp4x16_s4": {"kind": "planted", "sizes": [16, 16, 16, 16], "p_in": 0.5, "p_out": 0.05, "seed": 4},
    "pp8x40_s5": {"kind": "planted", "sizes": [40] * 8, "p_in": 0.3, "p_out": 0.01, "seed": 5},
    "pprand_s6": {"kind": "planted_random", "seed": 6, "n_min": 64, "n_max": 2000, "k_min": 4, "k_max": 64,
                  "p_in": 0.15, "p_out": 0.002},
    "pprand_s7": {"kind": "planted_random", "seed": 7, "n_min": 64, "n_max": 20000, "k_min": 4, "k_max": 64,
                  "p_in": 0.15, "p_out": 0.002},
    "c1": {"kind": "planted", "sizes": [100] * 100, "p_in": 0.1, "p_out": 0.0005, "seed": 1},
    "rmat10": {"kind": "rmat", "scale": 10, "ef": 16, "seed": 1},
    "rmat12": {"kind": "rmat", "scale": 12, "ef": 16, "seed": 2},
    "rmat14": {"kind": "rmat", "scale": 14, "ef": 16, "seed": 3},
    "rmat16": {"kind": "rmat", "scale": 16, "ef": 16, "seed": 1},
    "rmat18": {"kind": "rmat", "scale": 18, "ef": 16, "seed": 1},
    "c2": {"kind": "rmat", "scale": 20, "ef": 16, "seed": 1},
    "c3": {"kind": "rmat", "scale": 24, "ef": 16, "seed": 1},
}

SMALL = ["bowtie", "k5", "k9", "er24_s1", "er40_s2", "er64_s3", "pp4x16_s4", "pp8x40_s5", "pprand_s6",
         "rmat10", "rmat12"]
MEDIUM = ["pprand_s7", "c1", "rmat14", "rmat16"]
LARGE = ["rmat18", "c2"]
HUGE = ["c3"]  # written only when named (make_golden.py c3)


def make_edges(recipe):
    k = recipe["kind"]
    if k == "pairs":
        return scd_amd.from_edge_pairs(recipe["pairs"])
    if k == "complete":
        n = recipe["n"]
        return scd_amd.from_edge_pairs([(i, j) for i in range(n) for j in range(i + 1, n)])
    if k == "planted":
        return scd_amd.planted_partition(np.array(recipe["sizes"], np.uint32), recipe["p_in"], recipe["p_out"],
                                         recipe["seed"])
    if k == "planted_random":
        e, _, _ = scd_amd.planted_random(recipe["seed"], recipe["n_min"], recipe["n_max"], recipe["k_min"],
                                         recipe["k_max"], recipe["p_in"], recipe["p_out"])
        return e
    if k == "rmat":
        return scd_amd.rmat(recipe["scale"], recipe["ef"], 0.57, 0.19, 0.19, recipe["seed"])
    raise ValueError(k)


_cache = {}


def graph(name):
    """The converted graph of a named recipe (cached per process)."""
    if name not in _cache:
        _cache[name] = make_edges(GRAPHS[name]).convert()
    return _cache[name]
