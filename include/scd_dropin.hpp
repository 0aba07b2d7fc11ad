/*
 * scd_dropin.hpp — the reference's single-GPU interface (src/cuda.hpp:23-28)
 * with its exact names, C++ signatures and behaviour, implemented over the C
 * ABI of scd_hip.h by libscd_dropin.so.  A program written against the
 * reference header (src/main-gpu.cpp) links against this library unchanged.
 *
 * Behaviour kept from the reference:
 *   cuda_init            prints "selected device:" and the device name; an
 *                        invalid ordinal lists the devices on stderr and
 *                        returns false (src/cuda.cu:55-82)
 *   cuda_transfer_graph  alpha is ignored (src/cuda.cu:84); prints the memory
 *                        lines of src/cuda.cu:133-145; returns true
 *   cuda_improve_labels  prints the four phase timings (src/cuda.cu:351-360)
 *   any device error     message on stderr and exit(code) (src/cuda-common.cuh:67-77)
 *
 * The Graph type is the reference's class (src/graph.hpp:23-62): the library
 * reads its public members n, m, clustering_coef, triangle_deg, triangles,
 * indices and adj, declared in that order; scd_graph_layout.hpp holds a
 * declaration with that layout for programs that do not include graph.hpp.
 */
#ifndef SCD_CUDA_SHIM_HPP
#define SCD_CUDA_SHIM_HPP

class Graph;

bool cuda_init(int device_id, int work_group_size, bool profile_flag);
bool cuda_transfer_graph(const Graph *g, double alpha);
void cuda_labels_to_device(unsigned int *labels);
void cuda_improve_labels(double *wcc);
void cuda_store_labels();
void cuda_labels_from_device(unsigned int *labels);

/* Not in the reference: the context the drop-in library drives, for callers that mix
 * the reference names with the C ABI (scd_hip.h). */
struct scd_ctx;
scd_ctx *scd_dropin_context();

#endif
