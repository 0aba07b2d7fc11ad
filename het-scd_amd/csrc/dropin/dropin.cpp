// libscd_dropin.so: the six functions of the reference's src/cuda.hpp
// over the engine's C ABI, with one process-wide context as the reference
// keeps its device state in file-static globals (src/cuda.cu:29-52).
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include <hip/hip_runtime_api.h>

#include "scd_dropin.hpp"
#include "scd_graph_layout.hpp"
#include "scd_hip.h"

static scd_ctx *g_ctx = nullptr;

scd_ctx *scd_dropin_context() { return g_ctx; }

// Device errors end the process, like CUDA_CALL (src/cuda-common.cuh:67-77).
static void must(int rc, const char *what) {
    if (rc != SCD_OK) {
        std::cerr << what << ": " << scd_last_error() << std::endl;
        exit(rc);
    }
}

bool cuda_init(int device_id, int work_group_size, bool profile_flag) {
    int count = scd_device_count();
    if (device_id < 0 || device_id >= count) {
        std::cerr << "Please select a HIP device:" << std::endl << std::endl;
        for (int i = 0; i < count; i++) {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, i) == hipSuccess)
                std::cerr << i << ": " << prop.name << " (" << prop.gcnArchName << ")" << std::endl << std::endl;
        }
        return false;
    }
    // profile_flag is kept and unused, as in the reference (src/cuda.cu:79):
    // iterations always print their four phase times there, and here they
    // come from device clock stamps inside the captured, laned iteration, so
    // the drop-in runs the same launch sequence as scd_refine
    (void)profile_flag;
    int rc = scd_init(device_id, work_group_size, 0, &g_ctx);
    if (rc == SCD_OK) rc = scd_set_phase_timing(g_ctx, 1);
    if (rc != SCD_OK) {
        std::cerr << scd_last_error() << std::endl;
        return false;
    }
    hipDeviceProp_t prop;
    std::cout << "selected device:" << std::endl;
    if (hipGetDeviceProperties(&prop, device_id) == hipSuccess)
        std::cout << device_id << ": " << prop.name << " (" << prop.gcnArchName << ", "
                  << prop.multiProcessorCount << " CUs, " << (prop.totalGlobalMem >> 20) << " MiB)"
                  << std::endl;
    return true;
}

bool cuda_transfer_graph(const Graph *g, double alpha) {
    (void)alpha;  // the reference ignores it too (alpha = 1.0 on the device)
    size_t free_before = 0, total = 0, free_after = 0;
    (void)hipMemGetInfo(&free_before, &total);
    must(scd_upload(g_ctx, g->n, g->m, g->indices, g->adj, g->triangles, g->triangle_deg, g->clustering_coef),
         "cuda_transfer_graph");
    (void)hipMemGetInfo(&free_after, &total);
    if (total) {
        const size_t reserved = total - free_before, used = free_before - free_after, avail = free_after;
        std::cout << "memory:" << std::endl;
        std::cout << " - " << reserved << " bytes reserved (" << reserved * 100 / total << "%)" << std::endl;
        std::cout << " - " << used << " bytes used (" << used * 100 / total << "%)" << std::endl;
        std::cout << " - " << avail << " bytes available (" << avail * 100 / total << "%)" << std::endl;
    }
    return true;
}

void cuda_labels_to_device(unsigned int *labels) {
    must(scd_labels_to_device(g_ctx, labels), "cuda_labels_to_device");
    // with SCD_TUNE_CACHE naming the autotuner's text file, the kernel
    // shapes of this graph's size class from it (measured and appended
    // when the class is new; the labels stay as they are)
    if (const char *tc = getenv("SCD_TUNE_CACHE"))
        if (*tc) must(scd_autotune(g_ctx, tc, 3, nullptr, nullptr), "cuda_labels_to_device");
    // the driver's loop starts next: record its iterations now
    must(scd_capture_iterations(g_ctx), "cuda_labels_to_device");
}

void cuda_improve_labels(double *wcc) {
    must(scd_improve(g_ctx, wcc, nullptr), "cuda_improve_labels");
    // the phase split of src/cuda.cu:351-360
    double ms[4] = {0, 0, 0, 0};
    must(scd_last_phase_ms(g_ctx, ms), "cuda_improve_labels");
    std::cout << "improve partitioning: " << ms[0] << " ms" << std::endl;
    std::cout << "collect statistics: " << ms[1] << " ms" << std::endl;
    std::cout << "remove external edges: " << ms[2] << " ms" << std::endl;
    std::cout << "compute wcc: " << ms[3] << " ms" << std::endl;
}

void cuda_store_labels() { must(scd_store_best(g_ctx), "cuda_store_labels"); }

void cuda_labels_from_device(unsigned int *labels) {
    must(scd_labels_from_device(g_ctx, labels), "cuda_labels_from_device");
}
