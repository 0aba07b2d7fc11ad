// Community labels on the host: the printed count of the reference's
// compress_labels, a corrected compaction, and a labels file.
//
// Labels file: [u32 n][u32 num_communities][u32 labels n], little-endian,
// labels compact in 0..num_communities-1 (scd_labels_compact).
#include <cstdio>
#include <cstdlib>
#include <unordered_map>

#include "host_common.hpp"
#include "scd_host.h"

using scd::set_host_error;

// src/scd.cpp:32-55 inverts its map test, so every repeated label gets a
// fresh number and the printed count is n minus the number of distinct
// valid labels.  Kept for stdout parity.
extern "C" uint32_t scd_labels_printed_count(uint32_t n, const uint32_t *labels) {
    if (!labels) return 0;
    std::unordered_map<uint32_t, uint32_t> seen;
    seen.reserve(n);
    uint32_t distinct = 0;
    for (uint32_t v = 0; v < n; v++) {
        const uint32_t l = labels[v];
        if (l != SCD_INVALID_LABEL && seen.emplace(l, 0).second) distinct++;
    }
    return n - distinct;
}

// What src/scd.cpp:32-55 intends: distinct valid labels numbered in order of
// first appearance, every INVALID vertex a community of its own.
extern "C" int scd_labels_compact(uint32_t n, const uint32_t *labels, uint32_t *out, uint32_t *num_communities) {
    if ((n && (!labels || !out)) || !num_communities) {
        set_host_error("scd_labels_compact: null argument");
        return SCD_EINVAL;
    }
    std::unordered_map<uint32_t, uint32_t> id;
    id.reserve(n);
    uint32_t k = 0;
    for (uint32_t v = 0; v < n; v++) {
        const uint32_t l = labels[v];
        if (l == SCD_INVALID_LABEL) {
            out[v] = k++;
        } else {
            auto it = id.emplace(l, k);
            if (it.second) k++;
            out[v] = it.first->second;
        }
    }
    *num_communities = k;
    return SCD_OK;
}

extern "C" int scd_labels_write(const char *path, uint32_t n, const uint32_t *labels) {
    if (!path || (n && !labels)) {
        set_host_error("scd_labels_write: null argument");
        return SCD_EINVAL;
    }
    uint32_t *c = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    if (!c) {
        set_host_error("scd_labels_write: out of memory");
        return SCD_ENOMEM;
    }
    uint32_t k = 0;
    int rc = scd_labels_compact(n, labels, c, &k);
    if (rc) {
        free(c);
        return rc;
    }
    FILE *f = fopen(path, "wb");
    if (!f) {
        free(c);
        set_host_error("failed to open %s for writing", path);
        return SCD_EIO;
    }
    const uint32_t hdr[2] = {n, k};
    bool ok = fwrite(hdr, 4, 2, f) == 2 && fwrite(c, 4, n, f) == n;
    ok = (fclose(f) == 0) && ok;
    free(c);
    if (!ok) {
        set_host_error("an error occured while writing %s", path);
        return SCD_EIO;
    }
    return SCD_OK;
}

extern "C" int scd_labels_read(const char *path, uint32_t *n_out, uint32_t *k_out, uint32_t **labels_out) {
    if (!path || !n_out || !k_out || !labels_out) {
        set_host_error("scd_labels_read: null argument");
        return SCD_EINVAL;
    }
    *labels_out = nullptr;
    FILE *f = fopen(path, "rb");
    if (!f) {
        set_host_error("failed to open %s", path);
        return SCD_EIO;
    }
    uint32_t hdr[2];
    if (fread(hdr, 4, 2, f) != 2) {
        fclose(f);
        set_host_error("%s: short header", path);
        return SCD_EIO;
    }
    uint32_t *l = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)hdr[0] + 1));
    if (!l) {
        fclose(f);
        set_host_error("scd_labels_read: out of memory");
        return SCD_ENOMEM;
    }
    char extra;
    const bool ok = fread(l, 4, hdr[0], f) == hdr[0] && fread(&extra, 1, 1, f) == 0;
    fclose(f);
    bool range = true;
    for (uint32_t v = 0; ok && v < hdr[0]; v++) range = range && l[v] < hdr[1];
    if (!ok || !range) {
        free(l);
        set_host_error("%s: %s", path, ok ? "label out of range" : "size does not match the header");
        return SCD_EIO;
    }
    *n_out = hdr[0];
    *k_out = hdr[1];
    *labels_out = l;
    return SCD_OK;
}

extern "C" void scd_labels_free(uint32_t *labels) { free(labels); }
