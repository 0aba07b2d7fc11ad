// libscd_hip.so: the MI355X SCD refinement engine behind the C ABI of
// include/scd_hip.h.
//
// Device state per context (all resident in HBM for the life of the graph):
//   CSR        indices[n+1], adj[E], tri[n], tri_deg[n]          (uint32/int32)
//   labels     cur[n], next[n], best[n]                           (uint32)
//   stats      int4[n]: (size, ext, int, -) per community         (int32)
//   internal   int_adj[E] (same-label lists at the CSR offsets), int_deg[n]
//   wcc        terms[n] (fp64), partial sums, the mean
//   bins       vertex ids grouped by degree bin (host-built at upload)
//   hubs       per-hub descriptors, tiles, chunks and two scratch segments
// One iteration (scd_improve) is: move kernels over every bin (old labels
// -> next), swap, clear statistics, filter+statistics kernels, WCC kernels,
// two-step deterministic sum; then one device->host copy of the mean.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "dev.hpp"
#include "move_kernels.hpp"
#include "scd_hip.h"
#include "stats_kernels.hpp"
#include "ip_kernels.hpp"
#include "wcc_kernels.hpp"
#include "arena_kernels.hpp"
#include "bins_kernels.hpp"
#include "convert_kernels.hpp"

using namespace scd_dev;

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
static thread_local char g_err[512];

static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char *scd_last_error(void) { return g_err; }

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(e_ == hipErrorOutOfMemory ? SCD_ENOMEM : SCD_EDEVICE, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                     \
    } while (0)

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct ProfRec {
    std::string name;
    uint64_t launches = 0;
    double ms = 0.0;
};

struct PendingEv {
    int rec;
    hipEvent_t a, b;
};

struct scd_ctx {
    int device = 0;
    int block_size = 256;
    uint32_t num_cus = 256;
    std::string arch = "unknown";  // gcnArchName without feature suffixes
    bool profile = false;
    hipStream_t stream = nullptr;
    uint64_t dev_bytes = 0;

    // graph
    uint32_t n = 0, m = 0;
    uint64_t E = 0;
    double cc = 0.0;
    bool have_graph = false, have_labels = false;
    uint32_t num_comms = 0;
    uint32_t dmax = 0;  // largest degree (set with the degree bins)

    uint32_t *indices = nullptr, *adj = nullptr;
    int32_t *tri = nullptr, *tri_deg = nullptr;
    uint32_t *cur = nullptr, *next = nullptr, *best = nullptr;
    int4 *stats = nullptr;
    uint32_t *int_adj = nullptr, *int_deg = nullptr;
    uint8_t *iflag = nullptr;                              // per internal entry
    uint32_t *wdplus = nullptr, *wT = nullptr, *wtd = nullptr;  // internal triangle counts
    // per label: community member set changed by the last move
    uint8_t *wdirty = nullptr;
    // private copies of the triangle counts of the lowest ids (zero between
    // passes)
    uint32_t *hot_T = nullptr, *hot_td = nullptr;
    uint32_t hot_n = 0;
    // wT / wtd hold the counts of the labels in cur (set by every WCC
    // evaluation, cleared by everything that writes cur otherwise): the
    // next iteration then recounts only the communities its moves change
    bool wcc_valid = false;
    double *terms = nullptr, *partial = nullptr, *wcc_dev = nullptr;
    double *wcc_host = nullptr;  // pinned: [0] mean WCC of the last pass
    // input checks of the uploads: device word set by the check kernels,
    // its pinned host copy
    uint32_t *chk = nullptr;
    uint32_t *chk_host = nullptr;
    // device counters of the last pass: [0, K_NB) evaluated pairs per degree
    // bin, [K_NB] internal adjacency entries, [K_NB + 1] sum of internal
    // degree squared
    unsigned long long *counters = nullptr;
    unsigned long long *counters_host = nullptr;  // pinned, 16 entries
    uint64_t bin_edges[16] = {0};  // adjacency entries per bin
    bool counting = false;         // fill the counters (scd_set_counting)

    // batched arena (G > 0): graph of each vertex, per-graph clustering
    // coefficient / activity / store flags, per-graph WCC reduction items
    uint32_t G = 0;
    std::vector<uint64_t> h_vtx_off;  // G + 1
    std::vector<uint32_t> h_label_off;  // G: first global label of each graph
    uint32_t *gid = nullptr, *active_g = nullptr, *store_g = nullptr, *k_g = nullptr;
    double *cc_g = nullptr, *wcc_g = nullptr, *wcc_g_host = nullptr;
    size_t wcc_g_host_cap = 0;  // doubles held by wcc_g_host (pinned, kept across arenas)
    void *h_stage = nullptr;    // pinned staging for label downloads (kept, grown)
    size_t h_stage_bytes = 0;
    uint64_t *vtx_off_d = nullptr;
    uint32_t *witem_start = nullptr, *witem_end = nullptr, *g_item_begin = nullptr;
    uint32_t n_witems = 0;
    double *partial_g = nullptr;

    // snapshot for scd_snapshot / scd_restore
    uint32_t *snap_labels = nullptr;
    int4 *snap_stats = nullptr;
    uint32_t snap_num_comms = 0;
    bool have_snap = false;

    // bins (bins_partial: only the vertices of the graphs still refining)
    uint32_t *bin_verts = nullptr;
    uint32_t bin_begin[K_NB + 1] = {0};
    bool bins_partial = false;

    // work queue of the internal triangle count (vertices of the W1..HUB
    // bins, WTC_VC out-list entries per item): chunk counts, their scan
    // (device-only, scan levels preallocated), item total and counter
    uint32_t wq_M = 0;
    uint32_t *wq_nch = nullptr, *wq_off = nullptr, *wq_tot = nullptr, *wq_next = nullptr;
    // items of the triangle pass counted from the longer out-list's side
    uint32_t *vq_nch = nullptr, *vq_off = nullptr, *vq_tot = nullptr, *vq_next = nullptr;
    uint32_t tc_heavy = 128;  // WccArgs::vc_min of graph triangle counting (0: off)
    uint32_t wcc_heavy = 0;   // the same for the WCC's internal triangles
    uint32_t *wq_sum[4] = {nullptr}, *wq_scan[4] = {nullptr};
    // the WCC's lists of those vertices with 1..WTC_PACK_LO and more
    // internal out-neighbours (k_wtc_pack), and their two counters
    uint32_t *wq_lo = nullptr, *wq_hi = nullptr, *wq_pcnt = nullptr;

    // hubs
    uint32_t nh = 0, n_tiles = 0, n_chunks = 0, max_hub_deg = 0;
    uint32_t *hub_v = nullptr, *hub_d = nullptr, *hub_off = nullptr, *hub_chunk_begin = nullptr;
    uint32_t *tile_hub = nullptr, *tile_start = nullptr, *chunk_hub = nullptr, *chunk_start = nullptr;
    uint32_t *hub_a = nullptr, *hub_b = nullptr, *chunk_cnt = nullptr, *chunk_bl = nullptr,
             *chunk_has_rem = nullptr;
    double *chunk_bi = nullptr, *chunk_rem = nullptr;
    // hash tables of the hub candidates (move without export): slots, per-hub
    // offset and capacity, chunks of slots and their results
    unsigned long long *hub_slots = nullptr;
    uint64_t *tab_off = nullptr;
    uint32_t *tab_cap = nullptr, *tchunk_hub = nullptr, *tchunk_start = nullptr, *hub_tchunk_begin = nullptr;
    uint32_t *tchunk_bl = nullptr, *tchunk_has_rem = nullptr;
    double *tchunk_bi = nullptr, *tchunk_rem = nullptr;
    uint32_t n_tchunks = 0;
    // label-range buckets of the hubs (tune.hub_table 2): per hub the sample
    // stride, samples per bucket start and first slot; per slot its hub,
    // bucket start, entry count and result; partition chunks; per hub the
    // buckets in use and the overflow count; HB_CAP entries per slot
    uint32_t *hb_sigma = nullptr, *hb_kstep = nullptr, *hb_slot0 = nullptr, *hb_slot_hub = nullptr,
             *hb_pch_hub = nullptr, *hb_pch_start = nullptr, *hb_lo = nullptr, *hb_nb = nullptr, *hb_cnt = nullptr,
             *hb_ovf_cnt = nullptr, *hb_region = nullptr, *hb_bl = nullptr, *hb_has_rem = nullptr;
    double *hb_bi = nullptr, *hb_rem = nullptr;
    uint32_t hb_slots = 0, hb_pchunks = 0;
    uint32_t hb_cap = HB_CAP;  // SCD_HUB_BUCKET_CAP (tests) lowers it to reach the overflow path
    // chunks launched by each merge pass (a prefix of the chunk list: hubs
    // are ordered by degree, largest first)
    std::vector<uint32_t> merge_pass_chunks;
    // dense-community WCC (tune.dense_wcc): per-label row offsets and
    // local-id counters, per-vertex local ids, the list of dense-community
    // vertices, two counters and the row buffer (dn_units 256-byte units)
    uint32_t *dn_off = nullptr, *dn_vbase = nullptr, *dn_cursor = nullptr, *dn_lid = nullptr, *dn_list = nullptr,
             *dn_ctr = nullptr;
    unsigned long long *dn_bits = nullptr;
    uint32_t dn_units = 0;
    std::vector<DevBuf> hub_bufs;

    scd_tuning tune = {4096, 2048, 1024, 0, 0, 1, 0, 0};
    uint32_t last_ip_rounds = 0;

    // per-iteration hipGraphs, one per label-buffer parity (cur == lab_a or
    // lab_b); dropped whenever a launch parameter changes
    bool use_graph = true;
    bool capturing = false;
    uint32_t *lab_a = nullptr, *lab_b = nullptr;
    hipGraphExec_t gexec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [parity][incremental WCC]

    // profiling
    std::vector<ProfRec> prof;
    std::vector<PendingEv> pending;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;

    // independent per-bin kernels run on lanes: the context's stream plus
    // three more (default priority), forked and joined with events, so
    // small and tail-bound launches overlap (captured as parallel branches)
    hipStream_t lane[3] = {nullptr, nullptr, nullptr};
    // an iteration's community statistics run on this stream beside the
    // WCC's triangle counting, which does not read them; stats_ev (recorded
    // after them) is waited for before the first kernel that does
    hipStream_t side = nullptr;
    hipEvent_t stats_ev = nullptr;
    bool stats_pending = false;
    // fork and join events: every fork or join point of a launch sequence
    // records its own events (pool reset when the stream is idle, and at the
    // start of each capture; captures draw from a pool of their own)
    std::vector<hipEvent_t> lane_ev[2];  // [0] eager, [1] captured
    size_t lane_ev_used[2] = {0, 0};
    hipStream_t ls = nullptr;  // the stream LAUNCH uses
    bool use_lanes = true;
    bool one_stream = false;  // set for the length of a call that must not create events

    // phase timing of each iteration (cuda_improve_labels' four lines,
    // src/cuda.cu:347-360): device wall-clock stamps written by one-thread
    // kernels at the phase boundaries, inside captured iterations too
    bool phase_timing = false;
    unsigned long long *stamps = nullptr;       // device, kStamps entries
    unsigned long long *stamps_host = nullptr;  // pinned copy
    double wall_clock_khz = 0.0;
    double phase_ms[4] = {0, 0, 0, 0};

    std::vector<DevBuf> bufs;
    // buffers of the previous graph, reused by the next upload when a size
    // fits (a server re-assembles arenas often); freed one upload later
    std::vector<DevBuf> cache;
    // temporaries of the entry points, kept for later calls up to
    // kTmpCacheBytes (hipMalloc and hipFree cost more than many kernels)
    std::vector<DevBuf> tmp_cache;
    size_t tmp_cache_bytes = 0;

    scd_convert_info cv_info{};  // counts and stage times of the last convert
};

constexpr size_t kTmpCacheBytes = 4ull << 30;
constexpr double kTuneMargin = 0.03;  // autotuner: required gain of the dense-community WCC

template <typename T>
static int dalloc(scd_ctx *c, T **p, size_t count, std::vector<DevBuf> *list = nullptr) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    // graph buffers come from the previous graph's, temporaries and degree
    // bin lists from earlier ones (best fit, at most 2x + 1 MiB)
    std::vector<DevBuf> *pool = !list ? &c->cache : &c->tmp_cache;
    if (pool) {
        size_t best = pool->size();
        for (size_t i = 0; i < pool->size(); i++) {
            const size_t b = (*pool)[i].bytes;
            if (b >= bytes && b <= 2 * bytes + (1u << 20) && (best == pool->size() || b < (*pool)[best].bytes))
                best = i;
        }
        if (best < pool->size()) {
            const DevBuf d = (*pool)[best];
            *p = (T *)d.p;
            (list ? list : &c->bufs)->push_back(d);
            pool->erase(pool->begin() + best);
            if (pool == &c->tmp_cache) c->tmp_cache_bytes -= d.bytes;
            return SCD_OK;
        }
    }
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess)
        return fail(SCD_ENOMEM, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    *p = (T *)q;
    (list ? list : &c->bufs)->push_back({q, bytes});
    c->dev_bytes += bytes;
    return SCD_OK;
}

static void free_list(scd_ctx *c, std::vector<DevBuf> &list) {
    for (auto &b : list) {
        (void)hipFree(b.p);
        c->dev_bytes -= b.bytes;
    }
    list.clear();
}

// Temporaries go back to the context's pool (later work on the stream is
// ordered after the kernels that used them); beyond kTmpCacheBytes the
// largest are freed.
static void release_tmp(scd_ctx *c, std::vector<DevBuf> &tmp) {
    for (auto &b : tmp) {
        c->tmp_cache.push_back(b);
        c->tmp_cache_bytes += b.bytes;
    }
    tmp.clear();
    while (c->tmp_cache_bytes > kTmpCacheBytes && !c->tmp_cache.empty()) {
        size_t big = 0;
        for (size_t i = 1; i < c->tmp_cache.size(); i++)
            if (c->tmp_cache[i].bytes > c->tmp_cache[big].bytes) big = i;
        (void)hipFree(c->tmp_cache[big].p);
        c->dev_bytes -= c->tmp_cache[big].bytes;
        c->tmp_cache_bytes -= c->tmp_cache[big].bytes;
        c->tmp_cache.erase(c->tmp_cache.begin() + big);
    }
}

static void drop_graphs(scd_ctx *c) {
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            if (c->gexec[i][j]) (void)hipGraphExecDestroy(c->gexec[i][j]);
            c->gexec[i][j] = nullptr;
        }
}

static void free_graph(scd_ctx *c, bool keep_cache = true) {
    drop_graphs(c);
    if (keep_cache) release_tmp(c, c->hub_bufs);
    else free_list(c, c->hub_bufs);
    free_list(c, c->cache);
    if (keep_cache) c->cache.swap(c->bufs);
    else free_list(c, c->bufs);
    if (!keep_cache) {
        if (c->wcc_g_host) (void)hipHostFree(c->wcc_g_host);
        c->wcc_g_host = nullptr;
        c->wcc_g_host_cap = 0;
        if (c->h_stage) (void)hipHostFree(c->h_stage);
        c->h_stage = nullptr;
        c->h_stage_bytes = 0;
    }
    c->G = 0;
    c->gid = c->active_g = c->store_g = c->k_g = nullptr;
    c->cc_g = c->wcc_g = nullptr;
    c->have_graph = c->have_labels = false;
    c->wcc_valid = false;
    c->indices = c->adj = nullptr;
    c->have_snap = false;  // a snapshot belongs to its graph
    c->nh = c->n_tiles = c->n_chunks = 0;
}

// ---------------------------------------------------------------------------
// Profiling: a HIP event pair around each launch on the context's stream.
// ---------------------------------------------------------------------------
static int prof_slot(scd_ctx *c, const char *name) {
    for (size_t i = 0; i < c->prof.size(); i++)
        if (c->prof[i].name == name) return (int)i;
    c->prof.push_back({name, 0, 0.0});
    return (int)c->prof.size() - 1;
}

static hipEvent_t next_event(scd_ctx *c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

struct ProfScope {
    scd_ctx *c;
    int rec = -1;
    hipEvent_t a = nullptr;
    ProfScope(scd_ctx *cc, const char *name) : c(cc) {
        if (!c->profile) return;
        rec = prof_slot(c, name);
        a = next_event(c);
        if (a) (void)hipEventRecord(a, c->stream);
    }
    ~ProfScope() {
        if (!c->profile || !a) return;
        hipEvent_t b = next_event(c);
        if (!b) return;
        (void)hipEventRecord(b, c->stream);
        c->pending.push_back({rec, a, b});
    }
};

static int prof_collect(scd_ctx *c) {
    if (!c->profile) return SCD_OK;
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (auto &pe : c->pending) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, pe.a, pe.b));
        c->prof[pe.rec].launches++;
        c->prof[pe.rec].ms += ms;
    }
    c->pending.clear();
    c->ev_used = 0;
    return SCD_OK;
}

#define LAUNCH(c, name, grid, block, kern, ...)                                          \
    do {                                                                                 \
        if ((grid) > 0) {                                                                \
            ProfScope ps_(c, name);                                                      \
            hipLaunchKernelGGL((kern), dim3(grid), dim3(block), 0, (c)->ls, __VA_ARGS__);     \
        }                                                                                \
    } while (0)

// Lanes are off while profiling (per-kernel events are recorded on the
// context's stream).
static inline bool lanes_on(const scd_ctx *c) {
    return c->use_lanes && c->tune.lanes && c->E >= (1ull << 20) && !c->profile && c->lane[0] && !c->one_stream;
}

// The next unused fork / join event of the current pool (nullptr if the
// runtime cannot create one).
static hipEvent_t lane_event(scd_ctx *c) {
    const int k = c->capturing ? 1 : 0;
    auto &pool = c->lane_ev[k];
    if (c->lane_ev_used[k] == pool.size()) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        pool.push_back(e);
    }
    return pool[c->lane_ev_used[k]++];
}

static void fork_lanes(scd_ctx *c) {
    if (!lanes_on(c)) return;
    hipEvent_t e = lane_event(c);
    if (!e) {  // no event: stay on the context's stream from here on
        c->use_lanes = false;
        return;
    }
    (void)hipEventRecord(e, c->stream);
    for (int i = 0; i < 3; i++) (void)hipStreamWaitEvent(c->lane[i], e, 0);
}

// Subsequent LAUNCHes go to lane i (0 = the context's stream).
static void on_lane(scd_ctx *c, int i) { c->ls = (lanes_on(c) && i > 0) ? c->lane[i - 1] : c->stream; }

static void join_lanes(scd_ctx *c) {
    c->ls = c->stream;
    if (!lanes_on(c)) return;
    for (int i = 0; i < 3; i++) {
        hipEvent_t e = lane_event(c);
        if (!e) {  // join through a synchronization instead (eager launches only)
            (void)hipDeviceSynchronize();
            c->use_lanes = false;
            return;
        }
        (void)hipEventRecord(e, c->lane[i]);
        (void)hipStreamWaitEvent(c->stream, e, 0);
    }
}

// The context's stream waits for the statistics launched on the side
// stream (no-op when none are pending).
static void wait_stats(scd_ctx *c) {
    if (!c->stats_pending) return;
    (void)hipStreamWaitEvent(c->stream, c->stats_ev, 0);
    c->stats_pending = false;
}

// Phase boundary i of an iteration (0: start, 1: after the move, 2: after
// the filter kernels, 3: after the statistics, 4: after the WCC).
constexpr int kStamps = 8;
__global__ void k_stamp(unsigned long long *__restrict__ t, int i) {
    if (threadIdx.x == 0) t[i] = wall_clock64();
}

static void stamp(scd_ctx *c, int i) {
    if (c->phase_timing && c->stamps) hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, c->stream, c->stamps, i);
}

static int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SCD_EDEVICE, "%s: kernel launch failed: %s", what, hipGetErrorString(e));
    return SCD_OK;
}

static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// Degree bins and hub work lists (host-built, uploaded once per graph and
// again when the tuning changes).
// ---------------------------------------------------------------------------
// Degree bins of the resident graph (bins_kernels.hpp) from its CSR offsets
// on the device; the host receives per-workgroup counts, the degree sums and
// the largest degree, and sends back each workgroup's offsets.
// active_only (batched arenas): the vertices of the graphs still refining
// (active_g on the device), else all.
// Label-range buckets of the hubs (tune.hub_table 2), hubs as ordered in
// build_bins (degrees hd): per hub a sample stride that keeps its samples
// within HB_SAMPLES, kstep samples per bucket start (about HB_TARGET entries
// a bucket) and room for 1 + 2 (samples - 1) / kstep buckets (each
// candidate start may add one more bucket after a frequent label), at most
// HB_MAXB; partition chunks of HB_PCH entries.
static int build_hub_buckets(scd_ctx *c, const std::vector<uint32_t> &hd) {
    const uint32_t nh = c->nh;
    std::vector<uint32_t> sig(nh), ks(nh), slot0(nh + 1), slot_hub, ph, ps;
    uint64_t slots = 0;
    for (uint32_t h = 0; h < nh; h++) {
        const uint32_t d = hd[h];
        uint32_t sg = HB_TARGET / 16;  // 16 samples per bucket start: sizes within about 1/4 of the aim
        while ((d + sg - 1) / sg > HB_SAMPLES) sg *= 2;
        const uint32_t m = (d + sg - 1) / sg;
        uint32_t k = std::max<uint32_t>(1, HB_TARGET / sg);
        while (1 + 2 * ((m - 1) / k) > HB_MAXB) k *= 2;
        sig[h] = sg;
        ks[h] = k;
        slot0[h] = (uint32_t)slots;
        const uint32_t nsl = 1 + 2 * ((m - 1) / k);
        slots += nsl;
        for (uint32_t i = 0; i < nsl; i++) slot_hub.push_back(h);
        for (uint32_t st = 0; st < d; st += HB_PCH) {
            ph.push_back(h);
            ps.push_back(st);
        }
    }
    if (slots >= (1ull << 32) / HB_CAP) return fail(SCD_ELIMIT, "hub buckets exceed 2^32 entries");
    slot0[nh] = (uint32_t)slots;
    c->hb_slots = (uint32_t)slots;
    c->hb_pchunks = (uint32_t)ph.size();
    auto up = [&](uint32_t **dst, const std::vector<uint32_t> &src) -> int {
        int r = dalloc(c, dst, src.size(), &c->hub_bufs);
        if (r) return r;
        HIP_TRY(hipMemcpy(*dst, src.data(), sizeof(uint32_t) * src.size(), hipMemcpyHostToDevice));
        return SCD_OK;
    };
    int rc;
    if ((rc = up(&c->hb_sigma, sig)) || (rc = up(&c->hb_kstep, ks)) || (rc = up(&c->hb_slot0, slot0)) ||
        (rc = up(&c->hb_slot_hub, slot_hub)) || (rc = up(&c->hb_pch_hub, ph)) || (rc = up(&c->hb_pch_start, ps)) ||
        (rc = dalloc(c, &c->hb_lo, slots, &c->hub_bufs)) || (rc = dalloc(c, &c->hb_nb, nh, &c->hub_bufs)) ||
        (rc = dalloc(c, &c->hb_cnt, slots, &c->hub_bufs)) || (rc = dalloc(c, &c->hb_ovf_cnt, nh, &c->hub_bufs)) ||
        (rc = dalloc(c, &c->hb_region, slots * HB_CAP, &c->hub_bufs)) ||
        (rc = dalloc(c, &c->hb_bl, slots, &c->hub_bufs)) || (rc = dalloc(c, &c->hb_has_rem, slots, &c->hub_bufs)) ||
        (rc = dalloc(c, &c->hb_bi, slots, &c->hub_bufs)) || (rc = dalloc(c, &c->hb_rem, slots, &c->hub_bufs)))
        return rc;
    return SCD_OK;
}

static int build_bins(scd_ctx *c, bool active_only = false) {
    HIP_TRY(hipStreamSynchronize(c->stream));  // the old lists may still be read
    drop_graphs(c);
    release_tmp(c, c->hub_bufs);
    const uint32_t n = c->n, T = c->tune.lds_tile, C = c->tune.merge_chunk, Wm = c->tune.wave_max;
    const uint32_t *kgid = active_only && c->G ? c->gid : nullptr;
    c->bins_partial = kgid != nullptr;
    const uint32_t nblk = cdiv(n, BIN_CHUNK);
    std::vector<DevBuf> tmp;
    struct TmpGuard {
        scd_ctx *c;
        std::vector<DevBuf> &t;
        ~TmpGuard() { release_tmp(c, t); }
    } guard{c, tmp};
    uint32_t *cnt = nullptr, *offs = nullptr, *dmax = nullptr;
    unsigned long long *esum = nullptr;
    int rc;
    if ((rc = dalloc(c, &cnt, (size_t)nblk * K_NB, &tmp)) || (rc = dalloc(c, &offs, (size_t)nblk * K_NB, &tmp)) ||
        (rc = dalloc(c, &esum, K_NB, &tmp)) || (rc = dalloc(c, &dmax, 1, &tmp)))
        return rc;
    LAUNCH(c, "bins", 1, 64, k_fill_u64v, esum, (uint64_t)K_NB, 0ull);
    LAUNCH(c, "bins", 1, 64, k_fill_u32, dmax, 1u);
    LAUNCH(c, "bins", nblk, 256, k_bin_count, c->indices, n, T, Wm, kgid, c->active_g, cnt, esum, dmax);
    if ((rc = check_launch("degree bins"))) return rc;
    std::vector<uint32_t> h_cnt((size_t)nblk * K_NB), h_off((size_t)nblk * K_NB);
    unsigned long long h_esum[K_NB];
    uint32_t h_dmax = 0;
    HIP_TRY(hipMemcpyAsync(h_cnt.data(), cnt, 4ull * nblk * K_NB, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(h_esum, esum, sizeof(h_esum), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&h_dmax, dmax, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<uint32_t> tot(K_NB, 0);
    for (uint32_t k = 0; k < nblk; k++)
        for (int b = 0; b < K_NB; b++) tot[b] += h_cnt[(size_t)k * K_NB + b];
    c->bin_begin[0] = 0;
    for (int b = 0; b < K_NB; b++) c->bin_begin[b + 1] = c->bin_begin[b] + tot[b];
    for (int b = 0; b < K_NB; b++) {
        uint32_t o = c->bin_begin[b];
        for (uint32_t k = 0; k < nblk; k++) {
            h_off[(size_t)k * K_NB + b] = o;
            o += h_cnt[(size_t)k * K_NB + b];
        }
    }
    for (int b = 0; b < 16; b++) c->bin_edges[b] = b < K_NB ? h_esum[b] : 0;
    c->dmax = h_dmax;
    if ((rc = dalloc(c, &c->bin_verts, n, &c->hub_bufs))) return rc;
    HIP_TRY(hipMemcpyAsync(offs, h_off.data(), 4ull * nblk * K_NB, hipMemcpyHostToDevice, c->stream));
    LAUNCH(c, "bins", nblk, 256, k_bin_scatter, c->indices, n, T, Wm, kgid, c->active_g, offs, c->bin_verts);
    if ((rc = check_launch("degree bins"))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));

    // the WCC's packed lists (every binned vertex may land in one)
    if ((rc = dalloc(c, &c->wq_lo, c->bin_begin[K_NB], &c->hub_bufs)) ||
        (rc = dalloc(c, &c->wq_hi, c->bin_begin[K_NB], &c->hub_bufs)) ||
        (rc = dalloc(c, &c->wq_pcnt, 2, &c->hub_bufs)))
        return rc;
    // internal triangle work queue
    c->wq_M = c->bin_begin[K_NB] - c->bin_begin[K_W1];
    if (c->wq_M) {
        if ((rc = dalloc(c, &c->wq_nch, c->wq_M, &c->hub_bufs)) ||
            (rc = dalloc(c, &c->wq_off, c->wq_M, &c->hub_bufs)) || (rc = dalloc(c, &c->wq_tot, 1, &c->hub_bufs)) ||
            (rc = dalloc(c, &c->wq_next, 1, &c->hub_bufs)) ||
            (rc = dalloc(c, &c->vq_nch, c->wq_M, &c->hub_bufs)) ||
            (rc = dalloc(c, &c->vq_off, c->wq_M, &c->hub_bufs)) || (rc = dalloc(c, &c->vq_tot, 1, &c->hub_bufs)) ||
            (rc = dalloc(c, &c->vq_next, 1, &c->hub_bufs)))
            return rc;
        uint64_t len = c->wq_M;
        for (int l = 0; l < 4; l++) {
            const uint64_t nt = cdiv(len, 2048);
            if ((rc = dalloc(c, &c->wq_sum[l], nt, &c->hub_bufs)) ||
                (rc = dalloc(c, &c->wq_scan[l], nt, &c->hub_bufs)))
                return rc;
            if (nt == 1) break;
            if (l == 3) return fail(SCD_ELIMIT, "work queue scan deeper than 4 levels");
            len = nt;
        }
    }

    // dense-community WCC: rows take at most 2 bytes per adjacency entry
    // (density >= 1/16), so E / 128 units hold the rows of every dense
    // community of a partition (at least 1 MiB, at most 4 GiB)
    c->dn_off = c->dn_vbase = c->dn_cursor = c->dn_lid = c->dn_list = c->dn_ctr = nullptr;
    c->dn_bits = nullptr;
    c->dn_units = 0;
    if (c->tune.dense_wcc == 2 || (c->tune.dense_wcc == 1 && c->E >= (1ull << 20))) {
        const uint64_t units = std::min<uint64_t>(std::max<uint64_t>(c->E / 128, 4096), 1ull << 24);
        if ((rc = dalloc(c, &c->dn_off, n, &c->hub_bufs)) || (rc = dalloc(c, &c->dn_vbase, n, &c->hub_bufs)) ||
            (rc = dalloc(c, &c->dn_cursor, n, &c->hub_bufs)) || (rc = dalloc(c, &c->dn_lid, n, &c->hub_bufs)) ||
            (rc = dalloc(c, &c->dn_list, n, &c->hub_bufs)) || (rc = dalloc(c, &c->dn_ctr, DN_CTR, &c->hub_bufs)) ||
            (rc = dalloc(c, &c->dn_bits, units * DN_UNIT_WORDS, &c->hub_bufs)))
            return rc;
        c->dn_units = (uint32_t)units;
    }

    // hubs
    const uint32_t hb0 = c->bin_begin[K_HUB], nh = c->bin_begin[K_HUB + 1] - hb0;
    c->nh = nh;
    c->n_tiles = c->n_chunks = c->max_hub_deg = 0;
    c->hb_slots = c->hb_pchunks = 0;
    if (nh == 0) return SCD_OK;
    std::vector<uint32_t> hv(nh), hd(nh), hoff(nh), hcb(nh + 1), th, ts, ch, cs;
    {
        uint32_t *d_hd = nullptr;
        if ((rc = dalloc(c, &d_hd, nh, &tmp))) return rc;
        LAUNCH(c, "bins", cdiv(nh, 256), 256, k_bin_degrees, c->indices, c->bin_verts + hb0, nh, d_hd);
        if ((rc = check_launch("degree bins"))) return rc;
        HIP_TRY(hipMemcpyAsync(hv.data(), c->bin_verts + hb0, 4ull * nh, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(hd.data(), d_hd, 4ull * nh, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    // hubs by degree, largest first: the hubs that a merge pass of width w
    // still sorts (degree > w) then own a prefix of the chunk list
    {
        std::vector<uint32_t> ord(nh);
        for (uint32_t h = 0; h < nh; h++) ord[h] = h;
        std::stable_sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) { return hd[x] > hd[y]; });
        std::vector<uint32_t> v2(nh), d2(nh);
        for (uint32_t h = 0; h < nh; h++) {
            v2[h] = hv[ord[h]];
            d2[h] = hd[ord[h]];
        }
        hv.swap(v2);
        hd.swap(d2);
    }
    uint64_t off = 0;
    for (uint32_t h = 0; h < nh; h++) {
        const uint32_t d = hd[h];
        hoff[h] = (uint32_t)off;
        off += d;
        c->max_hub_deg = std::max(c->max_hub_deg, d);
        for (uint32_t s = 0; s < d; s += T) {
            th.push_back(h);
            ts.push_back(s);
        }
        hcb[h] = (uint32_t)ch.size();
        for (uint32_t s = 0; s < d; s += C) {
            ch.push_back(h);
            cs.push_back(s);
        }
    }
    hcb[nh] = (uint32_t)ch.size();
    if (off >= (1ull << 32)) return fail(SCD_ELIMIT, "hub segments exceed 2^32 entries");
    // chunks of each merge pass: those of the hubs with degree above its width
    c->merge_pass_chunks.clear();
    for (uint64_t w = T; w < c->max_hub_deg; w *= 2) {
        uint32_t k = 0;
        while (k < nh && hd[k] > w) k++;
        c->merge_pass_chunks.push_back(hcb[k]);
    }
    c->n_tiles = (uint32_t)th.size();
    c->n_chunks = (uint32_t)ch.size();
    auto up = [&](uint32_t **dst, const std::vector<uint32_t> &src) -> int {
        int r = dalloc(c, dst, src.size(), &c->hub_bufs);
        if (r) return r;
        HIP_TRY(hipMemcpy(*dst, src.data(), sizeof(uint32_t) * src.size(), hipMemcpyHostToDevice));
        return SCD_OK;
    };
    if ((rc = up(&c->hub_v, hv)) || (rc = up(&c->hub_d, hd)) || (rc = up(&c->hub_off, hoff)) ||
        (rc = up(&c->hub_chunk_begin, hcb)) || (rc = up(&c->tile_hub, th)) ||
        (rc = up(&c->tile_start, ts)) || (rc = up(&c->chunk_hub, ch)) || (rc = up(&c->chunk_start, cs)))
        return rc;
    if ((rc = dalloc(c, &c->hub_a, off, &c->hub_bufs)) || (rc = dalloc(c, &c->hub_b, off, &c->hub_bufs)) ||
        (rc = dalloc(c, &c->chunk_cnt, c->n_chunks, &c->hub_bufs)) ||
        (rc = dalloc(c, &c->chunk_bl, c->n_chunks, &c->hub_bufs)) ||
        (rc = dalloc(c, &c->chunk_has_rem, c->n_chunks, &c->hub_bufs)) ||
        (rc = dalloc(c, &c->chunk_bi, c->n_chunks, &c->hub_bufs)) ||
        (rc = dalloc(c, &c->chunk_rem, c->n_chunks, &c->hub_bufs)))
        return rc;

    // hash tables: 2 x degree slots rounded up to a power of two, C per chunk
    std::vector<uint64_t> toff(nh);
    std::vector<uint32_t> tcap(nh), tch, tcs, htb(nh + 1);
    uint64_t slots = 0;
    for (uint32_t hb = 0; hb < nh; hb++) {
        uint32_t cap = 1;
        while (cap < 2 * hd[hb]) cap <<= 1;
        toff[hb] = slots;
        tcap[hb] = cap;
        slots += cap;
        htb[hb] = (uint32_t)tch.size();
        for (uint32_t st = 0; st < cap; st += C) {
            tch.push_back(hb);
            tcs.push_back(st);
        }
    }
    htb[nh] = (uint32_t)tch.size();
    c->n_tchunks = (uint32_t)tch.size();
    if ((rc = dalloc(c, &c->hub_slots, slots, &c->hub_bufs)) || (rc = dalloc(c, &c->tab_off, nh, &c->hub_bufs)) ||
        (rc = up(&c->tab_cap, tcap)) || (rc = up(&c->tchunk_hub, tch)) || (rc = up(&c->tchunk_start, tcs)) ||
        (rc = up(&c->hub_tchunk_begin, htb)) || (rc = dalloc(c, &c->tchunk_bl, c->n_tchunks, &c->hub_bufs)) ||
        (rc = dalloc(c, &c->tchunk_has_rem, c->n_tchunks, &c->hub_bufs)) ||
        (rc = dalloc(c, &c->tchunk_bi, c->n_tchunks, &c->hub_bufs)) ||
        (rc = dalloc(c, &c->tchunk_rem, c->n_tchunks, &c->hub_bufs)))
        return rc;
    HIP_TRY(hipMemcpy(c->tab_off, toff.data(), 8ull * nh, hipMemcpyHostToDevice));
    LAUNCH(c, "hub_table_clear", cdiv(slots, 256), 256, k_fill_u64v, c->hub_slots, slots, HUB_EMPTY);
    if ((rc = check_launch("hub tables"))) return rc;
    return c->tune.hub_table == 2 ? build_hub_buckets(c, hd) : SCD_OK;
}

static HubArgs hub_args(scd_ctx *c) {
    HubArgs h;
    h.hub_v = c->hub_v;
    h.hub_d = c->hub_d;
    h.hub_off = c->hub_off;
    h.hub_chunk_begin = c->hub_chunk_begin;
    h.tile_hub = c->tile_hub;
    h.tile_start = c->tile_start;
    h.chunk_hub = c->chunk_hub;
    h.chunk_start = c->chunk_start;
    h.chunk_bi = c->chunk_bi;
    h.chunk_bl = c->chunk_bl;
    h.chunk_rem = c->chunk_rem;
    h.chunk_has_rem = c->chunk_has_rem;
    return h;
}

static inline uint32_t bin_n(const scd_ctx *c, int b) { return c->bin_begin[b + 1] - c->bin_begin[b]; }
static inline const uint32_t *bin_v(const scd_ctx *c, int b) { return c->bin_verts + c->bin_begin[b]; }

// ---------------------------------------------------------------------------
// Phases
// ---------------------------------------------------------------------------
// The workgroup-per-vertex bins: K_LDS (degrees up to lds_split(T)) with
// the kernel shape for lds_split(T) keys, K_LDS2 with the one for T.
template <int T>
static void launch_lds_and_tiles(scd_ctx *c, const MoveArgs &a, const HubArgs &h, const HubTable &t) {
    constexpr int TH = T >= 8192 ? T / 2 : T;  // lds_split(T)
    MoveArgs a2 = a;
    a2.counters = c->counting ? c->counters + K_LDS2 : nullptr;
    if (!a.ex_sorted && c->tune.lds_hash && T <= 4096) {
        LAUNCH(c, "move_lds2", bin_n(c, K_LDS2), 256, (k_move_lds_hash<T <= 4096 ? T : 4096>), a2, bin_v(c, K_LDS2),
               bin_n(c, K_LDS2));
        LAUNCH(c, "move_lds", bin_n(c, K_LDS), 256, (k_move_lds_hash<TH <= 4096 ? TH : 4096>), a, bin_v(c, K_LDS),
               bin_n(c, K_LDS));
    } else {
        LAUNCH(c, "move_lds2", bin_n(c, K_LDS2), 256, (k_move_lds<T>), a2, bin_v(c, K_LDS2), bin_n(c, K_LDS2));
        LAUNCH(c, "move_lds", bin_n(c, K_LDS), 256, (k_move_lds<TH>), a, bin_v(c, K_LDS), bin_n(c, K_LDS));
    }
    on_lane(c, 1);  // the hub chain: tiles, then merges / table, evaluation, finalize
    if (a.ex_sorted || c->tune.hub_table == 0)
        LAUNCH(c, "hub_tile_sort", c->n_tiles, 256, (k_hub_tile_sort<T>), a, h, c->hub_a);
    else if (c->tune.hub_table == 1)
        LAUNCH(c, "hub_tile_count", c->n_tiles, 256, (k_hub_tile_count<T>), a, h, t);
}

static HubTable hub_table(scd_ctx *c) {
    HubTable t;
    t.slots = c->hub_slots;
    t.tab_off = c->tab_off;
    t.tab_cap = c->tab_cap;
    t.tchunk_hub = c->tchunk_hub;
    t.tchunk_start = c->tchunk_start;
    return t;
}

// Move step for every vertex: old labels (cur) -> labels_out.  Export mode
// writes sorted keys / run lengths instead when ex_sorted is set.
static int phase_move(scd_ctx *c, uint32_t *labels_out, uint32_t *ex_sorted, uint32_t *ex_count) {
    MoveArgs a;
    a.indices = c->indices;
    a.adj = c->adj;
    a.labels = c->cur;
    a.stats = c->stats;
    a.p_ext = c->cc;
    a.labels_out = labels_out;
    a.ex_sorted = ex_sorted;
    a.ex_count = ex_count;
    a.gid = c->G ? c->gid : nullptr;
    a.cc_g = c->cc_g;
    a.active_g = c->active_g;
    auto at = [&](int bin) {
        MoveArgs b = a;
        b.counters = c->counting ? c->counters + bin : nullptr;
        return b;
    };
    // lanes: 3 small-degree bins | 2 middle bins | 1 hub chain | 0 lds bin;
    // the lds bin and the hub chain (the longest) go out first
    fork_lanes(c);
    on_lane(c, 0);
    a.counters = c->counting ? c->counters + K_LDS : nullptr;
    HubArgs h = hub_args(c);
    const HubTable t = hub_table(c);
    switch (c->tune.lds_tile) {
        case 2048: launch_lds_and_tiles<2048>(c, a, h, t); break;
        case 4096: launch_lds_and_tiles<4096>(c, a, h, t); break;
        case 8192: launch_lds_and_tiles<8192>(c, a, h, t); break;
        default: launch_lds_and_tiles<16384>(c, a, h, t); break;
    }
    on_lane(c, 1);
    if (c->nh && !ex_sorted && c->tune.hub_table == 2) {
        HubArgs hb = h;  // results per bucket slot
        hb.hub_chunk_begin = c->hb_slot0;
        hb.chunk_bi = c->hb_bi;
        hb.chunk_bl = c->hb_bl;
        hb.chunk_rem = c->hb_rem;
        hb.chunk_has_rem = c->hb_has_rem;
        HubBuckets k;
        k.sigma = c->hb_sigma;
        k.kstep = c->hb_kstep;
        k.slot0 = c->hb_slot0;
        k.slot_hub = c->hb_slot_hub;
        k.pch_hub = c->hb_pch_hub;
        k.pch_start = c->hb_pch_start;
        k.lo = c->hb_lo;
        k.nb = c->hb_nb;
        k.cnt = c->hb_cnt;
        k.ovf_cnt = c->hb_ovf_cnt;
        k.region = c->hb_region;
        k.ovf = c->hub_b;
        k.cap = c->hb_cap;
        LAUNCH(c, "hub_split", c->nh, 256, k_hb_split, a, hb, k);
        LAUNCH(c, "hub_part", c->hb_pchunks, 256, k_hb_part, a, hb, k);
        LAUNCH(c, "hub_eval", c->hb_slots, 256, k_hb_eval, at(K_HUB), hb, k);
        LAUNCH(c, "hub_finalize", cdiv(c->nh, 4), 256, k_hub_finalize, a, hb, c->nh);
    } else if (c->nh && !ex_sorted && c->tune.hub_table) {
        HubArgs ht = h;  // results per table chunk
        ht.hub_chunk_begin = c->hub_tchunk_begin;
        ht.chunk_bi = c->tchunk_bi;
        ht.chunk_bl = c->tchunk_bl;
        ht.chunk_rem = c->tchunk_rem;
        ht.chunk_has_rem = c->tchunk_has_rem;
        LAUNCH(c, "hub_eval", c->n_tchunks, 256, k_hub_table_eval, at(K_HUB), ht, t, c->tune.merge_chunk);
        LAUNCH(c, "hub_finalize", cdiv(c->nh, 4), 256, k_hub_finalize, a, ht, c->nh);
    } else if (c->nh) {
        // merge passes over the hubs still longer than the pass width; each
        // hub's sorted keys end in hub_a or hub_b by the parity of its own
        // pass count (k_hub_eval picks)
        const uint32_t C = c->tune.merge_chunk;
        uint32_t *src = c->hub_a, *dst = c->hub_b;
        uint64_t w = c->tune.lds_tile;
        for (uint32_t chunks : c->merge_pass_chunks) {
            LAUNCH(c, "hub_merge", chunks, 256, k_hub_merge, h, src, dst, (uint32_t)w, C);
            std::swap(src, dst);
            w *= 2;
        }
        LAUNCH(c, "hub_eval", c->n_chunks, 256, k_hub_eval, at(K_HUB), h, c->hub_a, c->hub_b, c->tune.lds_tile, C);
        LAUNCH(c, "hub_finalize", cdiv(c->nh, 4), 256, k_hub_finalize, a, h, c->nh);
    }
    on_lane(c, 3);
    LAUNCH(c, "move_sub8", cdiv(bin_n(c, K_SUB8), 32), 256, k_move_sub<8>, at(K_SUB8), bin_v(c, K_SUB8), bin_n(c, K_SUB8));
    LAUNCH(c, "move_sub16", cdiv(bin_n(c, K_SUB16), 16), 256, k_move_sub<16>, at(K_SUB16), bin_v(c, K_SUB16), bin_n(c, K_SUB16));
    LAUNCH(c, "move_sub32", cdiv(bin_n(c, K_SUB32), 8), 256, k_move_sub<32>, at(K_SUB32), bin_v(c, K_SUB32), bin_n(c, K_SUB32));
    LAUNCH(c, "move_wave1", cdiv(bin_n(c, K_W1), 4), 256, k_move_wave<1>, at(K_W1), bin_v(c, K_W1), bin_n(c, K_W1));
    on_lane(c, 2);
    LAUNCH(c, "move_wave2", cdiv(bin_n(c, K_W2), 4), 256, k_move_wave<2>, at(K_W2), bin_v(c, K_W2), bin_n(c, K_W2));
    LAUNCH(c, "move_wave4", cdiv(bin_n(c, K_W4), 4), 256, k_move_wave<4>, at(K_W4), bin_v(c, K_W4), bin_n(c, K_W4));
    LAUNCH(c, "move_wave8", cdiv(bin_n(c, K_W8), 4), 256, k_move_wave<8>, at(K_W8), bin_v(c, K_W8), bin_n(c, K_W8));
    LAUNCH(c, "move_wave16", cdiv(bin_n(c, K_W16), 4), 256, k_move_wave<16>, at(K_W16), bin_v(c, K_W16), bin_n(c, K_W16));
    join_lanes(c);
    return check_launch("move");
}

// Statistics + internal adjacency of the current labels.  iteration: within
// an iteration (the phase boundary after the filter kernels is stamped, and
// without phase stamps the statistics run on the side stream, beside the
// WCC's triangle counting; phase_wcc waits for them before the first kernel
// that reads them).
static int phase_stats(scd_ctx *c, bool iteration = false) {
    StatArgs s;
    s.indices = c->indices;
    s.adj = c->adj;
    s.labels = c->cur;
    s.stats = c->stats;
    s.int_adj = c->int_adj;
    s.int_deg = c->int_deg;
    s.iflag = c->iflag;
    s.gid = c->G ? c->gid : nullptr;
    s.active_g = c->active_g;
    s.counters = c->counting ? c->counters + K_NB : nullptr;
    // lanes: 3 small-degree bins | 2 middle bins | 1 hubs | 0 lds bin
    fork_lanes(c);
    on_lane(c, 3);
    LAUNCH(c, "filter_sub8", cdiv(bin_n(c, K_SUB8), 32), 256, k_filter_sub<8>, s, bin_v(c, K_SUB8), bin_n(c, K_SUB8));
    LAUNCH(c, "filter_sub16", cdiv(bin_n(c, K_SUB16), 16), 256, k_filter_sub<16>, s, bin_v(c, K_SUB16), bin_n(c, K_SUB16));
    LAUNCH(c, "filter_sub32", cdiv(bin_n(c, K_SUB32), 8), 256, k_filter_sub<32>, s, bin_v(c, K_SUB32), bin_n(c, K_SUB32));
    LAUNCH(c, "filter_wave1", cdiv(bin_n(c, K_W1), 4), 256, k_filter_wave<1>, s, bin_v(c, K_W1), bin_n(c, K_W1));
    on_lane(c, 2);
    LAUNCH(c, "filter_wave2", cdiv(bin_n(c, K_W2), 4), 256, k_filter_wave<2>, s, bin_v(c, K_W2), bin_n(c, K_W2));
    LAUNCH(c, "filter_wave4", cdiv(bin_n(c, K_W4), 4), 256, k_filter_wave<4>, s, bin_v(c, K_W4), bin_n(c, K_W4));
    LAUNCH(c, "filter_wave8", cdiv(bin_n(c, K_W8), 4), 256, k_filter_wave<8>, s, bin_v(c, K_W8), bin_n(c, K_W8));
    LAUNCH(c, "filter_wave16", cdiv(bin_n(c, K_W16), 4), 256, k_filter_wave<16>, s, bin_v(c, K_W16), bin_n(c, K_W16));
    on_lane(c, 0);
    // K_LDS2 (degrees above lds_split) with the shape for lds_tile keys,
    // K_LDS with the one for lds_split
    switch (c->tune.lds_tile) {
        case 2048:
            LAUNCH(c, "filter_block", bin_n(c, K_LDS), 256, k_filter_block<2048>, s, bin_v(c, K_LDS), bin_n(c, K_LDS));
            break;
        case 4096:
            LAUNCH(c, "filter_block", bin_n(c, K_LDS), 256, k_filter_block<4096>, s, bin_v(c, K_LDS), bin_n(c, K_LDS));
            break;
        case 8192:
            LAUNCH(c, "filter_block2", bin_n(c, K_LDS2), 256, k_filter_block<8192>, s, bin_v(c, K_LDS2), bin_n(c, K_LDS2));
            LAUNCH(c, "filter_block", bin_n(c, K_LDS), 256, k_filter_block<4096>, s, bin_v(c, K_LDS), bin_n(c, K_LDS));
            break;
        default:
            LAUNCH(c, "filter_block2", bin_n(c, K_LDS2), 256, k_filter_block<16384>, s, bin_v(c, K_LDS2), bin_n(c, K_LDS2));
            LAUNCH(c, "filter_block", bin_n(c, K_LDS), 256, k_filter_block<8192>, s, bin_v(c, K_LDS), bin_n(c, K_LDS));
            break;
    }
    if (c->nh) {
        on_lane(c, 1);
        const uint32_t C = c->tune.merge_chunk;
        LAUNCH(c, "hub_filter_compact", c->n_chunks, 256, k_hub_filter_compact, s, c->hub_v, c->hub_d, c->hub_off,
               c->chunk_hub, c->chunk_start, C, c->chunk_cnt, c->hub_a);
        LAUNCH(c, "hub_filter_place", c->n_chunks, 256, k_hub_filter_place, s, c->hub_v, c->hub_off,
               c->hub_chunk_begin, c->chunk_hub, c->chunk_start, c->chunk_cnt, c->hub_a);
    }
    join_lanes(c);
    if (iteration) stamp(c, 2);
    hipEvent_t e0 = (iteration && !c->phase_timing && c->side && lanes_on(c)) ? lane_event(c) : nullptr;
    hipEvent_t e1 = e0 ? lane_event(c) : nullptr;
    if (e1) {
        (void)hipEventRecord(e0, c->stream);
        (void)hipStreamWaitEvent(c->side, e0, 0);
        c->ls = c->side;
    }
    LAUNCH(c, "clear_stats", cdiv(c->num_comms, 256), 256, k_fill_int4, c->stats, c->num_comms);
    LAUNCH(c, "stats_agg", std::min<uint32_t>(cdiv(c->n, 1024), 2u * c->num_cus), 1024, k_stats_agg, c->indices,
           c->cur, c->int_deg, c->n, c->G ? c->gid : nullptr, c->active_g, c->stats);
    if (e1) {
        (void)hipEventRecord(e1, c->side);
        c->stats_ev = e1;
        c->stats_pending = true;
        c->ls = c->stream;
    }
    return check_launch("stats");
}

// Exclusive scan of n uint32 values (in -> out, no aliasing) on the stream
// without host involvement: tile sums, scan of the tile sums (one workgroup
// up to 2048 tiles, else recursively into the level buffers preallocated in
// wq_sum / wq_scan), tile scans; the total goes to total_dev when given.  Capturable: no host copies, no
// memset nodes.
static int scan_u32_dev(scd_ctx *c, const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *total_dev, int level) {
    const uint32_t nt = cdiv(n, 2048);
    uint32_t *sums = c->wq_sum[level], *scan = c->wq_scan[level];
    LAUNCH(c, "scan_tile_sums", nt, 256, k_scan_tile_sums, in, n, sums);
    if (nt <= 2048) {  // the tile sums' scan and the total in one workgroup
        LAUNCH(c, "scan_top", 1, 256, k_scan_top, sums, nt, scan, total_dev);
    } else {
        int rc = scan_u32_dev(c, sums, scan, nt, nullptr, level + 1);
        if (rc) return rc;
        if (total_dev) LAUNCH(c, "scan_total", 1, 64, k_scan_total, scan, sums, nt, total_dev);
    }
    LAUNCH(c, "scan_tiles", nt, 256, k_scan_tiles, in, n, scan, out);
    return SCD_OK;
}

// Triangles of the subgraph given by w.int_adj / w.int_deg (the internal
// subgraph of the current labels, or with int_deg == nullptr the whole
// graph): T[u] = triangles at u, td[u] = edges at u in a triangle; the
// kernels per degree bin and the work queue of long out-lists.
// fold_hot: fold the hot ids' private copies into T / td at the end (the
// WCC's terms kernel folds them instead).
static int count_internal_triangles(scd_ctx *c, const WccArgs &w0, bool fold_hot = true) {
    WccArgs w = w0;
    // dense communities of the WCC's internal subgraph: classified, local
    // ids and rows, counted on their own lane beside the list kernels
    // (graphs of at least 2^20 adjacency entries: below that the four extra
    // launches cost more than the list kernels they relieve)
    const bool dense = c->dn_bits && w.int_deg && c->num_comms;
    DenseArgs d;
    if (dense) {
        d.labels = w.labels;
        d.stats = w.stats;
        d.dirty = w.dirty;
        d.off = c->dn_off;
        d.vbase = c->dn_vbase;
        d.cursor = c->dn_cursor;
        d.lid = c->dn_lid;
        d.list = c->dn_list;
        d.ctr = c->dn_ctr;
        d.bits = c->dn_bits;
        d.units = c->dn_units;
        LAUNCH(c, "dn_clear", 1, 64, k_fill_u32, c->dn_ctr, DN_CTR);
        wait_stats(c);  // the classification reads the community sizes
        LAUNCH(c, "dn_classify", cdiv(c->num_comms, 256), 256, k_dn_classify, d, c->num_comms);
        LAUNCH(c, "dn_lid", cdiv(c->n, 256), 256, k_dn_lid, w, d, c->n);
        w.dn_off = c->dn_off;
    }
    // the WCC's internal subgraph: its vertices mostly have no or a few
    // internal out-neighbours whatever their degree bin, so the counting
    // kernels walk lists of those that have some (k_wtc_pack) instead of
    // S lanes or a wave per binned vertex
    const bool pack = w.int_deg != nullptr;
    const uint32_t nb = c->bin_begin[K_NB];  // binned vertices
    if (pack) w.pack_cnt = c->wq_pcnt;
    const uint32_t gcap = 8 * c->num_cus;  // grids striding over a packed list
    LAUNCH(c, "wcc_dplus", cdiv(c->n, c->block_size), c->block_size, k_wtc_dplus, w, c->n);
    // lanes: 3 packed lists (the whole graph: small-degree groups) | 2
    // out-lists up to 64 (whole graph) | 1 dense communities | 0 work queue
    fork_lanes(c);
    if (dense) {
        on_lane(c, 1);
        // persistent grids: up to 32 waves per CU, no more than a wave per vertex
        const uint32_t grid = std::min<uint32_t>(8 * c->num_cus, cdiv(c->n, 4));
        LAUNCH(c, "dn_rows", grid, 256, k_dn_rows, w, d);
        LAUNCH(c, "dn_count", grid, 256, k_dn_count, w, d);
        LAUNCH(c, "dn_halve", std::min<uint32_t>(4 * c->num_cus, cdiv(c->n, 256)), 256, k_dn_halve, w, d);
    }
    on_lane(c, 3);
    if (pack) {
        LAUNCH(c, "wcc_tri_pack", cdiv(nb, WTC_PACK_NT), WTC_PACK_NT, k_wtc_pack, w, c->bin_verts, nb, c->wq_lo,
               c->wq_hi);
        LAUNCH(c, "wcc_tri_lo", std::min(cdiv(nb, 32), gcap), 256, k_wtc_group<8>, w, c->wq_lo, nb, c->wq_pcnt);
        LAUNCH(c, "wcc_tri_hi", std::min(cdiv(nb, 4), gcap), 256, k_wtc_group<64>, w, c->wq_hi, nb, c->wq_pcnt + 1);
    } else {
        LAUNCH(c, "wcc_tri_sub8", cdiv(bin_n(c, K_SUB8), 32), 256, k_wtc_group<8>, w, bin_v(c, K_SUB8),
               bin_n(c, K_SUB8), nullptr);
        LAUNCH(c, "wcc_tri_sub16", cdiv(bin_n(c, K_SUB16), 16), 256, k_wtc_group<16>, w, bin_v(c, K_SUB16),
               bin_n(c, K_SUB16), nullptr);
        LAUNCH(c, "wcc_tri_sub32", cdiv(bin_n(c, K_SUB32), 8), 256, k_wtc_group<32>, w, bin_v(c, K_SUB32),
               bin_n(c, K_SUB32), nullptr);
    }
    if (c->wq_M) {
        const uint32_t M = c->wq_M;
        if (!pack) {
            on_lane(c, 2);
            LAUNCH(c, "wcc_tri_wave", cdiv(M, 4), 256, k_wtc_group<64>, w, bin_v(c, K_W1), M, nullptr);
        }
        on_lane(c, 0);
        LAUNCH(c, "wcc_tri_chunks", cdiv(M, 256), 256, k_wtc_count_chunks<false>, w, bin_v(c, K_W1), M, c->wq_nch,
               c->wq_next);
        int rc = scan_u32_dev(c, c->wq_nch, c->wq_off, M, c->wq_tot, 0);
        if (rc) return rc;
        LAUNCH(c, "wcc_tri_queue", 4 * c->num_cus, WTC_QT, (k_wtc_queue<2048, WTC_QT, 4, false>), w, bin_v(c, K_W1),
               M, c->wq_off, c->wq_tot, c->wq_next);
        if (w.vc_min) {  // edges counted from the longer out-list's side
            LAUNCH(c, "wcc_tri_chunks", cdiv(M, 256), 256, k_wtc_count_chunks<true>, w, bin_v(c, K_W1), M,
                   c->vq_nch, c->vq_next);
            if ((rc = scan_u32_dev(c, c->vq_nch, c->vq_off, M, c->vq_tot, 0))) return rc;
            LAUNCH(c, "wcc_tri_vqueue", 4 * c->num_cus, WTC_QT, (k_wtc_queue<2048, WTC_QT, 4, true>), w,
                   bin_v(c, K_W1), M, c->vq_off, c->vq_tot, c->vq_next);
        }
    }
    join_lanes(c);
    // edge marks are complete: count them, again on lanes
    fork_lanes(c);
    if (pack) {
        on_lane(c, 3);
        LAUNCH(c, "wcc_tri_final", std::min(cdiv(nb, 32), gcap), 256, k_wtc_final<8>, w, c->wq_lo, nb, 0, c->wq_pcnt);
        on_lane(c, 2);
        LAUNCH(c, "wcc_tri_final", std::min(cdiv(nb, 4), gcap), 256, k_wtc_final<64>, w, c->wq_hi, nb, 0,
               c->wq_pcnt + 1);
    } else {
        // the whole graph: bounded grids, hub counts summed in LDS per workgroup
        const uint32_t fin_max = WTC_FIN_BLOCKS_PER_CU * c->num_cus;
        on_lane(c, 3);
        LAUNCH(c, "wcc_tri_final", std::min(cdiv(bin_n(c, K_SUB8), 32), fin_max), 256, k_wtc_final<8>, w,
               bin_v(c, K_SUB8), bin_n(c, K_SUB8), 1, nullptr);
        on_lane(c, 2);
        LAUNCH(c, "wcc_tri_final", std::min(cdiv(bin_n(c, K_SUB16), 16), fin_max), 256, k_wtc_final<16>, w,
               bin_v(c, K_SUB16), bin_n(c, K_SUB16), 1, nullptr);
        on_lane(c, 1);
        LAUNCH(c, "wcc_tri_final", std::min(cdiv(bin_n(c, K_SUB32), 8), fin_max), 256, k_wtc_final<32>, w,
               bin_v(c, K_SUB32), bin_n(c, K_SUB32), 1, nullptr);
        on_lane(c, 0);
        LAUNCH(c, "wcc_tri_final", std::min(cdiv(c->wq_M, 4), fin_max), 256, k_wtc_final<64>, w, bin_v(c, K_W1),
               c->wq_M, 1, nullptr);
    }
    join_lanes(c);
    if (fold_hot) LAUNCH(c, "wcc_tri_hot", cdiv(w.hot_n, 256), 256, k_wtc_hot_sum, w);
    return SCD_OK;
}

// WCC terms of the current labels and their mean into wcc_dev.
static int phase_wcc(scd_ctx *c, int32_t *int_tri_out, int32_t *int_tri_deg_out, bool incremental = false) {
    WccArgs w;
    w.dirty = incremental ? c->wdirty : nullptr;
    w.indices = c->indices;
    w.labels = c->cur;
    w.stats = c->stats;
    w.int_adj = c->int_adj;
    w.int_deg = c->int_deg;
    w.tri = c->tri;
    w.tri_deg = c->tri_deg;
    w.dplus = c->wdplus;
    w.T = c->wT;
    w.td = c->wtd;
    w.flag = c->iflag;
    w.terms = c->terms;
    w.T_hot = c->hot_T;
    w.td_hot = c->hot_td;
    w.hot_n = c->hot_n;
    w.vc_min = c->wcc_heavy;
    w.int_tri_out = int_tri_out;
    w.int_tri_deg_out = int_tri_deg_out;
    int rc = count_internal_triangles(c, w, false);
    wait_stats(c);  // the terms read the community sizes
    if (rc) return rc;
    LAUNCH(c, "wcc_terms", cdiv(c->n, c->block_size), c->block_size, k_wcc_terms, w, c->n);
    if (c->G) {
        LAUNCH(c, "wcc_sum_partial", c->n_witems, 256, k_reduce_partial_g, c->terms, c->witem_start, c->witem_end,
               c->partial_g);
        LAUNCH(c, "wcc_sum_final", c->G, 256, k_reduce_final_g, c->partial_g, c->g_item_begin, c->vtx_off_d, c->wcc_g);
    } else {
        const uint32_t np = cdiv(c->n, 4096);
        LAUNCH(c, "wcc_sum_partial", np, 256, k_reduce_partial, c->terms, c->n, c->partial);
        LAUNCH(c, "wcc_sum_final", 1, 256, k_reduce_final, c->partial, np, c->n, c->wcc_dev);
    }
    return check_launch("wcc");
}

static int clear_counters(scd_ctx *c) {
    LAUNCH(c, "clear_counters", 1, 64, k_fill_u64, c->counters, 16u);
    return check_launch("clear counters");
}

static int finish(scd_ctx *c) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->lane_ev_used[0] = 0;  // every recorded fork / join event has completed
    return prof_collect(c);
}

// ---------------------------------------------------------------------------
// Static scratch for scd_count_triangles_noalloc: one array per GPU, held
// until the process ends.  Host code never uses the array's name as an
// address: scd_init looks up each device's copy once, with that device
// current, and the entry point works with that pointer.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = 64ull << 20;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kScratchDevices];
static hipError_t g_scratch_err[kScratchDevices];

// Device `dev` is current.
static void scratch_setup(int dev) {
    if (dev < 0 || dev >= kScratchDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[dev] = e;
    if (e == hipSuccess) g_scratch_ptr[dev] = p;
    else (void)hipGetLastError();
}

static uint8_t *scratch_of(int dev) {
    if (dev < 0 || dev >= kScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (uint8_t *)g_scratch_ptr[dev];
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int scd_device_count(void) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e == hipErrorNoDevice) return 0;
    if (e != hipSuccess) return fail(SCD_EDEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
    return n;
}

extern "C" int scd_init(int device_id, int work_group_size, int profile, scd_ctx **out) {
    if (!out) return fail(SCD_EINVAL, "scd_init: null output");
    *out = nullptr;
    if (work_group_size < 1 || work_group_size > 1024)
        return fail(SCD_EINVAL, "scd_init: work group size %d is not in [1, 1024]", work_group_size);
    int count = scd_device_count();
    if (count < 0) return count;
    if (device_id < 0 || device_id >= count)
        return fail(SCD_EDEVICE, "scd_init: device %d is not one of the %d visible HIP devices",
                    device_id, count);
    HIP_TRY(hipSetDevice(device_id));
    scratch_setup(device_id);
    scd_ctx *c = new scd_ctx();
    c->device = device_id;
    c->block_size = work_group_size;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device_id) == hipSuccess && prop.multiProcessorCount > 0) {
            c->num_cus = (uint32_t)prop.multiProcessorCount;
            c->arch = prop.gcnArchName;
            const size_t colon = c->arch.find(':');
            if (colon != std::string::npos) c->arch.resize(colon);
        }
    }
    c->profile = profile != 0;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(SCD_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    c->ls = c->stream;
    {
        bool ok = true;
        for (int i = 0; i < 3 && ok; i++) ok = hipStreamCreateWithFlags(&c->lane[i], hipStreamNonBlocking) == hipSuccess;
        if (ok && hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            c->side = nullptr;  // statistics stay on the context's stream
        }
        if (!ok) {  // run everything on the one stream
            (void)hipGetLastError();
            c->use_lanes = false;
        }
        // SCD_LANES=0: one stream whatever the tuning says (tools that trace
        // kernels may not follow work forked across captured streams)
        if (const char *env = getenv("SCD_LANES"))
            if (atoi(env) == 0) c->use_lanes = false;
    }
    // SCD_HUB_BUCKET_CAP: entries per hub bucket region below HB_CAP, so that
    // tests reach the buckets' overflow path
    if (const char *env = getenv("SCD_HUB_BUCKET_CAP")) {
        const long v = atol(env);
        if (v >= 1 && v < (long)HB_CAP) c->hb_cap = (uint32_t)v;
    }
    e = hipHostMalloc((void **)&c->wcc_host, 512, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(SCD_ENOMEM, "hipHostMalloc: %s", hipGetErrorString(e));
    }
    c->counters_host = (unsigned long long *)((char *)c->wcc_host + 64);
    c->chk_host = (uint32_t *)((char *)c->wcc_host + 192);
    c->stamps_host = (unsigned long long *)((char *)c->wcc_host + 256);
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device_id) == hipSuccess && khz > 0)
            c->wall_clock_khz = khz;
    }
    for (int i = 0; i < 16; i++) c->counters_host[i] = 0;
    *out = c;
    return SCD_OK;
}

extern "C" void scd_destroy(scd_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_graph(c, false);
    free_list(c, c->cache);
    free_list(c, c->tmp_cache);
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    for (int i = 0; i < 3; i++)
        if (c->lane[i]) (void)hipStreamDestroy(c->lane[i]);
    if (c->side) (void)hipStreamDestroy(c->side);
    for (auto &pool : c->lane_ev)
        for (auto e : pool) (void)hipEventDestroy(e);
    if (c->stamps) (void)hipFree(c->stamps);
    (void)hipHostFree(c->wcc_host);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

static int alloc_graph(scd_ctx *c, uint32_t n, uint32_t m, double clustering_coef);
static int finish_upload(scd_ctx *c);
template <typename T>
static int upload_tmp(scd_ctx *c, const T *h, uint32_t count, T **d, std::vector<DevBuf> &tmp);

// Input checks on the device before any kernel indexes by adjacency entries
// or labels: check_begin clears the check word, the check kernels follow on
// the stream, check_end waits and turns a violation into SCD_EINVAL (the
// half-uploaded graph is dropped).
static int check_begin(scd_ctx *c) {
    LAUNCH(c, "check_clear", 1, 64, k_fill_u32, c->chk, 1u);
    return check_launch("input check");
}

static int check_end(scd_ctx *c, const char *what) {
    int rc = check_launch(what);
    if (rc == SCD_OK) {
        hipError_t e = hipMemcpyAsync(c->chk_host, c->chk, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(SCD_EDEVICE, "%s: %s", what, hipGetErrorString(e));
        else if (c->chk_host[0]) rc = fail(SCD_EINVAL, "%s", what);
    }
    if (rc) free_graph(c);
    return rc;
}

static int check_adjacency(scd_ctx *c, const char *what) {
    int rc = check_begin(c);
    if (rc) return rc;
    if (c->E)
        LAUNCH(c, "check_adj", std::min<uint32_t>(cdiv(c->E, 256), 64u * c->num_cus), 256, k_check_adj, c->adj, c->E,
               c->n, c->chk);
    return check_end(c, what);
}

extern "C" int scd_upload(scd_ctx *c, uint32_t n, uint32_t m, const uint32_t *indices,
                          const uint32_t *adj, const int32_t *tri, const int32_t *tri_deg,
                          double clustering_coef) {
    if (!c) return fail(SCD_EINVAL, "scd_upload: null context");
    if (n == 0) return fail(SCD_EINVAL, "scd_upload: graph has no vertices");
    if (!indices || (m && !adj) || !tri || !tri_deg) return fail(SCD_EINVAL, "scd_upload: null array");
    if (m >= (1u << 31)) return fail(SCD_ELIMIT, "scd_upload: m = %u needs 2m < 2^32", m);
    const uint64_t E = 2ull * m;
    if (indices[0] != 0 || indices[n] != E)
        return fail(SCD_EINVAL, "scd_upload: indices[0] = %u, indices[n] = %u, expected 0 and %llu",
                    indices[0], indices[n], (unsigned long long)E);
    for (uint32_t v = 0; v < n; v++)
        if (indices[v + 1] < indices[v]) return fail(SCD_EINVAL, "scd_upload: indices decrease at %u", v);
    int rc = alloc_graph(c, n, m, clustering_coef);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(c->indices, indices, sizeof(uint32_t) * ((size_t)n + 1), hipMemcpyHostToDevice, c->stream));
    if (E) HIP_TRY(hipMemcpyAsync(c->adj, adj, sizeof(uint32_t) * E, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->tri, tri, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->tri_deg, tri_deg, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
    if ((rc = check_adjacency(c, "scd_upload: an adjacency entry is not a vertex id (>= n)"))) return rc;
    return finish_upload(c);
}

// Private counter copies of the triangle kernels (wtc_hot_slot).
static uint32_t hot_words(uint32_t n) {
    return WTC_HOT_COPIES * std::min(n, WTC_HOT_MAX) + WTC_HOT1_COPIES * WTC_HOT1_N;
}

// Frees the previous graph and allocates every per-vertex and per-entry
// buffer for n vertices and m edges.
static int alloc_graph(scd_ctx *c, uint32_t n, uint32_t m, double clustering_coef) {
    const uint64_t E = 2ull * m;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    free_graph(c);
    c->n = n;
    c->m = m;
    c->E = E;
    c->cc = clustering_coef;
    int rc;
    if ((rc = dalloc(c, &c->indices, (size_t)n + 1)) || (rc = dalloc(c, &c->adj, E)) ||
        (rc = dalloc(c, &c->tri, n)) || (rc = dalloc(c, &c->tri_deg, n)) ||
        (rc = dalloc(c, &c->cur, n)) || (rc = dalloc(c, &c->next, n)) || (rc = dalloc(c, &c->best, n)) ||
        (rc = dalloc(c, &c->stats, n)) || (rc = dalloc(c, &c->int_adj, E)) ||
        (rc = dalloc(c, &c->int_deg, n)) || (rc = dalloc(c, &c->terms, n)) ||
        (rc = dalloc(c, &c->iflag, E)) || (rc = dalloc(c, &c->wdplus, n)) || (rc = dalloc(c, &c->wT, n)) ||
        (rc = dalloc(c, &c->wtd, n)) || (rc = dalloc(c, &c->wdirty, n)) ||
        (rc = dalloc(c, &c->hot_T, hot_words(n))) || (rc = dalloc(c, &c->hot_td, hot_words(n))) ||
        (rc = dalloc(c, &c->partial, cdiv(n, 4096))) || (rc = dalloc(c, &c->wcc_dev, 1)) ||
        (rc = dalloc(c, &c->counters, 16)) || (rc = dalloc(c, &c->snap_labels, n)) ||
        (rc = dalloc(c, &c->snap_stats, n)) || (rc = dalloc(c, &c->chk, 64))) {
        free_graph(c);
        return rc;
    }
    c->lab_a = c->cur;
    c->lab_b = c->next;
    c->hot_n = std::min(n, WTC_HOT_MAX);
    LAUNCH(c, "hot_clear", cdiv(hot_words(n), 256), 256, k_fill_u32, c->hot_T, hot_words(n));
    LAUNCH(c, "hot_clear", cdiv(hot_words(n), 256), 256, k_fill_u32, c->hot_td, hot_words(n));
    return check_launch("allocation");
}

// After the graph arrays are on the device (enqueued on the stream): host
// degrees from the host copy of indices, degree bins.
static int finish_upload(scd_ctx *c) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    int rc = build_bins(c);
    if (rc) {
        free_graph(c);
        return rc;
    }
    c->have_graph = true;
    return SCD_OK;
}

static int need_graph(scd_ctx *c, const char *fn) {
    if (!c) return fail(SCD_EINVAL, "%s: null context", fn);
    if (!c->have_graph) return fail(SCD_ESTATE, "%s: no graph uploaded", fn);
    (void)hipSetDevice(c->device);
    return SCD_OK;
}

static int need_labels(scd_ctx *c, const char *fn) {
    int rc = need_graph(c, fn);
    if (rc) return rc;
    if (!c->have_labels) return fail(SCD_ESTATE, "%s: no labels on the device", fn);
    return SCD_OK;
}

extern "C" int scd_labels_to_device(scd_ctx *c, const uint32_t *labels) {
    int rc = need_graph(c, "scd_labels_to_device");
    if (rc) return rc;
    if (!labels) return fail(SCD_EINVAL, "scd_labels_to_device: null labels");
    uint32_t mx = 0;
    for (uint32_t v = 0; v < c->n; v++) {
        if (labels[v] == SCD_INVALID_LABEL)
            return fail(SCD_EINVAL, "scd_labels_to_device: vertex %u has the INVALID label", v);
        mx = std::max(mx, labels[v]);
    }
    if (mx >= c->n) return fail(SCD_EINVAL, "scd_labels_to_device: label %u >= n = %u", mx, c->n);
    if (c->num_comms != mx + 1) drop_graphs(c);
    c->num_comms = mx + 1;
    HIP_TRY(hipMemcpyAsync(c->cur, labels, sizeof(uint32_t) * c->n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->best, labels, sizeof(uint32_t) * c->n, hipMemcpyHostToDevice, c->stream));
    if ((rc = clear_counters(c)) || (rc = phase_stats(c)) || (rc = finish(c))) return rc;
    c->have_labels = true;
    c->wcc_valid = false;
    return SCD_OK;
}

// ---------------------------------------------------------------------------
// A caller's partition (the CPU call surface, src/scd.hpp:25-31): labels
// with INVALID allowed, the caller's statistics, one move step alone, the
// WCC of given labels and sizes.
// ---------------------------------------------------------------------------
__global__ void k_stats_pack(const int32_t *__restrict__ sizes, const int32_t *__restrict__ ext,
                             const int32_t *__restrict__ in, uint32_t nc, int4 *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nc) out[i] = make_int4(sizes[i], ext[i], in[i], 0);
}

__global__ void k_stats_sizes(const int32_t *__restrict__ sizes, uint32_t nc, int4 *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nc) out[i].x = sizes[i];
}

// labels -> current and best labels after the checks of a caller's partition.
static int put_labels(scd_ctx *c, const char *fn, const uint32_t *labels, uint32_t num_comms) {
    if (!labels) return fail(SCD_EINVAL, "%s: null labels", fn);
    if (num_comms > c->n) return fail(SCD_EINVAL, "%s: %u communities for n = %u vertices", fn, num_comms, c->n);
    for (uint32_t v = 0; v < c->n; v++)
        if (labels[v] != SCD_INVALID_LABEL && labels[v] >= num_comms)
            return fail(SCD_EINVAL, "%s: vertex %u has label %u, not below %u communities", fn, v, labels[v],
                        num_comms);
    if (c->num_comms != num_comms) drop_graphs(c);
    c->num_comms = num_comms;
    HIP_TRY(hipMemcpyAsync(c->cur, labels, sizeof(uint32_t) * c->n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->best, labels, sizeof(uint32_t) * c->n, hipMemcpyHostToDevice, c->stream));
    c->have_labels = true;
    c->wcc_valid = false;
    return SCD_OK;
}

// Uploads count entries of each of the given arrays to device temporaries.
static int put_stats(scd_ctx *c, uint32_t nc, const int32_t *sizes, const int32_t *ext, const int32_t *in,
                     std::vector<DevBuf> &tmp) {
    int32_t *ds = nullptr, *de = nullptr, *di = nullptr;
    int rc;
    if ((rc = upload_tmp(c, sizes, std::max<uint32_t>(nc, 1), &ds, tmp))) return rc;
    if (ext && in) {
        if ((rc = upload_tmp(c, ext, std::max<uint32_t>(nc, 1), &de, tmp)) ||
            (rc = upload_tmp(c, in, std::max<uint32_t>(nc, 1), &di, tmp)))
            return rc;
        LAUNCH(c, "stats_pack", cdiv(nc, 256), 256, k_stats_pack, ds, de, di, nc, c->stats);
    } else {
        LAUNCH(c, "stats_sizes", cdiv(nc, 256), 256, k_stats_sizes, ds, nc, c->stats);
    }
    return check_launch("statistics upload");
}

extern "C" int scd_set_partition(scd_ctx *c, const uint32_t *labels, uint32_t num_comms, const int32_t *sizes,
                                 const int32_t *ext_edges, const int32_t *int_edges) {
    int rc = need_graph(c, "scd_set_partition");
    if (rc) return rc;
    const int given = (sizes != nullptr) + (ext_edges != nullptr) + (int_edges != nullptr);
    if (given != 0 && given != 3)
        return fail(SCD_EINVAL, "scd_set_partition: give all three statistics arrays or none");
    if ((rc = put_labels(c, "scd_set_partition", labels, num_comms))) return rc;
    std::vector<DevBuf> tmp;
    if (given) rc = put_stats(c, num_comms, sizes, ext_edges, int_edges, tmp);
    else if (!(rc = clear_counters(c))) rc = phase_stats(c);
    if (!rc) rc = finish(c);
    release_tmp(c, tmp);
    return rc;
}

extern "C" int scd_move_step(scd_ctx *c, uint32_t *labels_out) {
    int rc = need_labels(c, "scd_move_step");
    if (rc) return rc;
    if ((c->counting && (rc = clear_counters(c))) || (rc = phase_move(c, c->next, nullptr, nullptr))) return rc;
    std::swap(c->cur, c->next);
    c->wcc_valid = false;
    if (labels_out)
        HIP_TRY(hipMemcpyAsync(labels_out, c->cur, sizeof(uint32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    return finish(c);
}

extern "C" int scd_compute_wcc_of(scd_ctx *c, const uint32_t *labels, uint32_t num_comms, const int32_t *sizes,
                                  double *wcc) {
    int rc = need_graph(c, "scd_compute_wcc_of");
    if (rc) return rc;
    if ((rc = put_labels(c, "scd_compute_wcc_of", labels, num_comms))) return rc;
    std::vector<DevBuf> tmp;
    // the internal adjacency and statistics of the labels, then the sizes
    // the caller gives in place of the computed ones
    if (!(rc = clear_counters(c)) && !(rc = phase_stats(c)) && sizes)
        rc = put_stats(c, num_comms, sizes, nullptr, nullptr, tmp);
    if (!rc) rc = phase_wcc(c, nullptr, nullptr);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(c->wcc_host, c->wcc_dev, sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e != hipSuccess) rc = fail(SCD_EDEVICE, "scd_compute_wcc_of: %s", hipGetErrorString(e));
    }
    if (!rc) rc = finish(c);
    release_tmp(c, tmp);
    if (rc) return rc;
    if (wcc) *wcc = c->wcc_host[0];
    return SCD_OK;
}

// One iteration's device work on the stream: move (cur -> next), swap,
// statistics, WCC, mean to the pinned host word.  Leaves cur/next swapped.
static int enqueue_iteration(scd_ctx *c, bool incremental) {
    int rc;
    if (c->counting && (rc = clear_counters(c))) return rc;
    stamp(c, 0);
    if ((rc = phase_move(c, c->next, nullptr, nullptr))) return rc;
    stamp(c, 1);
    if (incremental) {
        const uint32_t B = c->block_size;
        LAUNCH(c, "wcc_dirty_clear", cdiv(c->n, 256), 256, k_fill_u8, c->wdirty, (uint64_t)c->n);
        LAUNCH(c, "wcc_dirty_mark", cdiv(c->n, B), B, k_mark_dirty, c->cur, c->next, c->n, c->wdirty);
    }
    std::swap(c->cur, c->next);
    if ((rc = phase_stats(c, true))) {
        wait_stats(c);
        return rc;
    }
    stamp(c, 3);
    if ((rc = phase_wcc(c, nullptr, nullptr, incremental))) return rc;
    stamp(c, 4);
    if (c->phase_timing && c->stamps)
        HIP_TRY(hipMemcpyAsync(c->stamps_host, c->stamps, sizeof(unsigned long long) * kStamps,
                               hipMemcpyDeviceToHost, c->stream));
    if (c->G) HIP_TRY(hipMemcpyAsync(c->wcc_g_host, c->wcc_g, 8ull * c->G, hipMemcpyDeviceToHost, c->stream));
    else HIP_TRY(hipMemcpyAsync(c->wcc_host, c->wcc_dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (c->counting)
        HIP_TRY(hipMemcpyAsync(c->counters_host, c->counters, 16 * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, c->stream));
    return SCD_OK;
}

// Captures enqueue_iteration for the current parity and WCC mode into
// gexec[par][inc].
static int capture_iteration(scd_ctx *c, int par, int inc) {
    uint32_t *cur0 = c->cur, *next0 = c->next;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    c->lane_ev_used[1] = 0;
    int rc = enqueue_iteration(c, inc != 0);
    c->capturing = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->cur = cur0;
    c->next = next0;
    if (rc == SCD_OK && e != hipSuccess) rc = fail(SCD_EDEVICE, "stream capture: %s", hipGetErrorString(e));
    if (rc == SCD_OK) {
        e = hipGraphInstantiate(&c->gexec[par][inc], g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            c->gexec[par][inc] = nullptr;
            rc = fail(SCD_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
        }
    }
    if (g) (void)hipGraphDestroy(g);
    // move the executable graph to the device now, not at its first launch
    if (rc == SCD_OK) (void)hipGraphUpload(c->gexec[par][inc], c->stream);
    if (rc != SCD_OK) (void)hipGetLastError();
    return rc;
}

extern "C" int scd_capture_iterations(scd_ctx *c) {
    int rc = need_labels(c, "scd_capture_iterations");
    if (rc) return rc;
    if (!c->use_graph || c->profile) return SCD_OK;
    // the next iteration (its WCC mode as it stands), then both label
    // parities with the incremental WCC every later iteration uses
    const int par = c->cur == c->lab_a ? 0 : 1;
    const int want[3][2] = {{par, c->wcc_valid ? 1 : 0}, {1 - par, 1}, {par, 1}};
    for (const auto &w : want) {
        const int p = w[0], inc = w[1];
        if (c->gexec[p][inc]) continue;
        // capture with the label buffers of that parity in place
        uint32_t *cur0 = c->cur, *next0 = c->next;
        c->cur = p == 0 ? c->lab_a : c->lab_b;
        c->next = p == 0 ? c->lab_b : c->lab_a;
        rc = capture_iteration(c, p, inc);
        c->cur = cur0;
        c->next = next0;
        if (rc) return rc;
    }
    return finish(c);
}

extern "C" int scd_next_step_captures(const scd_ctx *c) {
    if (!c || !c->use_graph || c->profile || !c->have_labels) return 0;
    const int inc = c->wcc_valid ? 1 : 0;
    const int par = c->cur == c->lab_a ? 0 : 1;
    return c->gexec[par][inc] ? 0 : 1;
}

extern "C" int scd_improve(scd_ctx *c, double *wcc, uint32_t *labels_out) {
    int rc = need_labels(c, "scd_improve");
    if (rc) return rc;
    // per-kernel event timing needs eager launches (events recorded by graph
    // nodes carry no usable timestamps)
    const int inc = c->wcc_valid ? 1 : 0;
    c->wcc_valid = false;
    if (c->use_graph && !c->profile) {
        const int par = c->cur == c->lab_a ? 0 : 1;
        if (!c->gexec[par][inc] && (rc = capture_iteration(c, par, inc))) return rc;
        HIP_TRY(hipGraphLaunch(c->gexec[par][inc], c->stream));
        std::swap(c->cur, c->next);
    } else if ((rc = enqueue_iteration(c, inc != 0))) {
        return rc;
    }
    c->wcc_valid = true;
    if (labels_out)
        HIP_TRY(hipMemcpyAsync(labels_out, c->cur, sizeof(uint32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    if ((rc = finish(c))) return rc;
    if (wcc) *wcc = c->wcc_host[0];
    if (c->phase_timing && c->stamps && c->wall_clock_khz > 0) {
        // reference order: move, statistics, removal of external edges, WCC
        const unsigned long long *t = c->stamps_host;
        const int from[4] = {0, 2, 1, 3}, to[4] = {1, 3, 2, 4};
        for (int i = 0; i < 4; i++) c->phase_ms[i] = (double)(t[to[i]] - t[from[i]]) / c->wall_clock_khz;
    }
    return SCD_OK;
}

extern "C" int scd_set_phase_timing(scd_ctx *c, int on) {
    if (!c) return fail(SCD_EINVAL, "scd_set_phase_timing: null context");
    (void)hipSetDevice(c->device);
    if (on && !c->stamps) {
        hipError_t e = hipMalloc((void **)&c->stamps, sizeof(unsigned long long) * kStamps);
        if (e != hipSuccess) {
            c->stamps = nullptr;
            return fail(SCD_ENOMEM, "scd_set_phase_timing: hipMalloc: %s", hipGetErrorString(e));
        }
    }
    if (on && c->wall_clock_khz <= 0) return fail(SCD_EDEVICE, "scd_set_phase_timing: no device wall clock rate");
    if (c->phase_timing != (on != 0)) drop_graphs(c);  // captured iterations hold the stamp kernels or not
    c->phase_timing = on != 0;
    for (double &x : c->phase_ms) x = 0.0;
    return SCD_OK;
}

extern "C" int scd_last_phase_ms(const scd_ctx *c, double ms[4]) {
    if (!c || !ms) return fail(SCD_EINVAL, "scd_last_phase_ms: null argument");
    for (int i = 0; i < 4; i++) ms[i] = c->phase_ms[i];
    return SCD_OK;
}

extern "C" int scd_set_profiling(scd_ctx *c, int on) {
    if (!c) return fail(SCD_EINVAL, "scd_set_profiling: null context");
    c->profile = on != 0;
    return SCD_OK;
}

extern "C" int scd_set_graph_mode(scd_ctx *c, int on) {
    if (!c) return fail(SCD_EINVAL, "scd_set_graph_mode: null context");
    if (!on) drop_graphs(c);
    c->use_graph = on != 0;
    return SCD_OK;
}

extern "C" int scd_store_best(scd_ctx *c) {
    int rc = need_labels(c, "scd_store_best");
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(c->best, c->cur, sizeof(uint32_t) * c->n, hipMemcpyDeviceToDevice, c->stream));
    return finish(c);
}

extern "C" int scd_labels_from_device(scd_ctx *c, uint32_t *labels) {
    int rc = need_labels(c, "scd_labels_from_device");
    if (rc) return rc;
    if (!labels) return fail(SCD_EINVAL, "scd_labels_from_device: null labels");
    HIP_TRY(hipMemcpyAsync(labels, c->best, sizeof(uint32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    return finish(c);
}

extern "C" int scd_current_labels(scd_ctx *c, uint32_t *labels) {
    int rc = need_labels(c, "scd_current_labels");
    if (rc) return rc;
    if (!labels) return fail(SCD_EINVAL, "scd_current_labels: null labels");
    HIP_TRY(hipMemcpyAsync(labels, c->cur, sizeof(uint32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    return finish(c);
}

extern "C" int scd_refine(scd_ctx *c, double initial_wcc, uint32_t max_iters, double *best_wcc,
                          uint32_t *iterations, double *wcc_hist) {
    int rc = need_labels(c, "scd_refine");
    if (rc) return rc;
    const double threshold = 0.01;
    double best = initial_wcc;
    uint32_t it = 0;
    while (it < max_iters) {
        it++;
        double wcc = 0.0;
        if ((rc = scd_improve(c, &wcc, nullptr))) return rc;
        if (wcc_hist) wcc_hist[it - 1] = wcc;
        if (wcc > best && (rc = scd_store_best(c))) return rc;
        if (wcc < best * (1 + threshold)) break;
        best = wcc;
    }
    if (best_wcc) *best_wcc = best;
    if (iterations) *iterations = it;
    return SCD_OK;
}

extern "C" int scd_compute_wcc(scd_ctx *c, double *wcc, double *terms_out, int32_t *int_tri_out,
                               int32_t *int_tri_deg_out) {
    int rc = need_labels(c, "scd_compute_wcc");
    if (rc) return rc;
    int32_t *d_it = nullptr, *d_itd = nullptr;
    std::vector<DevBuf> tmp;
    if (int_tri_out || int_tri_deg_out) {
        if ((rc = dalloc(c, &d_it, c->n, &tmp)) || (rc = dalloc(c, &d_itd, c->n, &tmp))) {
            release_tmp(c, tmp);
            return rc;
        }
    }
    if ((rc = clear_counters(c)) || (rc = phase_stats(c)) || (rc = phase_wcc(c, d_it, d_itd))) {
        release_tmp(c, tmp);
        return rc;
    }
    HIP_TRY(hipMemcpyAsync(c->wcc_host, c->wcc_dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (terms_out)
        HIP_TRY(hipMemcpyAsync(terms_out, c->terms, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
    if (int_tri_out)
        HIP_TRY(hipMemcpyAsync(int_tri_out, d_it, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    if (int_tri_deg_out)
        HIP_TRY(hipMemcpyAsync(int_tri_deg_out, d_itd, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    rc = finish(c);
    release_tmp(c, tmp);
    if (rc) return rc;
    c->wcc_valid = true;
    if (wcc) *wcc = c->wcc_host[0];
    return SCD_OK;
}

extern "C" int scd_get_stats(scd_ctx *c, uint32_t *num_comms, int32_t *sizes, int32_t *ext_edges,
                             int32_t *int_edges) {
    int rc = need_labels(c, "scd_get_stats");
    if (rc) return rc;
    std::vector<int4> h(c->num_comms);
    HIP_TRY(hipMemcpyAsync(h.data(), c->stats, sizeof(int4) * c->num_comms, hipMemcpyDeviceToHost, c->stream));
    if ((rc = finish(c))) return rc;
    for (uint32_t i = 0; i < c->num_comms; i++) {
        if (sizes) sizes[i] = h[i].x;
        if (ext_edges) ext_edges[i] = h[i].y;
        if (int_edges) int_edges[i] = h[i].z;
    }
    if (num_comms) *num_comms = c->num_comms;
    return SCD_OK;
}

// Exclusive scan of n uint32 values (in -> out, may alias) on the stream;
// returns the total through *total (host).
static int scan_u32(scd_ctx *c, const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *total,
                    std::vector<DevBuf> &tmp) {
    const uint32_t nt = cdiv(n, 2048);
    uint32_t *sums = nullptr, *sums_scan = nullptr;
    int rc;
    if ((rc = dalloc(c, &sums, nt, &tmp)) || (rc = dalloc(c, &sums_scan, nt, &tmp))) return rc;
    LAUNCH(c, "scan_tile_sums", nt, 256, k_scan_tile_sums, in, n, sums);
    if (nt > 1) {
        uint32_t dummy;
        if ((rc = scan_u32(c, sums, sums_scan, nt, &dummy, tmp))) return rc;
    } else {
        HIP_TRY(hipMemsetAsync(sums_scan, 0, sizeof(uint32_t), c->stream));
    }
    if ((rc = check_launch("scan"))) return rc;
    // total = last exclusive offset + last tile sum (read before out overwrites in)
    uint32_t last[2];
    HIP_TRY(hipMemcpyAsync(&last[0], sums_scan + nt - 1, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&last[1], sums + nt - 1, 4, hipMemcpyDeviceToHost, c->stream));
    LAUNCH(c, "scan_tiles", nt, 256, k_scan_tiles, in, n, sums_scan, out);
    HIP_TRY(hipStreamSynchronize(c->stream));
    *total = last[0] + last[1];
    return check_launch("scan");
}

extern "C" int scd_neighbor_label_pairs(scd_ctx *c, uint32_t *sorted_out, uint32_t *run_offsets,
                                        uint32_t *run_label, uint32_t *run_count, uint64_t *num_pairs) {
    int rc = need_labels(c, "scd_neighbor_label_pairs");
    if (rc) return rc;
    std::vector<DevBuf> tmp;
    uint32_t *d_sorted = nullptr, *d_cnt = nullptr, *d_pos = nullptr, *d_rl = nullptr, *d_rc = nullptr,
             *d_off = nullptr;
    const uint64_t E = c->E;
    if ((rc = dalloc(c, &d_sorted, E, &tmp)) || (rc = dalloc(c, &d_cnt, E, &tmp)) ||
        (rc = dalloc(c, &d_pos, E, &tmp)) || (rc = dalloc(c, &d_off, (size_t)c->n + 1, &tmp))) {
        release_tmp(c, tmp);
        return rc;
    }
    uint32_t total = 0;
    rc = clear_counters(c);
    if (!rc) rc = phase_move(c, nullptr, d_sorted, d_cnt);
    if (!rc && E) {
        LAUNCH(c, "run_flags", cdiv(E, 256), 256, k_run_flags, d_cnt, E, d_pos);
        rc = scan_u32(c, d_pos, d_pos, E, &total, tmp);
    }
    if (!rc) {
        if ((rc = dalloc(c, &d_rl, total, &tmp)) || (rc = dalloc(c, &d_rc, total, &tmp))) {
            release_tmp(c, tmp);
            return rc;
        }
        if (E) LAUNCH(c, "run_scatter", cdiv(E, 256), 256, k_run_scatter, d_sorted, d_cnt, d_pos, E, d_rl, d_rc);
        LAUNCH(c, "run_offsets", cdiv((uint64_t)c->n + 1, 256), 256, k_run_offsets, c->indices, d_pos, c->n, E,
               total, d_off);
        rc = check_launch("pairs");
    }
    if (!rc) {
        hipStream_t s = c->stream;
        hipError_t e = hipSuccess;
        if (sorted_out && E) e = hipMemcpyAsync(sorted_out, d_sorted, 4 * E, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess && run_offsets)
            e = hipMemcpyAsync(run_offsets, d_off, 4 * ((size_t)c->n + 1), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess && run_label && total) e = hipMemcpyAsync(run_label, d_rl, 4ull * total, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess && run_count && total) e = hipMemcpyAsync(run_count, d_rc, 4ull * total, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) rc = fail(SCD_EDEVICE, "scd_neighbor_label_pairs: copy: %s", hipGetErrorString(e));
    }
    if (!rc) rc = finish(c);
    release_tmp(c, tmp);
    if (rc) return rc;
    if (num_pairs) *num_pairs = total;
    return SCD_OK;
}

extern "C" int scd_get_tuning(const scd_ctx *c, scd_tuning *out) {
    if (!c || !out) return fail(SCD_EINVAL, "scd_get_tuning: null argument");
    *out = c->tune;
    return SCD_OK;
}

extern "C" int scd_set_tuning(scd_ctx *c, const scd_tuning *t) {
    if (!c || !t) return fail(SCD_EINVAL, "scd_set_tuning: null argument");
    const uint32_t T = t->lds_tile, C = t->merge_chunk;
    if (T != 2048 && T != 4096 && T != 8192 && T != 16384)
        return fail(SCD_EINVAL, "scd_set_tuning: lds_tile %u is not 2048, 4096, 8192 or 16384", T);
    if (C < 256 || C > 2 * T || (C & (C - 1)))
        return fail(SCD_EINVAL, "scd_set_tuning: merge_chunk %u must be a power of two in [256, 2*lds_tile]", C);
    const uint32_t W = t->wave_max ? t->wave_max : 1024u;
    if (W != 256 && W != 512 && W != 1024)
        return fail(SCD_EINVAL, "scd_set_tuning: wave_max %u is not 256, 512 or 1024", W);
    if (t->hub_table > 2) return fail(SCD_EINVAL, "scd_set_tuning: hub_table %u is not 0, 1 or 2", t->hub_table);
    if (t->lds_hash > 1 || (t->lds_hash && T > 4096))
        return fail(SCD_EINVAL, "scd_set_tuning: lds_hash %u needs lds_tile <= 4096", t->lds_hash);
    if (t->lanes > 1) return fail(SCD_EINVAL, "scd_set_tuning: lanes %u is not 0 or 1", t->lanes);
    if (t->dense_wcc > 2) return fail(SCD_EINVAL, "scd_set_tuning: dense_wcc %u is not 0, 1 or 2", t->dense_wcc);
    if (c->tune.hub_table != t->hub_table || c->tune.lds_hash != t->lds_hash || c->tune.lanes != t->lanes ||
        c->tune.dense_wcc != t->dense_wcc)
        drop_graphs(c);
    c->tune = *t;
    c->tune.wave_max = W;
    if (c->have_graph) {
        (void)hipSetDevice(c->device);
        HIP_TRY(hipStreamSynchronize(c->stream));
        int rc = build_bins(c);
        if (rc) return rc;
    }
    return SCD_OK;
}

extern "C" int scd_set_triangle_split(scd_ctx *c, uint32_t tc_heavy, uint32_t wcc_heavy) {
    if (!c) return fail(SCD_EINVAL, "scd_set_triangle_split: null context");
    for (uint32_t h : {tc_heavy, wcc_heavy})
        if (h && (h < WTC_GROUP_MAX || h >= WTC_VC_CAP))
            return fail(SCD_EINVAL, "scd_set_triangle_split: %u is not 0 or in [64, 2048)", h);
    if (c->wcc_heavy != wcc_heavy) drop_graphs(c);  // captured iterations hold the WCC's value
    c->tc_heavy = tc_heavy;
    c->wcc_heavy = wcc_heavy;
    return SCD_OK;
}

extern "C" int scd_last_counters(scd_ctx *c, uint64_t out[4]) {
    if (!c || !out) return fail(SCD_EINVAL, "scd_last_counters: null argument");
    out[0] = 0;
    for (int b = 0; b < K_NB; b++) out[0] += c->counters_host[b];
    out[1] = c->counters_host[K_NB];
    out[2] = c->counters_host[K_NB + 1];
    out[3] = c->nh;
    return SCD_OK;
}

extern "C" int scd_dense_wcc_info(scd_ctx *c, uint32_t out[2]) {
    if (!c || !out) return fail(SCD_EINVAL, "scd_dense_wcc_info: null argument");
    out[0] = out[1] = 0;
    if (!c->dn_ctr) return SCD_OK;
    (void)hipSetDevice(c->device);
    HIP_TRY(hipMemcpyAsync(c->wcc_host + 4, c->dn_ctr, 8, hipMemcpyDeviceToHost, c->stream));
    int rc = finish(c);
    if (rc) return rc;
    memcpy(out, c->wcc_host + 4, 8);
    return SCD_OK;
}

extern "C" int scd_bin_info(scd_ctx *c, int max_bins, char (*names)[16], uint64_t *vertices,
                            uint64_t *entries, uint64_t *pairs) {
    static const char *bin_names[K_NB] = {"sub8", "sub16", "sub32", "wave1", "wave2",
                                          "wave4", "wave8", "wave16", "lds", "lds2", "hub"};
    if (!c) return fail(SCD_EINVAL, "scd_bin_info: null context");
    if (!c->have_graph) return fail(SCD_ESTATE, "scd_bin_info: no graph uploaded");
    const int k = std::min<int>(max_bins, K_NB);
    for (int b = 0; b < k; b++) {
        if (names) snprintf(names[b], 16, "%s", bin_names[b]);
        if (vertices) vertices[b] = bin_n(c, b);
        if (entries) entries[b] = c->bin_edges[b];
        if (pairs) pairs[b] = c->counters_host[b];
    }
    return K_NB;
}

extern "C" int scd_set_counting(scd_ctx *c, int on) {
    if (!c) return fail(SCD_EINVAL, "scd_set_counting: null context");
    if (c->counting != (on != 0)) drop_graphs(c);
    c->counting = on != 0;
    return SCD_OK;
}

extern "C" int scd_snapshot(scd_ctx *c) {
    int rc = need_labels(c, "scd_snapshot");
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(c->snap_labels, c->cur, 4ull * c->n, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->snap_stats, c->stats, sizeof(int4) * c->num_comms, hipMemcpyDeviceToDevice,
                           c->stream));
    c->snap_num_comms = c->num_comms;
    c->have_snap = true;
    return finish(c);
}

extern "C" int scd_restore(scd_ctx *c) {
    int rc = need_labels(c, "scd_restore");
    if (rc) return rc;
    if (!c->have_snap) return fail(SCD_ESTATE, "scd_restore: no snapshot taken");
    HIP_TRY(hipMemcpyAsync(c->cur, c->snap_labels, 4ull * c->n, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->best, c->snap_labels, 4ull * c->n, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->stats, c->snap_stats, sizeof(int4) * c->snap_num_comms, hipMemcpyDeviceToDevice,
                           c->stream));
    if (c->num_comms != c->snap_num_comms) drop_graphs(c);
    c->num_comms = c->snap_num_comms;
    c->wcc_valid = false;
    return finish(c);
}

// Measurement helper: a plain streaming copy kernel over `bytes` bytes of
// device memory, for calibrating hardware traffic counters.
__global__ void k_calib_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" int scd_calibration_copy(scd_ctx *c, uint64_t bytes, int reps, double *ms_per_copy) {
    if (!c) return fail(SCD_EINVAL, "scd_calibration_copy: null context");
    (void)hipSetDevice(c->device);
    std::vector<DevBuf> tmp;
    uint4 *a = nullptr, *b = nullptr;
    const uint64_t n = bytes / 16;
    int rc;
    if ((rc = dalloc(c, &a, n, &tmp)) || (rc = dalloc(c, &b, n, &tmp))) {
        release_tmp(c, tmp);
        return rc;
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const uint32_t grid = (uint32_t)std::min<uint64_t>(cdiv(n, 256), 256ull * 64);
    HIP_TRY(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_calib_copy, dim3(grid), dim3(256), 0, c->stream, a, b, n);
    HIP_TRY(hipEventRecord(e1, c->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    release_tmp(c, tmp);
    if (ms_per_copy) *ms_per_copy = reps > 0 ? ms / reps : 0.0;
    return check_launch("calibration copy");
}

extern "C" int scd_prof_count(scd_ctx *c) { return c ? (int)c->prof.size() : 0; }

extern "C" int scd_prof_get(scd_ctx *c, int i, const char **name, uint64_t *launches, double *total_ms) {
    if (!c || i < 0 || i >= (int)c->prof.size()) return fail(SCD_EINVAL, "scd_prof_get: bad index");
    if (name) *name = c->prof[i].name.c_str();
    if (launches) *launches = c->prof[i].launches;
    if (total_ms) *total_ms = c->prof[i].ms;
    return SCD_OK;
}

extern "C" int scd_prof_reset(scd_ctx *c) {
    if (!c) return fail(SCD_EINVAL, "scd_prof_reset: null context");
    for (auto &p : c->prof) {
        p.launches = 0;
        p.ms = 0.0;
    }
    return SCD_OK;
}

extern "C" uint64_t scd_device_bytes(const scd_ctx *c) { return c ? c->dev_bytes : 0; }

// ---------------------------------------------------------------------------
// The move step's arithmetic on caller rows (dev.hpp's improvement and
// candidate set-up, the functions every move kernel inlines).
// ---------------------------------------------------------------------------
__global__ void k_eval_improvement(uint32_t count, const int32_t *__restrict__ r, const int32_t *__restrict__ d_in,
                                   const int32_t *__restrict__ d_out, const int32_t *__restrict__ c_out,
                                   const double *__restrict__ p_in, const double *__restrict__ p_ext,
                                   double *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = improvement(r[i], d_in[i], d_out[i], c_out[i], p_in[i], p_ext[i], 1.0);
}

__global__ void k_eval_candidate(uint32_t count, const int4 *__restrict__ s, const uint32_t *__restrict__ L,
                                 const uint32_t *__restrict__ freq, const uint32_t *__restrict__ deg,
                                 const uint32_t *__restrict__ my, double p_ext, double *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = candidate_improvement_s(s[i], L[i], freq[i], deg[i], my[i], p_ext);
}

// Copies `count` elements of each host array to a device temporary.
template <typename T>
static int upload_tmp(scd_ctx *c, const T *h, uint32_t count, T **d, std::vector<DevBuf> &tmp) {
    int rc = dalloc(c, d, count, &tmp);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(*d, h, sizeof(T) * count, hipMemcpyHostToDevice, c->stream));
    return SCD_OK;
}

extern "C" int scd_eval_improvement(scd_ctx *c, uint32_t count, const int32_t *r, const int32_t *d_in,
                                    const int32_t *d_out, const int32_t *c_out, const double *p_in,
                                    const double *p_ext, double *out) {
    if (!c) return fail(SCD_EINVAL, "scd_eval_improvement: null context");
    if (count == 0) return SCD_OK;
    if (!r || !d_in || !d_out || !c_out || !p_in || !p_ext || !out)
        return fail(SCD_EINVAL, "scd_eval_improvement: null array");
    (void)hipSetDevice(c->device);
    std::vector<DevBuf> tmp;
    int32_t *dr = nullptr, *di = nullptr, *dd = nullptr, *dc = nullptr;
    double *dpi = nullptr, *dpe = nullptr, *dout = nullptr;
    int rc;
    if ((rc = upload_tmp(c, r, count, &dr, tmp)) || (rc = upload_tmp(c, d_in, count, &di, tmp)) ||
        (rc = upload_tmp(c, d_out, count, &dd, tmp)) || (rc = upload_tmp(c, c_out, count, &dc, tmp)) ||
        (rc = upload_tmp(c, p_in, count, &dpi, tmp)) || (rc = upload_tmp(c, p_ext, count, &dpe, tmp)) ||
        (rc = dalloc(c, &dout, count, &tmp))) {
        release_tmp(c, tmp);
        return rc;
    }
    hipLaunchKernelGGL(k_eval_improvement, dim3(cdiv(count, 256)), dim3(256), 0, c->stream, count, dr, di, dd, dc,
                       dpi, dpe, dout);
    if ((rc = check_launch("scd_eval_improvement")) == SCD_OK) {
        hipError_t e = hipMemcpyAsync(out, dout, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(SCD_EDEVICE, "scd_eval_improvement: %s", hipGetErrorString(e));
    }
    release_tmp(c, tmp);
    return rc;
}

extern "C" int scd_eval_candidate(scd_ctx *c, uint32_t count, const int32_t *size, const int32_t *ext_edges,
                                  const int32_t *int_edges, const uint32_t *label, const uint32_t *freq,
                                  const uint32_t *degree, const uint32_t *my_label, double clustering_coef,
                                  double *out) {
    if (!c) return fail(SCD_EINVAL, "scd_eval_candidate: null context");
    if (count == 0) return SCD_OK;
    if (!size || !ext_edges || !int_edges || !label || !freq || !degree || !my_label || !out)
        return fail(SCD_EINVAL, "scd_eval_candidate: null array");
    (void)hipSetDevice(c->device);
    std::vector<int4> s(count);
    for (uint32_t i = 0; i < count; i++) s[i] = make_int4(size[i], ext_edges[i], int_edges[i], 0);
    std::vector<DevBuf> tmp;
    int4 *ds = nullptr;
    uint32_t *dl = nullptr, *df = nullptr, *dg = nullptr, *dm = nullptr;
    double *dout = nullptr;
    int rc;
    if ((rc = upload_tmp(c, s.data(), count, &ds, tmp)) || (rc = upload_tmp(c, label, count, &dl, tmp)) ||
        (rc = upload_tmp(c, freq, count, &df, tmp)) || (rc = upload_tmp(c, degree, count, &dg, tmp)) ||
        (rc = upload_tmp(c, my_label, count, &dm, tmp)) || (rc = dalloc(c, &dout, count, &tmp))) {
        release_tmp(c, tmp);
        return rc;
    }
    hipLaunchKernelGGL(k_eval_candidate, dim3(cdiv(count, 256)), dim3(256), 0, c->stream, count, ds, dl, df, dg, dm,
                       clustering_coef, dout);
    if ((rc = check_launch("scd_eval_candidate")) == SCD_OK) {
        hipError_t e = hipMemcpyAsync(out, dout, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(SCD_EDEVICE, "scd_eval_candidate: %s", hipGetErrorString(e));
    }
    release_tmp(c, tmp);
    return rc;
}

// ---------------------------------------------------------------------------
// Triangle statistics (src/convert.cpp:271-340) on the device.
// ---------------------------------------------------------------------------
// t / vt (remove_edges_without_triangles, src/convert.cpp:271-351): the
// internal-triangle kernels over the whole adjacency, tri = 2 x triangles.
// Triangles at every vertex of the context's graph (T), triangle edges at
// every vertex (td) and the flag of the oriented slot
// u -> w (w < u) of every edge in a triangle.
static int triangles_core(scd_ctx *c, uint32_t **T_out, uint32_t *d_td, uint8_t **flag_out, std::vector<DevBuf> &tmp) {
    const uint32_t n = c->n;
    const uint64_t E = c->E;
    uint32_t *dplus = nullptr, *T = nullptr;
    uint8_t *flag = nullptr;
    int rc;
    if ((rc = dalloc(c, &dplus, n, &tmp)) || (rc = dalloc(c, &T, n, &tmp)) || (rc = dalloc(c, &flag, E, &tmp)))
        return rc;
    WccArgs w;
    memset(&w, 0, sizeof(w));
    w.indices = c->indices;
    w.int_adj = c->adj;
    w.int_deg = nullptr;
    w.dplus = dplus;
    w.T = T;
    w.td = d_td;
    w.flag = flag;
    w.T_hot = c->hot_T;
    w.td_hot = c->hot_td;
    w.hot_n = c->hot_n;
    w.vc_min = c->tc_heavy;
    LAUNCH(c, "tc_clear", cdiv(E, 256), 256, k_fill_u8, flag, E);
    if ((rc = count_internal_triangles(c, w))) return rc;
    *T_out = T;
    if (flag_out) *flag_out = flag;
    return check_launch("triangles");
}

static int triangles_dev(scd_ctx *c, int32_t *d_tri, uint32_t *d_td, std::vector<DevBuf> &tmp) {
    uint32_t *T = nullptr;
    int rc = triangles_core(c, &T, d_td, nullptr, tmp);
    if (rc) return rc;
    LAUNCH(c, "tc_double", cdiv(c->n, 256), 256, k_wtc_double, T, c->n, d_tri);
    return check_launch("triangles");
}

extern "C" int scd_count_triangles(scd_ctx *c, int32_t *tri_out, int32_t *tri_deg_out, int replace) {
    int rc = need_graph(c, "scd_count_triangles");
    if (rc) return rc;
    std::vector<DevBuf> tmp;
    int32_t *d_tri = nullptr;
    uint32_t *d_td = nullptr;
    if ((rc = dalloc(c, &d_tri, c->n, &tmp)) || (rc = dalloc(c, &d_td, c->n, &tmp)) ||
        (rc = triangles_dev(c, d_tri, d_td, tmp))) {
        release_tmp(c, tmp);
        return rc;
    }
    hipError_t e = hipSuccess;
    if (tri_out) e = hipMemcpyAsync(tri_out, d_tri, 4ull * c->n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && tri_deg_out)
        e = hipMemcpyAsync(tri_deg_out, d_td, 4ull * c->n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && replace) {
        e = hipMemcpyAsync(c->tri, d_tri, 4ull * c->n, hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(c->tri_deg, d_td, 4ull * c->n, hipMemcpyDeviceToDevice, c->stream);
    }
    if (e != hipSuccess) rc = fail(SCD_EDEVICE, "scd_count_triangles: copy: %s", hipGetErrorString(e));
    if (!rc) rc = finish(c);
    release_tmp(c, tmp);
    return rc;
}

// scd_count_triangles with device outputs and no allocation: the
// out-degrees, triangle counts and edge flags of the pass live in the static
// scratch array (every call rewrites what it reads: k_wtc_dplus zeroes the
// counts, tc_clear the flags); the per-bin kernels run on the context's
// stream only (lane fork / join events are not created here).
extern "C" int scd_count_triangles_noalloc(scd_ctx *c, int32_t *tri_out, int32_t *tri_deg_out) {
    int rc = need_graph(c, "scd_count_triangles_noalloc");
    if (rc) return rc;
    if (c->profile)
        return fail(SCD_ESTATE, "scd_count_triangles_noalloc: profiling creates an event per launch; turn it off");
    uint8_t *base = scratch_of(c->device);
    if (!base) {
        const hipError_t e = c->device >= 0 && c->device < kScratchDevices ? g_scratch_err[c->device] : hipSuccess;
        return fail(SCD_EDEVICE, "scd_count_triangles_noalloc: no scratch address for device %d (%s)", c->device,
                    hipGetErrorString(e));
    }
    const uint32_t n = c->n;
    const uint64_t E = c->E;
    auto al = [](uint64_t b) { return (b + 255) & ~255ull; };
    const uint64_t need = 2 * al(4ull * n) + (tri_deg_out ? 0 : al(4ull * n)) + al(E);
    if (need > kScratchBytes)
        return fail(SCD_ELIMIT,
                    "scd_count_triangles_noalloc: n = %u, 2m = %llu need %llu bytes of scratch; the static array "
                    "holds %llu (use scd_count_triangles)",
                    n, (unsigned long long)E, (unsigned long long)need, (unsigned long long)kScratchBytes);
    uint64_t o = 0;
    uint32_t *dplus = (uint32_t *)(base + o);
    o += al(4ull * n);
    uint32_t *T = (uint32_t *)(base + o);
    o += al(4ull * n);
    uint32_t *td = tri_deg_out ? (uint32_t *)tri_deg_out : (uint32_t *)(base + o);
    if (!tri_deg_out) o += al(4ull * n);
    uint8_t *flag = base + o;
    WccArgs w;
    memset(&w, 0, sizeof(w));
    w.indices = c->indices;
    w.int_adj = c->adj;
    w.int_deg = nullptr;
    w.dplus = dplus;
    w.T = T;
    w.td = td;
    w.flag = flag;
    w.T_hot = c->hot_T;
    w.td_hot = c->hot_td;
    w.hot_n = c->hot_n;
    w.vc_min = c->tc_heavy;
    c->one_stream = true;
    LAUNCH(c, "tc_clear", cdiv(E, 256), 256, k_fill_u8, flag, E);
    rc = count_internal_triangles(c, w);
    if (!rc && tri_out) LAUNCH(c, "tc_double", cdiv(n, 256), 256, k_wtc_double, T, n, tri_out);
    c->one_stream = false;
    if (!rc) rc = check_launch("scd_count_triangles_noalloc");
    if (!rc) rc = finish(c);
    return rc;
}

// ---------------------------------------------------------------------------
// Convert (src/convert.cpp:466-537) on the device: an edge list becomes the
// binary graph layout, byte-identical to the host scd_convert, and stays
// uploaded in the context.
// ---------------------------------------------------------------------------
// Stable LSD radix sort of n 64-bit keys with 8-bit digits; digits equal in
// every key are skipped.  *keys and *alt swap with each pass: the sorted
// keys end in *keys.
static int sort_u64(scd_ctx *c, uint64_t **keys, uint64_t **alt, uint64_t n, std::vector<DevBuf> &tmp) {
    if (n < 2) return SCD_OK;
    unsigned long long *bits = nullptr;
    uint32_t *hist = nullptr, *hoff = nullptr;
    const uint32_t ntiles = cdiv(n, CV_TILE);
    int rc;
    if ((rc = dalloc(c, &bits, 2, &tmp)) || (rc = dalloc(c, &hist, 256ull * ntiles, &tmp)) ||
        (rc = dalloc(c, &hoff, 256ull * ntiles, &tmp)))
        return rc;
    LAUNCH(c, "cv_sort_bits", 1, 64, k_cv_bits_init, bits);
    LAUNCH(c, "cv_sort_bits", std::min(cdiv(n, 256), CV_REDUCE_BLOCKS), 256, k_cv_bits, *keys, n, bits);
    if ((rc = check_launch("convert sort"))) return rc;
    unsigned long long hb[2];
    HIP_TRY(hipMemcpyAsync(hb, bits, sizeof(hb), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const uint64_t diff = hb[0] ^ hb[1];
    for (int pass = 0; pass < 8; pass++) {
        if (((diff >> (8 * pass)) & 0xFFu) == 0) continue;
        LAUNCH(c, "cv_sort_hist", ntiles, 256, k_cv_hist, *keys, n, 8 * pass, ntiles, hist);
        if ((rc = check_launch("convert sort"))) return rc;
        uint32_t total = 0;
        if ((rc = scan_u32(c, hist, hoff, 256ull * ntiles, &total, tmp))) return rc;
        LAUNCH(c, "cv_sort_scatter", ntiles, 256, k_cv_scatter, *keys, n, 8 * pass, ntiles, hoff, *alt);
        std::swap(*keys, *alt);
    }
    return check_launch("convert sort");
}

// CSR (indices of n + 1 entries, adj) of E sorted directed keys over n
// vertices; h_idx (may be null) receives the host copy of indices.
static int csr_of_keys(scd_ctx *c, const uint64_t *dk, uint64_t E, uint32_t n, uint32_t *indices, uint32_t *adj,
                       std::vector<uint32_t> *h_idx, std::vector<DevBuf> &tmp) {
    uint32_t *deg = nullptr;
    int rc;
    if ((rc = dalloc(c, &deg, n, &tmp))) return rc;
    LAUNCH(c, "cv_csr", cdiv(n, 256), 256, k_fill_u32, deg, n);
    LAUNCH(c, "cv_csr", cdiv(E, 256), 256, k_cv_csr, dk, E, adj, deg);
    if ((rc = check_launch("convert csr"))) return rc;
    uint32_t total = 0;
    if ((rc = scan_u32(c, deg, indices, n, &total, tmp))) return rc;
    if (total != E) return fail(SCD_EDEVICE, "convert: CSR holds %u of %llu entries", total, (unsigned long long)E);
    const uint32_t last = (uint32_t)E;
    HIP_TRY(hipMemcpyAsync(indices + n, &last, 4, hipMemcpyHostToDevice, c->stream));
    if (h_idx) {
        h_idx->resize((size_t)n + 1);
        (*h_idx)[n] = last;
        HIP_TRY(hipMemcpyAsync(h_idx->data(), indices, 4ull * n, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SCD_OK;
}

static double wall_ms() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static int convert_impl(scd_ctx *c, uint64_t count, const uint32_t *u, const uint32_t *v, int on_device,
                        uint32_t *n_out, uint32_t *m_out, double *cc_out, std::vector<DevBuf> &tmp) {
    int rc;
    *n_out = 0;
    *m_out = 0;
    *cc_out = std::numeric_limits<double>::quiet_NaN();  // no vertices: 0.0 / 0
    scd_convert_info &info = c->cv_info;
    info = scd_convert_info{};
    info.edges_read = count;
    if (count == 0) return SCD_OK;
    // stage times: wall clock between stream synchronizations
    double t_last = wall_ms();
    auto stage = [&](double &slot) -> int {
        HIP_TRY(hipStreamSynchronize(c->stream));
        const double t = wall_ms();
        slot = t - t_last;
        t_last = t;
        return SCD_OK;
    };
    if (!on_device) {
        uint32_t *du = nullptr, *dv = nullptr;
        if ((rc = dalloc(c, &du, count, &tmp)) || (rc = dalloc(c, &dv, count, &tmp))) return rc;
        HIP_TRY(hipMemcpyAsync(du, u, 4ull * count, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(dv, v, 4ull * count, hipMemcpyHostToDevice, c->stream));
        u = du;
        v = dv;
    }
    // 1. ids by first appearance (src/convert.cpp:141-170)
    uint32_t *mx = nullptr, hmx = 0;
    if ((rc = dalloc(c, &mx, 1, &tmp))) return rc;
    LAUNCH(c, "cv_ids", 1, 64, k_fill_u32, mx, 1u);
    LAUNCH(c, "cv_ids", std::min(cdiv(count, 256), CV_REDUCE_BLOCKS), 256, k_cv_max, u, v, count, mx);
    if ((rc = check_launch("convert"))) return rc;
    HIP_TRY(hipMemcpyAsync(&hmx, mx, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const uint64_t nid = (uint64_t)hmx + 1;
    info.id_bound = (uint32_t)nid;  // 32-bit, as the reference's num_nodes
    unsigned long long *first = nullptr;
    uint32_t *seq = nullptr;
    if ((rc = dalloc(c, &first, nid, &tmp)) || (rc = dalloc(c, &seq, 2 * count, &tmp))) return rc;
    LAUNCH(c, "cv_ids", cdiv(nid, 256), 256, k_fill_u64v, first, nid, ~0ull);
    LAUNCH(c, "cv_ids", cdiv(count, 256), 256, k_cv_first, u, v, count, first);
    LAUNCH(c, "cv_ids", cdiv(2 * count, 256), 256, k_fill_u32, seq, (uint32_t)(2 * count));
    LAUNCH(c, "cv_ids", cdiv(nid, 256), 256, k_cv_mark, first, nid, seq);
    if ((rc = check_launch("convert"))) return rc;
    uint32_t nv = 0;
    if ((rc = scan_u32(c, seq, seq, 2 * count, &nv, tmp))) return rc;
    info.vertices = nv;
    // 2. undirected keys sorted, loops and duplicates dropped
    //    (src/convert.cpp:252-268, :501)
    uint64_t *ka = nullptr, *kb = nullptr;
    uint32_t *keep = nullptr, *pos = nullptr;
    if ((rc = dalloc(c, &ka, count, &tmp)) || (rc = dalloc(c, &kb, count, &tmp)) ||
        (rc = dalloc(c, &keep, count, &tmp)) || (rc = dalloc(c, &pos, count, &tmp)))
        return rc;
    LAUNCH(c, "cv_keys", cdiv(count, 256), 256, k_cv_keys, u, v, count, first, seq, ka);
    if ((rc = stage(info.ms_reindex)) || (rc = sort_u64(c, &ka, &kb, count, tmp)) || (rc = stage(info.ms_sort)))
        return rc;
    LAUNCH(c, "cv_unique", cdiv(count, 256), 256, k_cv_unique, ka, count, keep);
    if ((rc = check_launch("convert"))) return rc;
    uint32_t mu = 0;
    if ((rc = scan_u32(c, keep, pos, count, &mu, tmp)) || (rc = stage(info.ms_dedup))) return rc;
    info.edges_simple = mu;
    if (mu == 0) return SCD_OK;  // loops only: nothing survives
    // 3. the simple graph in degree order (hubs get the low ids the triangle
    //    kernels orient toward); both directions sorted: the CSR
    uint32_t *deg = nullptr, *rank = nullptr, *perm = nullptr;
    uint64_t *oa = nullptr, *ob = nullptr;
    if ((rc = dalloc(c, &deg, nv, &tmp)) || (rc = dalloc(c, &rank, nv, &tmp)) || (rc = dalloc(c, &perm, nv, &tmp)) ||
        (rc = dalloc(c, &oa, nv, &tmp)) || (rc = dalloc(c, &ob, nv, &tmp)))
        return rc;
    LAUNCH(c, "cv_degree", cdiv(nv, 256), 256, k_fill_u32, deg, nv);
    LAUNCH(c, "cv_degree", cdiv(count, 256), 256, k_cv_deg, ka, count, keep, deg);
    LAUNCH(c, "cv_degree", cdiv(nv, 256), 256, k_cv_rank_keys, deg, nv, oa);
    if ((rc = sort_u64(c, &oa, &ob, nv, tmp))) return rc;
    LAUNCH(c, "cv_degree", cdiv(nv, 256), 256, k_cv_rank, oa, nv, rank, perm);
    const uint64_t E1 = 2ull * mu;
    uint64_t *da = nullptr, *db = nullptr;
    if ((rc = dalloc(c, &da, E1, &tmp)) || (rc = dalloc(c, &db, E1, &tmp))) return rc;
    LAUNCH(c, "cv_directed", cdiv(count, 256), 256, k_cv_directed, ka, count, keep, pos, rank, da);
    if ((rc = sort_u64(c, &da, &db, E1, tmp))) return rc;
    if ((rc = alloc_graph(c, nv, mu, 0.0))) return rc;
    if ((rc = csr_of_keys(c, da, E1, nv, c->indices, c->adj, nullptr, tmp))) return rc;
    LAUNCH(c, "cv_csr", cdiv(nv, 256), 256, k_fill_u32, (uint32_t *)c->tri, nv);
    LAUNCH(c, "cv_csr", cdiv(nv, 256), 256, k_fill_u32, (uint32_t *)c->tri_deg, nv);
    if ((rc = finish_upload(c))) return rc;
    // 4. triangles per vertex and triangle edges (src/convert.cpp:321-340)
    uint32_t *T = nullptr, *td = nullptr;
    uint8_t *flag = nullptr;
    if ((rc = dalloc(c, &td, nv, &tmp)) || (rc = triangles_core(c, &T, td, &flag, tmp))) return rc;
    // 5. relabel by (triangle edges, id) descending; vertices without a
    //    triangle edge and triangle-free edges dropped (src/convert.cpp:354-414)
    uint32_t *newid = nullptr, *cnt = nullptr;
    if ((rc = dalloc(c, &newid, nv, &tmp)) || (rc = dalloc(c, &cnt, 1, &tmp))) return rc;
    LAUNCH(c, "cv_relabel", cdiv(nv, 256), 256, k_cv_order, td, perm, nv, oa);
    if ((rc = sort_u64(c, &oa, &ob, nv, tmp))) return rc;
    LAUNCH(c, "cv_relabel", cdiv(nv, 256), 256, k_cv_newid, oa, nv, rank, newid);
    LAUNCH(c, "cv_relabel", 1, 64, k_fill_u32, cnt, 1u);
    LAUNCH(c, "cv_relabel", cdiv(nv, 256), 256, k_cv_count_pos, td, nv, cnt);
    uint32_t *kept = nullptr;
    if ((rc = dalloc(c, &kept, E1, &tmp))) return rc;
    LAUNCH(c, "cv_relabel", cdiv(E1, 256), 256, k_cv_flag32, flag, E1, kept);
    if ((rc = check_launch("convert"))) return rc;
    uint32_t n_new = 0, mk = 0;
    HIP_TRY(hipMemcpyAsync(&n_new, cnt, 4, hipMemcpyDeviceToHost, c->stream));
    if ((rc = scan_u32(c, kept, kept, E1, &mk, tmp)) || (rc = stage(info.ms_triangles))) return rc;
    info.edges_final = mk;
    info.vertices_final = n_new;
    if (n_new == 0 || mk == 0) {
        free_graph(c, false);
        return SCD_OK;
    }
    const uint64_t E2 = 2ull * mk;
    uint64_t *fa = nullptr, *fb = nullptr;
    uint32_t *f_idx = nullptr, *f_adj = nullptr;
    int32_t *f_tri = nullptr, *f_td = nullptr;
    if ((rc = dalloc(c, &fa, E2, &tmp)) || (rc = dalloc(c, &fb, E2, &tmp)) ||
        (rc = dalloc(c, &f_idx, (size_t)n_new + 1, &tmp)) || (rc = dalloc(c, &f_adj, E2, &tmp)) ||
        (rc = dalloc(c, &f_tri, n_new, &tmp)) || (rc = dalloc(c, &f_td, n_new, &tmp)))
        return rc;
    LAUNCH(c, "cv_final", cdiv(E1, 256), 256, k_cv_final_keys, da, flag, E1, newid, kept, fa);
    LAUNCH(c, "cv_final", cdiv(nv, 256), 256, k_cv_final_tri, T, td, newid, nv, n_new, f_tri, f_td);
    if ((rc = sort_u64(c, &fa, &fb, E2, tmp))) return rc;
    std::vector<uint32_t> h_idx;
    if ((rc = csr_of_keys(c, fa, E2, n_new, f_idx, f_adj, &h_idx, tmp))) return rc;
    // clustering coefficient in vertex order (src/graph.cpp:57-68)
    std::vector<int32_t> h_tri(n_new);
    HIP_TRY(hipMemcpyAsync(h_tri.data(), f_tri, 4ull * n_new, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double cc = 0.0;
    for (uint32_t p = 0; p < n_new; p++) {
        const uint32_t d = h_idx[p + 1] - h_idx[p];
        if (d > 1) cc += h_tri[p] / (d * (d - 1.0));
    }
    cc /= n_new;
    // 6. the converted graph replaces the intermediate one
    if ((rc = alloc_graph(c, n_new, mk, cc))) return rc;
    HIP_TRY(hipMemcpyAsync(c->indices, f_idx, 4ull * (n_new + 1), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->adj, f_adj, 4ull * E2, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->tri, f_tri, 4ull * n_new, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->tri_deg, f_td, 4ull * n_new, hipMemcpyDeviceToDevice, c->stream));
    if ((rc = finish_upload(c)) || (rc = stage(info.ms_relabel))) return rc;
    *n_out = n_new;
    *m_out = mk;
    *cc_out = cc;
    return finish(c);
}

extern "C" int scd_convert_edges(scd_ctx *c, uint64_t count, const uint32_t *u, const uint32_t *v,
                                 int edges_on_device, uint32_t *n_out, uint32_t *m_out, double *cc_out) {
    if (!c) return fail(SCD_EINVAL, "scd_convert_edges: null context");
    if (count && (!u || !v)) return fail(SCD_EINVAL, "scd_convert_edges: null edge array");
    if (count >= (1ull << 31) - (1ull << 16))
        return fail(SCD_ELIMIT, "scd_convert_edges: %llu edges, at most 2^31 - 2^16 - 1",
                    (unsigned long long)count);
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    free_graph(c);
    std::vector<DevBuf> tmp;
    uint32_t n = 0, m = 0;
    double cc = 0.0;
    int rc = convert_impl(c, count, u, v, edges_on_device, &n, &m, &cc, tmp);
    (void)hipStreamSynchronize(c->stream);
    release_tmp(c, tmp);
    if (rc) {
        free_graph(c);
        return rc;
    }
    if (n_out) *n_out = n;
    if (m_out) *m_out = m;
    if (cc_out) *cc_out = cc;
    return SCD_OK;
}

extern "C" int scd_convert_last_info(const scd_ctx *c, scd_convert_info *out) {
    if (!c || !out) return fail(SCD_EINVAL, "scd_convert_last_info: null argument");
    *out = c->cv_info;
    return SCD_OK;
}

extern "C" int scd_download_graph(scd_ctx *c, uint32_t *indices, uint32_t *adj, int32_t *tri, int32_t *tri_deg) {
    int rc = need_graph(c, "scd_download_graph");
    if (rc) return rc;
    if (indices) HIP_TRY(hipMemcpyAsync(indices, c->indices, 4ull * (c->n + 1), hipMemcpyDeviceToHost, c->stream));
    if (adj && c->E) HIP_TRY(hipMemcpyAsync(adj, c->adj, 4ull * c->E, hipMemcpyDeviceToHost, c->stream));
    if (tri) HIP_TRY(hipMemcpyAsync(tri, c->tri, 4ull * c->n, hipMemcpyDeviceToHost, c->stream));
    if (tri_deg) HIP_TRY(hipMemcpyAsync(tri_deg, c->tri_deg, 4ull * c->n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SCD_OK;
}

extern "C" int scd_graph_shape(const scd_ctx *c, uint32_t *n, uint32_t *m, double *clustering_coef) {
    if (!c) return fail(SCD_EINVAL, "scd_graph_shape: null context");
    const bool g = c->have_graph;
    if (n) *n = g ? c->n : 0;
    if (m) *m = g ? c->m : 0;
    if (clustering_coef) *clustering_coef = g ? c->cc : 0.0;
    return SCD_OK;
}

// ---------------------------------------------------------------------------
// Initial partition (src/scd.cpp:113-171) on the device.
// ---------------------------------------------------------------------------
extern "C" int scd_initial_partition(scd_ctx *c, uint32_t *labels_out, uint32_t *num_labels_out) {
    int rc = need_graph(c, "scd_initial_partition");
    if (rc) return rc;
    const uint32_t n = c->n;
    std::vector<DevBuf> tmp;
    uint64_t *khi[2] = {nullptr, nullptr};
    uint32_t *klo[2] = {nullptr, nullptr}, *val[2] = {nullptr, nullptr}, *kg[2] = {nullptr, nullptr};
    uint32_t *flags = nullptr, *hist = nullptr, *hoff = nullptr, *rank = nullptr, *sf = nullptr, *resume = nullptr;
    unsigned long long *bits = nullptr;
    uint8_t *state = nullptr;
    const uint32_t ntiles = cdiv(n, RS_TILE);
    auto cleanup = [&](int r) {
        release_tmp(c, tmp);
        return r;
    };
    if ((rc = dalloc(c, &khi[0], n, &tmp)) || (rc = dalloc(c, &khi[1], n, &tmp)) ||
        (rc = dalloc(c, &klo[0], n, &tmp)) || (rc = dalloc(c, &klo[1], n, &tmp)) ||
        (rc = dalloc(c, &val[0], n, &tmp)) || (rc = dalloc(c, &val[1], n, &tmp)) ||
        (rc = dalloc(c, &kg[0], n, &tmp)) || (rc = dalloc(c, &kg[1], n, &tmp)) ||
        (rc = dalloc(c, &flags, 4, &tmp)) || (rc = dalloc(c, &bits, 6, &tmp)) ||
        (rc = dalloc(c, &hist, 256ull * ntiles, &tmp)) || (rc = dalloc(c, &hoff, 256ull * ntiles, &tmp)) ||
        (rc = dalloc(c, &rank, n, &tmp)) || (rc = dalloc(c, &sf, n, &tmp)) || (rc = dalloc(c, &state, n, &tmp)) ||
        (rc = dalloc(c, &resume, n, &tmp)))
        return cleanup(rc);
    LAUNCH(c, "ip_clear", 1, 64, k_fill_u32, flags, 4u);
    LAUNCH(c, "ip_clear", 1, 64, k_ip_bits_init, bits);
    LAUNCH(c, "ip_keys", std::min<uint32_t>(cdiv(n, 256), IP_KEYS_BLOCKS), 256, k_ip_keys, c->indices, c->tri,
           c->G ? c->gid : nullptr, n, kg[0],
           khi[0], klo[0], val[0], flags, bits);
    uint32_t hflags[4];
    unsigned long long hbits[6];
    if ((rc = check_launch("initial partition"))) return cleanup(rc);
    if (hipMemcpyAsync(hflags, flags, 16, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipMemcpyAsync(hbits, bits, 48, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return cleanup(fail(SCD_EDEVICE, "scd_initial_partition: readback failed"));
    if (hflags[0])
        return cleanup(fail(SCD_EINVAL, "scd_initial_partition: a vertex of degree < 2 makes the reference's "
                                        "order key NaN (src/scd.cpp:130); convert never produces one"));
    int cur = 0;
    for (int pass = 0; pass < 16; pass++) {
        const uint64_t x = pass < 4    ? (hbits[2] ^ hbits[3]) >> (8 * pass)
                           : pass < 12 ? (hbits[0] ^ hbits[1]) >> (8 * (pass - 4))
                                       : (hbits[4] ^ hbits[5]) >> (8 * (pass - 12));
        if ((x & 0xFF) == 0) continue;  // digit equal in every key
        LAUNCH(c, "ip_sort_hist", ntiles, 256, k_rs_hist, kg[cur], khi[cur], klo[cur], n, pass, ntiles, hist);
        if ((rc = check_launch("radix sort"))) return cleanup(rc);
        uint32_t total = 0;
        if ((rc = scan_u32(c, hist, hoff, 256ull * ntiles, &total, tmp))) return cleanup(rc);
        LAUNCH(c, "ip_sort_scatter", ntiles, 256, k_rs_scatter, kg[cur], khi[cur], klo[cur], val[cur], n, pass,
               ntiles, hoff, kg[1 - cur], khi[1 - cur], klo[1 - cur], val[1 - cur]);
        cur = 1 - cur;
    }
    LAUNCH(c, "ip_rank", cdiv(n, 256), 256, k_ip_rank, val[cur], n, rank, state);
    const uint32_t nmid = c->bin_begin[K_LDS] - c->bin_begin[K_W1];  // 32 < degree <= wave_max: a wave each
    const uint32_t nbig = c->bin_begin[K_NB] - c->bin_begin[K_LDS];  // degree > wave_max: a workgroup each
    // a round after the last decision changes nothing, so on small graphs,
    // where a round costs less than a host check, four rounds go out per
    // check of the open flag (which the last of them sets); a round over 2^24
    // adjacency entries or more costs more than the check it would save
    const uint32_t per_check = c->E < (1ull << 24) ? 4u : 1u;
    LAUNCH(c, "ip_clear", cdiv(n, 256), 256, k_fill_u32, resume, n);
    uint32_t rounds = 0;
    while (true) {
        for (uint32_t r = 0; r < per_check; r++) {
            rounds++;
            LAUNCH(c, "ip_clear", 1, 64, k_fill_u32, flags, 4u);
            LAUNCH(c, "ip_round", cdiv(bin_n(c, K_SUB8), 32), 256, k_ip_round<8>, c->indices, c->adj, rank, state,
                   bin_v(c, K_SUB8), bin_n(c, K_SUB8), flags, resume);
            LAUNCH(c, "ip_round", cdiv(bin_n(c, K_SUB16), 16), 256, k_ip_round<16>, c->indices, c->adj, rank,
                   state, bin_v(c, K_SUB16), bin_n(c, K_SUB16), flags, resume);
            LAUNCH(c, "ip_round", cdiv(bin_n(c, K_SUB32), 8), 256, k_ip_round<32>, c->indices, c->adj, rank,
                   state, bin_v(c, K_SUB32), bin_n(c, K_SUB32), flags, resume);
            if (nmid >= (1u << 19))  // enough vertices to fill the device at 64 per wave
                LAUNCH(c, "ip_round", cdiv(nmid, 256), 256, k_ip_round_wave, c->indices, c->adj, rank, state,
                       bin_v(c, K_W1), nmid, flags, resume);
            else
                LAUNCH(c, "ip_round", cdiv(nmid, 4), 256, k_ip_round<64>, c->indices, c->adj, rank, state,
                       bin_v(c, K_W1), nmid, flags, resume);
            LAUNCH(c, "ip_round", nbig, 256, k_ip_round_block, c->indices, c->adj, rank, state, bin_v(c, K_LDS),
                   nbig, flags, resume);
        }
        if ((rc = check_launch("initial partition"))) return cleanup(rc);
        if (hipMemcpyAsync(hflags, flags, 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return cleanup(fail(SCD_EDEVICE, "scd_initial_partition: readback failed"));
        if (hflags[0] == 0) break;
        if (rounds > n + per_check) return cleanup(fail(SCD_EDEVICE, "scd_initial_partition: no progress"));
    }
    LAUNCH(c, "ip_seedflags", cdiv(n, 256), 256, k_ip_seedflags, val[cur], state, n, sf);
    uint32_t k = 0;
    if ((rc = scan_u32(c, sf, sf, n, &k, tmp))) return cleanup(rc);
    if (c->G) LAUNCH(c, "ip_graph_seeds", cdiv(c->G, 256), 256, k_ip_graph_seeds, sf, c->vtx_off_d, c->G, n, k, c->k_g);
    LAUNCH(c, "ip_labels", cdiv(bin_n(c, K_SUB8), 32), 256, k_ip_labels<8>, c->indices, c->adj, rank, state, sf,
           bin_v(c, K_SUB8), bin_n(c, K_SUB8), c->cur);
    LAUNCH(c, "ip_labels", cdiv(bin_n(c, K_SUB16), 16), 256, k_ip_labels<16>, c->indices, c->adj, rank, state, sf,
           bin_v(c, K_SUB16), bin_n(c, K_SUB16), c->cur);
    LAUNCH(c, "ip_labels", cdiv(bin_n(c, K_SUB32), 8), 256, k_ip_labels<32>, c->indices, c->adj, rank, state, sf,
           bin_v(c, K_SUB32), bin_n(c, K_SUB32), c->cur);
    LAUNCH(c, "ip_labels", cdiv(nmid, 4), 256, k_ip_labels<64>, c->indices, c->adj, rank, state, sf, bin_v(c, K_W1),
           nmid, c->cur);
    LAUNCH(c, "ip_labels", nbig, 256, k_ip_labels_block, c->indices, c->adj, rank, state, sf, bin_v(c, K_LDS), nbig,
           c->cur);
    if ((rc = check_launch("initial partition"))) return cleanup(rc);
    HIP_TRY(hipMemcpyAsync(c->best, c->cur, 4ull * n, hipMemcpyDeviceToDevice, c->stream));
    if (c->num_comms != k) drop_graphs(c);
    c->num_comms = k;
    c->last_ip_rounds = rounds;
    if ((rc = clear_counters(c)) || (rc = phase_stats(c))) return cleanup(rc);
    if (labels_out) HIP_TRY(hipMemcpyAsync(labels_out, c->cur, 4ull * n, hipMemcpyDeviceToHost, c->stream));
    if ((rc = finish(c))) return cleanup(rc);
    c->have_labels = true;
    c->wcc_valid = false;
    if (num_labels_out) *num_labels_out = k;
    return cleanup(SCD_OK);
}

// ---------------------------------------------------------------------------
// Batched arenas: many graphs refined together as one disjoint union.
// Labels are global on the device (graph g's labels start at h_label_off[g],
// preserving each graph's label order and hence its tie-breaks) and local
// (0-based per graph) at the API.
// ---------------------------------------------------------------------------
static int setup_batch(scd_ctx *c, uint32_t G, const uint64_t *vtx_off, const double *cc, bool host_gid);

extern "C" int scd_batch_upload(scd_ctx *c, uint32_t G, const uint64_t *vtx_off, const uint32_t *indices,
                                const uint32_t *adj, const int32_t *tri, const int32_t *tri_deg, const double *cc) {
    if (!c || !vtx_off || !indices || !cc) return fail(SCD_EINVAL, "scd_batch_upload: null argument");
    if (G == 0) return fail(SCD_EINVAL, "scd_batch_upload: no graphs");
    if (vtx_off[0] != 0) return fail(SCD_EINVAL, "scd_batch_upload: vtx_off[0] must be 0");
    for (uint32_t g = 0; g < G; g++)
        if (vtx_off[g + 1] <= vtx_off[g])
            return fail(SCD_EINVAL, "scd_batch_upload: graph %u has no vertices", g);
    if (vtx_off[G] >= 0xFFFFFFFFull) return fail(SCD_ELIMIT, "scd_batch_upload: too many vertices");
    const uint32_t n = (uint32_t)vtx_off[G];
    const uint32_t E = indices[n];
    if (E & 1u) return fail(SCD_EINVAL, "scd_batch_upload: odd adjacency length %u", E);
    int rc = scd_upload(c, n, E / 2, indices, adj, tri, tri_deg, 0.0);
    if (rc) return rc;
    if ((rc = setup_batch(c, G, vtx_off, cc, true))) return rc;
    if ((rc = check_begin(c))) return rc;
    LAUNCH(c, "check_arena", cdiv(n, 4), 256, k_check_arena, c->indices, c->adj, c->gid, c->vtx_off_d, n, c->chk);
    return check_end(c, "scd_batch_upload: an adjacency list leaves its graph's vertex range");
}

// Per-graph state of an arena whose union graph is uploaded: graph of each
// vertex (written here when host_gid is set, else by the caller's kernel),
// clustering coefficients, activity and store flags, WCC reduction items.
// Labels offsets start compact (set by the initial partition).
static int setup_batch(scd_ctx *c, uint32_t G, const uint64_t *vtx_off, const double *cc, bool host_gid) {
    const uint32_t n = c->n;
    int rc;
    std::vector<uint32_t> ws, we, gib(G + 1);
    for (uint32_t g = 0; g < G; g++) {
        gib[g] = (uint32_t)ws.size();
        for (uint64_t s = vtx_off[g]; s < vtx_off[g + 1]; s += 4096) {
            ws.push_back((uint32_t)s);
            we.push_back((uint32_t)vtx_off[g + 1]);
        }
    }
    gib[G] = (uint32_t)ws.size();
    c->n_witems = (uint32_t)ws.size();
    if ((rc = dalloc(c, &c->gid, n)) || (rc = dalloc(c, &c->cc_g, G)) || (rc = dalloc(c, &c->active_g, G)) ||
        (rc = dalloc(c, &c->store_g, G)) || (rc = dalloc(c, &c->k_g, G)) || (rc = dalloc(c, &c->wcc_g, G)) ||
        (rc = dalloc(c, &c->vtx_off_d, G + 1)) || (rc = dalloc(c, &c->witem_start, c->n_witems)) ||
        (rc = dalloc(c, &c->witem_end, c->n_witems)) || (rc = dalloc(c, &c->g_item_begin, G + 1)) ||
        (rc = dalloc(c, &c->partial_g, c->n_witems))) {
        free_graph(c);
        return rc;
    }
    if (c->wcc_g_host_cap < G) {
        const size_t cap = std::max<size_t>(G, 2 * c->wcc_g_host_cap);
        if (c->wcc_g_host) (void)hipHostFree(c->wcc_g_host);
        c->wcc_g_host = nullptr;
        c->wcc_g_host_cap = 0;
        hipError_t e = hipHostMalloc((void **)&c->wcc_g_host, 8ull * cap, hipHostMallocDefault);
        if (e != hipSuccess) {
            c->wcc_g_host = nullptr;
            free_graph(c);
            return fail(SCD_ENOMEM, "hipHostMalloc: %s", hipGetErrorString(e));
        }
        c->wcc_g_host_cap = cap;
    }
    if (host_gid) {
        std::vector<uint32_t> gid(n);
        for (uint32_t g = 0; g < G; g++)
            for (uint64_t v = vtx_off[g]; v < vtx_off[g + 1]; v++) gid[v] = g;
        HIP_TRY(hipMemcpy(c->gid, gid.data(), 4ull * n, hipMemcpyHostToDevice));
    }
    // the small per-graph tables in one copy
    std::vector<uint32_t> pack;
    pack.reserve(4ull * G + 2 * ws.size() + G + 1);
    pack.insert(pack.end(), G, 1u);  // active
    pack.insert(pack.end(), G, 0u);  // store
    pack.insert(pack.end(), ws.begin(), ws.end());
    pack.insert(pack.end(), we.begin(), we.end());
    pack.insert(pack.end(), gib.begin(), gib.end());
    HIP_TRY(hipMemcpy(c->cc_g, cc, 8ull * G, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->vtx_off_d, vtx_off, 8ull * (G + 1), hipMemcpyHostToDevice));
    std::vector<DevBuf> tmp;
    uint32_t *d_pack = nullptr;
    if ((rc = dalloc(c, &d_pack, pack.size(), &tmp))) return rc;
    HIP_TRY(hipMemcpy(d_pack, pack.data(), 4ull * pack.size(), hipMemcpyHostToDevice));
    size_t o = 0;
    HIP_TRY(hipMemcpyAsync(c->active_g, d_pack + o, 4ull * G, hipMemcpyDeviceToDevice, c->stream));
    o += G;
    HIP_TRY(hipMemcpyAsync(c->store_g, d_pack + o, 4ull * G, hipMemcpyDeviceToDevice, c->stream));
    o += G;
    HIP_TRY(hipMemcpyAsync(c->witem_start, d_pack + o, 4ull * ws.size(), hipMemcpyDeviceToDevice, c->stream));
    o += ws.size();
    HIP_TRY(hipMemcpyAsync(c->witem_end, d_pack + o, 4ull * we.size(), hipMemcpyDeviceToDevice, c->stream));
    o += we.size();
    HIP_TRY(hipMemcpyAsync(c->g_item_begin, d_pack + o, 4ull * (G + 1), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    release_tmp(c, tmp);
    c->G = G;
    c->h_vtx_off.assign(vtx_off, vtx_off + G + 1);
    c->h_label_off.assign(G, 0);
    return SCD_OK;
}

// Pinned host staging of at least `bytes` (kept and grown).
static int host_stage(scd_ctx *c, size_t bytes) {
    if (c->h_stage_bytes >= bytes) return SCD_OK;
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    c->h_stage = nullptr;
    const size_t want = std::max(bytes, 2 * c->h_stage_bytes);
    c->h_stage_bytes = 0;
    hipError_t e = hipHostMalloc(&c->h_stage, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        c->h_stage = nullptr;
        return fail(SCD_ENOMEM, "hipHostMalloc: %s", hipGetErrorString(e));
    }
    c->h_stage_bytes = want;
    return SCD_OK;
}

static int need_batch(scd_ctx *c, const char *fn, bool labels) {
    int rc = labels ? need_labels(c, fn) : need_graph(c, fn);
    if (rc) return rc;
    if (!c->G) return fail(SCD_ESTATE, "%s: no batch uploaded", fn);
    return SCD_OK;
}

// Device labels (global) -> local labels per graph.
static int labels_to_local(scd_ctx *c, const uint32_t *dev, uint32_t *out) {
    HIP_TRY(hipMemcpyAsync(out, dev, 4ull * c->n, hipMemcpyDeviceToHost, c->stream));
    int rc = finish(c);
    if (rc) return rc;
    for (uint32_t g = 0; g < c->G; g++) {
        const uint32_t off = c->h_label_off[g];
        for (uint64_t v = c->h_vtx_off[g]; v < c->h_vtx_off[g + 1]; v++)
            if (out[v] != SCD_INVALID_LABEL) out[v] -= off;
    }
    return SCD_OK;
}

extern "C" int scd_batch_initial_partition(scd_ctx *c, uint32_t *labels_out, uint32_t *k_out) {
    int rc = need_batch(c, "scd_batch_initial_partition", false);
    if (rc) return rc;
    uint32_t ktot = 0;
    if ((rc = scd_initial_partition(c, nullptr, &ktot))) return rc;
    std::vector<uint32_t> k(c->G);
    HIP_TRY(hipMemcpy(k.data(), c->k_g, 4ull * c->G, hipMemcpyDeviceToHost));
    uint32_t off = 0;
    for (uint32_t g = 0; g < c->G; g++) {
        c->h_label_off[g] = off;
        off += k[g];
    }
    if (k_out) memcpy(k_out, k.data(), 4ull * c->G);
    if (labels_out) return labels_to_local(c, c->cur, labels_out);
    return SCD_OK;
}

extern "C" int scd_batch_compute_wcc(scd_ctx *c, double *wcc_out) {
    int rc = need_batch(c, "scd_batch_compute_wcc", true);
    if (rc) return rc;
    if ((rc = clear_counters(c)) || (rc = phase_stats(c)) || (rc = phase_wcc(c, nullptr, nullptr))) return rc;
    HIP_TRY(hipMemcpyAsync(c->wcc_g_host, c->wcc_g, 8ull * c->G, hipMemcpyDeviceToHost, c->stream));
    if ((rc = finish(c))) return rc;
    c->wcc_valid = true;
    if (wcc_out) memcpy(wcc_out, c->wcc_g_host, 8ull * c->G);
    return SCD_OK;
}

extern "C" int scd_batch_refine(scd_ctx *c, const double *initial_wcc, uint32_t max_iters, double *best_out,
                                uint32_t *iters_out, double *done_ms, double *wcc_hist) {
    int rc = need_batch(c, "scd_batch_refine", true);
    if (rc) return rc;
    if (!initial_wcc) return fail(SCD_EINVAL, "scd_batch_refine: null initial WCC");
    const uint32_t G = c->G;
    const double threshold = 0.01;
    std::vector<double> best(initial_wcc, initial_wcc + G), done(G, 0.0);
    std::vector<uint32_t> active(G, 1u), store(G, 0u), iters(G, 0u);
    uint32_t remaining = G, it = 0;
    // vertices of the graphs still refining, and in the degree bins
    uint64_t active_v = c->n, binned_v = c->n;
    // on an error: every vertex binned again (best effort), the error kept
    auto unbin = [c](int err) {
        if (c->bins_partial) (void)build_bins(c, false);
        return err;
    };
    const double t0 = wall_ms();
    while (remaining > 0 && it < max_iters) {
        it++;
        HIP_TRY(hipMemcpyAsync(c->active_g, active.data(), 4ull * G, hipMemcpyHostToDevice, c->stream));
        if (2 * active_v <= binned_v) {
            // half of the binned vertices belong to stopped graphs: bin only
            // the others (every kernel over the bins then skips the stopped
            // graphs at no cost), with the stopped graphs' labels in both
            // label buffers
            HIP_TRY(hipMemcpyAsync(c->next, c->cur, 4ull * c->n, hipMemcpyDeviceToDevice, c->stream));
            if ((rc = build_bins(c, true))) return unbin(rc);
            binned_v = active_v;
        }
        if ((rc = scd_improve(c, nullptr, nullptr))) return unbin(rc);
        const double now = wall_ms() - t0;
        bool any_store = false;
        for (uint32_t g = 0; g < G; g++) {
            store[g] = 0;
            if (!active[g]) continue;
            const double w = c->wcc_g_host[g];
            if (wcc_hist) wcc_hist[(size_t)g * max_iters + iters[g]] = w;
            iters[g]++;
            if (w > best[g]) {
                store[g] = 1;
                any_store = true;
            }
            if (w < best[g] * (1 + threshold)) {
                active[g] = 0;
                remaining--;
                active_v -= c->h_vtx_off[g + 1] - c->h_vtx_off[g];
                done[g] = now;
            } else {
                best[g] = w;
            }
        }
        if (any_store) {
            HIP_TRY(hipMemcpyAsync(c->store_g, store.data(), 4ull * G, hipMemcpyHostToDevice, c->stream));
            LAUNCH(c, "store_best", cdiv(c->n, 256), 256, k_store_best_g, c->cur, c->gid, c->store_g, c->n, c->best);
            if ((rc = check_launch("store best")) || (rc = finish(c))) return unbin(rc);
        }
    }
    const double end = wall_ms() - t0;
    for (uint32_t g = 0; g < G; g++)
        if (active[g]) done[g] = end;
    if (best_out) memcpy(best_out, best.data(), 8ull * G);
    if (iters_out) memcpy(iters_out, iters.data(), 4ull * G);
    if (done_ms) memcpy(done_ms, done.data(), 8ull * G);
    // leave every graph active and binned for later calls
    std::vector<uint32_t> ones(G, 1u);
    HIP_TRY(hipMemcpyAsync(c->active_g, ones.data(), 4ull * G, hipMemcpyHostToDevice, c->stream));
    if (c->bins_partial) {
        HIP_TRY(hipMemcpyAsync(c->next, c->cur, 4ull * c->n, hipMemcpyDeviceToDevice, c->stream));
        if ((rc = build_bins(c, false))) return rc;
    }
    return finish(c);
}

extern "C" int scd_batch_labels(scd_ctx *c, uint32_t *labels_out) {
    int rc = need_batch(c, "scd_batch_labels", true);
    if (rc) return rc;
    if (!labels_out) return fail(SCD_EINVAL, "scd_batch_labels: null output");
    return labels_to_local(c, c->best, labels_out);
}

extern "C" int scd_batch_current_labels(scd_ctx *c, uint32_t *labels_out) {
    int rc = need_batch(c, "scd_batch_current_labels", true);
    if (rc) return rc;
    if (!labels_out) return fail(SCD_EINVAL, "scd_batch_current_labels: null output");
    return labels_to_local(c, c->cur, labels_out);
}

// ---------------------------------------------------------------------------
// Arenas assembled from device-resident graphs, step-wise batch refinement
// and per-graph label export: the pieces a server needs for continuous
// batching (graphs join and leave between iterations).
// ---------------------------------------------------------------------------

extern "C" int scd_batch_upload_parts(scd_ctx *c, uint32_t G, const scd_part *parts,
                                      const uint32_t *const *cur_labels, const uint32_t *const *best_labels) {
    if (!c || !parts) return fail(SCD_EINVAL, "scd_batch_upload_parts: null argument");
    if (G == 0) return fail(SCD_EINVAL, "scd_batch_upload_parts: no graphs");
    std::vector<uint64_t> vtx_off(G + 1, 0), eoff(G + 1, 0);
    std::vector<double> cc(G);
    for (uint32_t g = 0; g < G; g++) {
        const scd_part &p = parts[g];
        if (p.n == 0) return fail(SCD_EINVAL, "scd_batch_upload_parts: graph %u has no vertices", g);
        if (!p.indices || (p.m && !p.adj) || !p.tri || !p.tri_deg)
            return fail(SCD_EINVAL, "scd_batch_upload_parts: graph %u: null array", g);
        if (cur_labels && !cur_labels[g])
            return fail(SCD_EINVAL, "scd_batch_upload_parts: graph %u has no labels", g);
        vtx_off[g + 1] = vtx_off[g] + p.n;
        eoff[g + 1] = eoff[g] + 2ull * p.m;
        cc[g] = p.clustering_coef;
    }
    if (vtx_off[G] >= 0xFFFFFFFFull) return fail(SCD_ELIMIT, "scd_batch_upload_parts: too many vertices");
    if (eoff[G] >= (1ull << 32)) return fail(SCD_ELIMIT, "scd_batch_upload_parts: 2m >= 2^32");
    const uint32_t n = (uint32_t)vtx_off[G];
    const uint32_t Eall = (uint32_t)eoff[G];
    int rc = alloc_graph(c, n, Eall / 2, 0.0);
    if (rc) return rc;
    std::vector<ArenaPart> table(G);
    for (uint32_t g = 0; g < G; g++) {
        const scd_part &p = parts[g];
        const uint32_t *cl = cur_labels ? cur_labels[g] : nullptr;
        const uint32_t *bl = cur_labels && best_labels ? best_labels[g] : nullptr;
        table[g] = {p.indices, p.adj, p.tri, p.tri_deg, cl, bl, p.n, (uint32_t)vtx_off[g], 2u * p.m, (uint32_t)eoff[g]};
    }
    std::vector<DevBuf> tmp;
    ArenaPart *d_table = nullptr;
    if ((rc = dalloc(c, &d_table, G, &tmp))) return rc;
    HIP_TRY(hipMemcpyAsync(d_table, table.data(), sizeof(ArenaPart) * G, hipMemcpyHostToDevice, c->stream));
    if ((rc = setup_batch(c, G, vtx_off.data(), cc.data(), false))) {
        release_tmp(c, tmp);
        return rc;
    }
    if ((rc = check_begin(c))) {
        release_tmp(c, tmp);
        return rc;
    }
    // the assembly kernels check the parts as they copy them: CSR offsets
    // from 0 to 2m, adjacency entries and labels below the part's n
    LAUNCH(c, "arena_vertices", cdiv(n + 1ull, ARENA_CHUNK), 256, k_arena_vertices, d_table, G, n, Eall,
           cur_labels ? 1 : 0, c->indices, c->tri, c->tri_deg, c->gid, c->cur, c->best, c->chk);
    if (Eall)
        LAUNCH(c, "arena_entries", cdiv(Eall, ARENA_CHUNK), 256, k_arena_entries, d_table, G, Eall, c->adj, c->chk);
    if ((rc = check_end(c, "scd_batch_upload_parts: a part's CSR offsets, adjacency entries or labels are out "
                           "of range"))) {
        release_tmp(c, tmp);
        return rc;
    }
    rc = finish_upload(c);
    release_tmp(c, tmp);
    if (rc) return rc;
    if (!cur_labels) return SCD_OK;
    // labels of graph g live at [vtx_off[g], vtx_off[g+1]) of the label space
    for (uint32_t g = 0; g < G; g++) c->h_label_off[g] = (uint32_t)vtx_off[g];
    c->num_comms = n;
    if ((rc = clear_counters(c)) || (rc = phase_stats(c)) || (rc = finish(c))) return rc;
    c->have_labels = true;
    c->wcc_valid = false;
    return SCD_OK;
}

// Labels of the graphs listed in outs (already holding vo, n and label
// offsets, outputs and positions) leave the arena in one launch.
static int export_table(scd_ctx *c, const std::vector<ArenaOut> &outs, uint32_t total) {
    if (outs.empty() || total == 0) return SCD_OK;
    std::vector<DevBuf> tmp;
    ArenaOut *d_outs = nullptr;
    int rc = dalloc(c, &d_outs, outs.size(), &tmp);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(d_outs, outs.data(), sizeof(ArenaOut) * outs.size(), hipMemcpyHostToDevice, c->stream));
    LAUNCH(c, "arena_export", cdiv(total, ARENA_CHUNK), 256, k_arena_export, d_outs, (uint32_t)outs.size(), total,
           c->cur, c->best);
    rc = check_launch("label export");
    if (!rc) rc = finish(c);
    release_tmp(c, tmp);
    return rc;
}

extern "C" int scd_batch_export_labels(scd_ctx *c, uint32_t *const *cur_out, uint32_t *const *best_out) {
    int rc = need_batch(c, "scd_batch_export_labels", true);
    if (rc) return rc;
    std::vector<ArenaOut> outs;
    uint32_t total = 0;
    for (uint32_t g = 0; g < c->G; g++) {
        uint32_t *co = cur_out ? cur_out[g] : nullptr, *bo = best_out ? best_out[g] : nullptr;
        if (!co && !bo) continue;
        const uint32_t vo = (uint32_t)c->h_vtx_off[g], ng = (uint32_t)(c->h_vtx_off[g + 1] - vo);
        outs.push_back({co, bo, vo, ng, c->h_label_off[g], total});
        total += ng;
    }
    return export_table(c, outs, total);
}

extern "C" int scd_batch_graph_labels(scd_ctx *c, uint32_t g, int best, uint32_t *labels_out) {
    return scd_batch_graphs_labels(c, 1, &g, best, &labels_out);
}

extern "C" int scd_batch_graphs_labels(scd_ctx *c, uint32_t count, const uint32_t *graphs, int best,
                                       uint32_t *const *labels_out) {
    int rc = need_batch(c, "scd_batch_graphs_labels", true);
    if (rc) return rc;
    if (count && (!graphs || !labels_out)) return fail(SCD_EINVAL, "scd_batch_graphs_labels: null argument");
    std::vector<ArenaOut> outs(count);
    uint64_t total = 0;
    for (uint32_t i = 0; i < count; i++) {
        const uint32_t g = graphs[i];
        if (g >= c->G || !labels_out[i])
            return fail(SCD_EINVAL, "scd_batch_graphs_labels: bad graph or null output (entry %u)", i);
        const uint32_t vo = (uint32_t)c->h_vtx_off[g], ng = (uint32_t)(c->h_vtx_off[g + 1] - vo);
        outs[i] = {nullptr, nullptr, vo, ng, c->h_label_off[g], (uint32_t)total};
        total += ng;
        if (total >= (1ull << 32)) return fail(SCD_ELIMIT, "scd_batch_graphs_labels: 2^32 labels or more at once");
    }
    if (total == 0) return SCD_OK;
    // one device staging area and one copy to pinned host memory for all
    std::vector<DevBuf> tmp;
    uint32_t *stage = nullptr;
    if ((rc = dalloc(c, &stage, total, &tmp))) return rc;
    for (uint32_t i = 0; i < count; i++) (best ? outs[i].best : outs[i].cur) = stage + outs[i].so;
    if ((rc = export_table(c, outs, (uint32_t)total)) || (rc = host_stage(c, 4 * total))) {
        release_tmp(c, tmp);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(c->h_stage, stage, 4 * total, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    release_tmp(c, tmp);
    if (e != hipSuccess) return fail(SCD_EDEVICE, "scd_batch_graphs_labels: %s", hipGetErrorString(e));
    const uint32_t *h = (const uint32_t *)c->h_stage;
    for (uint32_t i = 0; i < count; i++) memcpy(labels_out[i], h + outs[i].so, 4ull * outs[i].n);
    return SCD_OK;
}

extern "C" int scd_batch_step(scd_ctx *c, const uint32_t *active, double *wcc_out) {
    int rc = need_batch(c, "scd_batch_step", true);
    if (rc) return rc;
    if (!active) return fail(SCD_EINVAL, "scd_batch_step: null activity flags");
    HIP_TRY(hipMemcpyAsync(c->active_g, active, 4ull * c->G, hipMemcpyHostToDevice, c->stream));
    if ((rc = scd_improve(c, nullptr, nullptr))) return rc;
    if (wcc_out) memcpy(wcc_out, c->wcc_g_host, 8ull * c->G);
    return SCD_OK;
}

extern "C" int scd_batch_store_best(scd_ctx *c, const uint32_t *store) {
    int rc = need_batch(c, "scd_batch_store_best", true);
    if (rc) return rc;
    if (!store) return fail(SCD_EINVAL, "scd_batch_store_best: null flags");
    HIP_TRY(hipMemcpyAsync(c->store_g, store, 4ull * c->G, hipMemcpyHostToDevice, c->stream));
    LAUNCH(c, "store_best", cdiv(c->n, 256), 256, k_store_best_g, c->cur, c->gid, c->store_g, c->n, c->best);
    if ((rc = check_launch("store best"))) return rc;
    return finish(c);
}

// ---------------------------------------------------------------------------
// Autotuner: the tuning (degree-bin thresholds, tile sizes) with the fastest
// refinement iteration, per graph-size class.
// ---------------------------------------------------------------------------

namespace {
struct TuneEntry {
    scd_tuning t;
    double ms;
};
std::mutex g_tune_mu;
std::map<std::string, TuneEntry> g_tune_cache;

uint32_t ilog2u(uint64_t x) {
    uint32_t r = 0;
    while (x > 1) {
        x >>= 1;
        r++;
    }
    return r;
}

bool read_tune_file(const char *path, const std::string &key, TuneEntry &out) {
    FILE *f = fopen(path, "r");
    if (!f) return false;
    char line[512], k[256];
    bool found = false;
    while (fgets(line, sizeof(line), f)) {
        // key lds_tile merge_chunk wave_max hub_table lds_hash lanes dense_wcc ms
        // (lines without dense_wcc read it as 0, the default)
        double v[8];
        const int got = sscanf(line, "%255s %lf %lf %lf %lf %lf %lf %lf %lf", k, &v[0], &v[1], &v[2], &v[3], &v[4],
                               &v[5], &v[6], &v[7]);
        if (got != 8 && got != 9) continue;
        const unsigned Dn = got == 9 ? (unsigned)v[6] : 0u;
        const double ms = got == 9 ? v[7] : v[6];
        if (key == k) {  // the last line of a class wins
            out.t = scd_tuning{(unsigned)v[0], (unsigned)v[1], (unsigned)v[2], (unsigned)v[3], (unsigned)v[4],
                               (unsigned)v[5], Dn, 0};
            out.ms = ms;
            found = true;
        }
    }
    fclose(f);
    return found;
}
}  // namespace

extern "C" int scd_autotune(scd_ctx *c, const char *cache_path, int reps, scd_tuning *chosen,
                            double *iteration_ms) {
    int rc = need_labels(c, "scd_autotune");
    if (rc) return rc;
    if (reps < 1) reps = 3;
    const uint32_t dmax = c->dmax;
    const uint64_t dsum = c->E;
    char key[256];
    snprintf(key, sizeof(key), "%s:%ucu:n%u:dmax%u:davg%u", c->arch.c_str(), c->num_cus, ilog2u(c->n),
             ilog2u(dmax), ilog2u(dsum / std::max<uint32_t>(c->n, 1)));
    TuneEntry hit;
    bool have = false;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            hit = it->second;
            have = true;
        }
    }
    if (!have && cache_path && read_tune_file(cache_path, key, hit)) {
        have = true;
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = hit;
    }
    if (have) {
        if ((rc = scd_set_tuning(c, &hit.t))) return rc;
        if (chosen) *chosen = c->tune;
        if (iteration_ms) *iteration_ms = hit.ms;
        return SCD_OK;
    }

    // candidates: only knobs that change a bin or tile for these degrees,
    // the defaults first
    std::vector<uint32_t> Ws = {1024};
    if (dmax > 256) Ws = {1024, 512, 256};
    std::vector<uint32_t> Ts = {2048};
    if (dmax > 2048) Ts = {4096, 2048, 8192, 16384};
    std::vector<scd_tuning> cand;
    for (uint32_t T : Ts)
        for (uint32_t W : Ws) {
            std::vector<uint32_t> Cs = {std::min<uint32_t>(2048, 2 * T)};
            std::vector<uint32_t> Hs = {0};
            if (dmax > T) {
                Cs = {1024, 2048, 4096};
                Hs = {0, 1, 2};
            }
            std::vector<uint32_t> Ls = {0};
            if (T <= 4096 && dmax > W) Ls = {0, 1};
            std::vector<uint32_t> Ns = {1};
            if (c->E >= (1ull << 20)) Ns = {0, 1};
            for (uint32_t C : Cs)
                for (uint32_t H : Hs) {
                    // with the buckets the chunk size only shapes the hub
                    // statistics filter: its two largest
                    if (H == 2 && Cs.size() > 1 && C == Cs[0]) continue;
                    for (uint32_t Lh : Ls)
                        for (uint32_t N : Ns) cand.push_back(scd_tuning{T, C, W, H, Lh, N, 0, 0});
                }
        }

    // the labels, statistics and best labels to start every timed iteration from
    std::vector<DevBuf> tmp;
    uint32_t *s_cur = nullptr, *s_best = nullptr;
    int4 *s_stats = nullptr;
    const uint32_t nc = c->num_comms;
    if ((rc = dalloc(c, &s_cur, c->n, &tmp)) || (rc = dalloc(c, &s_best, c->n, &tmp)) ||
        (rc = dalloc(c, &s_stats, std::max<uint32_t>(nc, 1), &tmp))) {
        release_tmp(c, tmp);
        return rc;
    }
    auto copy_state = [&](bool save) -> int {
        if (save) {
            HIP_TRY(hipMemcpyAsync(s_cur, c->cur, 4ull * c->n, hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(hipMemcpyAsync(s_best, c->best, 4ull * c->n, hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(hipMemcpyAsync(s_stats, c->stats, sizeof(int4) * nc, hipMemcpyDeviceToDevice, c->stream));
        } else {
            HIP_TRY(hipMemcpyAsync(c->cur, s_cur, 4ull * c->n, hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(hipMemcpyAsync(c->best, s_best, 4ull * c->n, hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(hipMemcpyAsync(c->stats, s_stats, sizeof(int4) * nc, hipMemcpyDeviceToDevice, c->stream));
            if (c->num_comms != nc) drop_graphs(c);
            c->num_comms = nc;
            c->wcc_valid = false;
        }
        return finish(c);
    };
    const scd_tuning orig = c->tune;
    scd_tuning best_t = orig;
    double best_ms = 1e300;
    if ((rc = copy_state(true))) {
        release_tmp(c, tmp);
        return rc;
    }
    // median of `runs` timed iterations under tuning t (an untimed run
    // first captures the iteration's graphs)
    auto time_tuning = [&](const scd_tuning &t, int runs, double *med) -> int {
        int r = SCD_OK;
        if ((r = scd_set_tuning(c, &t)) || (r = copy_state(false)) || (r = scd_improve(c, nullptr, nullptr)))
            return r;
        std::vector<double> ms;
        for (int k = 0; k < runs; k++) {
            if ((r = copy_state(false))) return r;
            const double t0 = wall_ms();
            if ((r = scd_improve(c, nullptr, nullptr))) return r;
            ms.push_back(wall_ms() - t0);
        }
        std::sort(ms.begin(), ms.end());
        *med = ms[ms.size() / 2];
        return SCD_OK;
    };
    // every shape once, then the three fastest again with twice the runs
    // (the fastest median of those wins: shapes a few percent apart are
    // told apart by the longer medians, not by the order they were tried
    // in), then the winner with the dense-community WCC, which must be
    // kTuneMargin faster to replace it
    std::vector<std::pair<double, scd_tuning>> timed;
    for (size_t i = 0; i < cand.size() && rc == SCD_OK; i++) {
        double med = 0.0;
        if ((rc = time_tuning(cand[i], reps, &med)) == SCD_OK) timed.emplace_back(med, c->tune);
        else if (rc == SCD_ENOMEM || rc == SCD_ELIMIT) rc = SCD_OK;  // a shape whose buffers do not fit: not a candidate
    }
    if (rc == SCD_OK) {
        std::stable_sort(timed.begin(), timed.end(),
                         [](const std::pair<double, scd_tuning> &x, const std::pair<double, scd_tuning> &y) {
                             return x.first < y.first;
                         });
        for (size_t k = 0; k < std::min<size_t>(3, timed.size()) && rc == SCD_OK; k++) {
            double med = 0.0;
            if ((rc = time_tuning(timed[k].second, 2 * reps + 1, &med)) == SCD_OK && med < best_ms) {
                best_ms = med;
                best_t = c->tune;
            }
        }
    }
    if (rc == SCD_OK) {
        scd_tuning t = best_t;
        t.dense_wcc = 1;
        double med = 0.0;
        if ((rc = time_tuning(t, reps, &med)) == SCD_OK && med < best_ms * (1.0 - kTuneMargin)) {
            best_ms = med;
            best_t = c->tune;
        }
    }
    if (rc == SCD_OK) rc = scd_set_tuning(c, &best_t);
    else (void)scd_set_tuning(c, &orig);
    const int rc2 = copy_state(false);
    release_tmp(c, tmp);
    if (rc) return rc;
    if (rc2) return rc2;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = TuneEntry{best_t, best_ms};
    }
    if (cache_path) {
        FILE *f = fopen(cache_path, "a");
        if (!f) return fail(SCD_EIO, "scd_autotune: cannot append to %s", cache_path);
        fprintf(f, "%s %u %u %u %u %u %u %u %.6f\n", key, best_t.lds_tile, best_t.merge_chunk, best_t.wave_max,
                best_t.hub_table, best_t.lds_hash, best_t.lanes, best_t.dense_wcc, best_ms);
        fclose(f);
    }
    if (chosen) *chosen = best_t;
    if (iteration_ms) *iteration_ms = best_ms;
    return SCD_OK;
}

