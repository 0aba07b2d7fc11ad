// Serving loop over the batched engine (include/scd_serve.h).
//
// Every admitted graph is held on the device in its own buffers (CSR with
// local ids, current and best labels).  Two engine contexts share the
// device: `prep` computes the initial partition and initial WCC of newly
// admitted graphs as one arena, `main` refines all joined graphs as one
// arena.  A join re-assembles `main` from the resident graphs (the finished
// ones drop out, the new ones come in with their initial labels, the others
// carry their current and best labels over), so the iteration loop itself
// never changes shape.  Each graph follows the stopping rule of
// src/main-gpu.cpp:71-96 on its own WCC values, which are those of the
// single-graph path (the batched engine guarantees it), so its result does
// not depend on which graphs it was served with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "scd_hip.h"
#include "scd_host.h"
#include "scd_serve.h"

namespace {

thread_local char g_err[512];

int sfail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

// an engine call failed: keep its message
int engine_fail(int rc, const char *what) {
    snprintf(g_err, sizeof(g_err), "%s: %s", what, scd_last_error());
    return rc;
}

enum State { QUEUED = 0, RESIDENT = 1, REFINING = 2, DONE = 3, FAILED = 4, UPLOADING = 5 };

// One device allocation shared by the graphs of an intake batch; it goes
// back to the pool when the last of them leaves.
struct Block {
    void *p = nullptr;
    uint64_t bytes = 0;
    uint32_t refs = 0;
};

struct Req {
    scd_request in{};
    int state = QUEUED;
    struct Block *blk = nullptr;  // the device allocation holding the arrays below
    uint32_t *d_ind = nullptr, *d_adj = nullptr, *d_cur = nullptr, *d_best = nullptr;
    int32_t *d_tri = nullptr, *d_td = nullptr;
    uint64_t bytes = 0;
    double best = 0.0;  // best WCC so far (the loop's `best`)
    scd_result res{};
    std::vector<uint32_t> labels;
};

const double kThreshold = 0.01;  // src/main-gpu.cpp:88, stop below best * (1 + 0.01)
// intake batches: small enough that the first graphs refine while later
// ones are copied and prepared, large enough to amortize each batch's
// preparation and arena assembly (measured on a 200-graph mix of planted
// graphs and R-MAT s18-22: 64 / 2 GiB beat 256 / 8 GiB by 19% in capacity)
const size_t kIntakeGraphs = 64;
const uint64_t kIntakeBytes = 2ull << 30;
// uploads go through two pinned staging buffers in turn: host threads fill
// one while the other is copied to the device
const size_t kStageBytes = 64ull << 20;
// device blocks of finished graphs are kept for later admissions (hipFree
// waits for the device, so the serving loop never calls it): at most
// kPoolBytes (or an eighth of the HBM budget) stay pooled, the intake
// thread frees the rest
const uint64_t kPoolBytes = 4ull << 30;
const int kCopyThreads = 12;

}  // namespace

struct scd_server {
    scd_server_opts o{};
    scd_ctx *main = nullptr, *prep = nullptr;
    std::vector<Req> reqs;
    std::vector<size_t> arena;     // request of each graph of the main arena
    std::vector<uint32_t> active;  // per arena graph
    uint32_t n_active = 0;
    std::vector<size_t> waiting;   // resident, not joined yet (arrival order)
    uint64_t resident = 0;
    scd_server_stats st{};
    std::chrono::steady_clock::time_point t0;

    // uploads and preparation run on the intake thread (own stream, the
    // preparation context), overlapping the refinement on the device; the
    // main loop drains finished ones
    std::thread up_thread;
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::deque<size_t> up_queue;                 // requests to upload (guarded by mu)
    std::vector<std::pair<size_t, int>> up_done;  // (request, status) (guarded by mu)
    // uploaded, unprepared batches: the intake thread hands a batch over
    // as it is when the main loop has nothing to refine (main_idle), and
    // the main loop prepares it in its own arena, which then needs no
    // second assembly (guarded by mu)
    std::vector<std::vector<size_t>> up_raw;
    std::atomic<bool> main_idle{false};
    std::mutex prep_mu;  // the preparation context, used by either thread
    bool up_stop = false;                        // guarded by mu
    hipStream_t up_stream = nullptr;
    void *pin[2] = {nullptr, nullptr};        // pinned staging (intake thread)
    hipEvent_t pin_ev[2] = {nullptr, nullptr};  // last copy out of each
    bool pin_busy[2] = {false, false};
    int pin_next = 0;
    std::mutex pool_mu;
    std::vector<std::pair<void *, uint64_t>> pool;  // free device blocks (guarded by pool_mu)
    uint64_t pool_bytes = 0;                        // guarded by pool_mu
    // with graph_mode the main context captures iterations, and device
    // allocations or copies issued by another thread during a capture make
    // it fail: intake batches and capturing steps exclude each other
    std::mutex capture_mu;
    size_t uploading = 0;  // main thread only
    double upload_ms = 0, prep_ms = 0;  // intake thread only
    double main_prep_ms = 0;            // main thread only

    double now() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

extern "C" const char *scd_server_last_error(void) { return g_err; }

extern "C" uint64_t scd_server_footprint(uint32_t n, uint32_t m) {
    // resident: CSR (n + 1 + 2m) and tri, tri_deg, two label arrays (4n);
    // arena share: the engine's per-vertex (about 100 B) and per-entry
    // (adj, internal adjacency, flags: 9 B) buffers.  The hub buffers come on
    // top (about 40 B per adjacency entry of a vertex with more than
    // lds_tile neighbours: two scratch copies and the bucket regions), as
    // the degrees are not known here.
    const uint64_t N = n, E = 2ull * m;
    return 4 * (N + 1 + E) + 16 * N + 104 * N + 9 * E;
}

extern "C" int scd_server_create(const scd_server_opts *opts, scd_server **out) {
    if (!opts || !out) return sfail(SCD_EINVAL, "scd_server_create: null argument");
    *out = nullptr;
    auto *s = new scd_server();
    s->o = *opts;
    if (s->o.max_iters == 0) s->o.max_iters = 1000;
    if (s->o.join_wait_ms < 0) s->o.join_wait_ms = 0;
    int rc = scd_init(opts->device, 256, 0, &s->main);
    if (rc == SCD_OK) rc = scd_init(opts->device, 256, 0, &s->prep);
    if (rc != SCD_OK) {
        rc = engine_fail(rc, "scd_server_create");
        scd_server_destroy(s);
        return rc;
    }
    scd_set_graph_mode(s->main, opts->graph_mode ? 1 : 0);
    scd_set_graph_mode(s->prep, 0);
    // arenas of many graphs fill the device bin by bin: the per-bin kernels
    // run on one stream (concurrent bin lanes only added synchronization:
    // 8% slower iterations on a 185-graph planted arena), the hubs of
    // several R-MAT graphs take the label-range buckets (on a 200-graph mix
    // of planted graphs and R-MAT s18-22 the hash tables made arena
    // iterations 15% faster than merge passes over all hub chunks, and the
    // buckets 8% faster again), and the planted graphs' dense communities
    // count their internal triangles by bitmap rows
    for (scd_ctx *c : {s->main, s->prep}) {
        scd_tuning t;
        if (scd_get_tuning(c, &t) == SCD_OK) {
            t.lanes = 0;
            t.hub_table = 2;
            t.dense_wcc = 1;
            (void)scd_set_tuning(c, &t);
        }
    }
    if (hipSetDevice(opts->device) != hipSuccess ||
        hipStreamCreateWithFlags(&s->up_stream, hipStreamNonBlocking) != hipSuccess) {
        scd_server_destroy(s);
        return sfail(SCD_EDEVICE, "scd_server_create: cannot create the upload stream");
    }
    for (int b = 0; b < 2; b++) {
        if (hipHostMalloc(&s->pin[b], kStageBytes, hipHostMallocDefault) != hipSuccess) s->pin[b] = nullptr;
        if (s->pin[b] && hipEventCreateWithFlags(&s->pin_ev[b], hipEventDisableTiming) != hipSuccess)
            s->pin_ev[b] = nullptr;
        if (!s->pin[b] || !s->pin_ev[b]) {
            scd_server_destroy(s);
            return sfail(SCD_ENOMEM, "scd_server_create: pinned staging allocation failed");
        }
    }
    if (s->o.hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipSetDevice(opts->device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess) {
            scd_server_destroy(s);
            return sfail(SCD_EDEVICE, "scd_server_create: cannot query device memory");
        }
        s->o.hbm_budget = (uint64_t)(0.6 * (double)fr);
    }
    *out = s;
    return SCD_OK;
}

static void stop_uploader(scd_server *s);

static uint64_t pool_cap(const scd_server *s) { return std::min(kPoolBytes, s->o.hbm_budget / 8); }

// Drops a request's share of its device block (the block returns to the
// pool with its last graph) and of the budget.
static void release(scd_server *s, Req &r) {
    if (r.blk && --r.blk->refs == 0) {
        std::lock_guard<std::mutex> lk(s->pool_mu);
        s->pool.emplace_back(r.blk->p, r.blk->bytes);
        s->pool_bytes += r.blk->bytes;
        delete r.blk;
    }
    r.blk = nullptr;
    r.d_ind = r.d_adj = r.d_cur = r.d_best = nullptr;
    r.d_tri = r.d_td = nullptr;
    if (r.bytes) s->resident -= r.bytes;
    r.bytes = 0;
}

// A pooled block of at least `bytes` (at most twice that plus 1 MiB), or null.
static void *pool_take(scd_server *s, uint64_t bytes, uint64_t *got) {
    std::lock_guard<std::mutex> lk(s->pool_mu);
    size_t best = s->pool.size();
    for (size_t i = 0; i < s->pool.size(); i++) {
        const uint64_t b = s->pool[i].second;
        if (b >= bytes && b <= 2 * bytes + (1u << 20) && (best == s->pool.size() || b < s->pool[best].second))
            best = i;
    }
    if (best == s->pool.size()) return nullptr;
    void *p = s->pool[best].first;
    *got = s->pool[best].second;
    s->pool_bytes -= *got;
    s->pool[best] = s->pool.back();
    s->pool.pop_back();
    return p;
}

// Frees pooled blocks, largest first, down to `keep` bytes (intake thread
// and destruction only).
static void pool_trim(scd_server *s, uint64_t keep) {
    std::vector<void *> drop;
    {
        std::lock_guard<std::mutex> lk(s->pool_mu);
        std::sort(s->pool.begin(), s->pool.end(), [](const auto &a, const auto &b) { return a.second < b.second; });
        while (s->pool_bytes > keep && !s->pool.empty()) {
            drop.push_back(s->pool.back().first);
            s->pool_bytes -= s->pool.back().second;
            s->pool.pop_back();
        }
    }
    for (void *p : drop) (void)hipFree(p);
}

extern "C" void scd_server_destroy(scd_server *s) {
    if (!s) return;
    stop_uploader(s);
    for (auto &r : s->reqs) release(s, r);
    pool_trim(s, 0);
    for (int b = 0; b < 2; b++) {
        if (s->pin[b]) (void)hipHostFree(s->pin[b]);
        if (s->pin_ev[b]) (void)hipEventDestroy(s->pin_ev[b]);
    }
    if (s->up_stream) (void)hipStreamDestroy(s->up_stream);
    if (s->main) scd_destroy(s->main);
    if (s->prep) scd_destroy(s->prep);
    delete s;
}

extern "C" int scd_server_submit(scd_server *s, const scd_request *req, uint64_t *id_out) {
    if (!s || !req) return sfail(SCD_EINVAL, "scd_server_submit: null argument");
    const scd_request &q = *req;
    if (q.n == 0) return sfail(SCD_EINVAL, "scd_server_submit: graph has no vertices");
    if (!q.indices || (q.m && !q.adj) || !q.tri || !q.tri_deg)
        return sfail(SCD_EINVAL, "scd_server_submit: null array");
    if (q.m >= (1u << 31)) return sfail(SCD_ELIMIT, "scd_server_submit: m = %u needs 2m < 2^32", q.m);
    if (q.indices[0] != 0 || q.indices[q.n] != 2ull * q.m)
        return sfail(SCD_EINVAL, "scd_server_submit: indices[0] = %u, indices[n] = %u, expected 0 and %llu",
                     q.indices[0], q.indices[q.n], 2ull * q.m);
    // the initial partition's key needs degree >= 2 (src/scd.cpp:130)
    for (uint32_t v = 0; v < q.n; v++)
        if (q.indices[v + 1] < q.indices[v] + 2)
            return sfail(SCD_EINVAL, "scd_server_submit: vertex %u has degree < 2 (not a convert output)", v);
    // every adjacency entry must be a vertex id: the kernels index by them
    const int64_t E = 2ll * q.m;
    int64_t bad = E;
#pragma omp parallel for reduction(min : bad) if (E > (1 << 20))
    for (int64_t j = 0; j < E; j++)
        if (q.adj[j] >= q.n && j < bad) bad = j;
    if (bad < E)
        return sfail(SCD_EINVAL, "scd_server_submit: adjacency entry %lld is %u, not below n = %u", (long long)bad,
                     q.adj[bad], q.n);
    Req r;
    r.in = q;
    r.res.status = SCD_OK;
    r.res.arrival_ms = q.arrival_ms;
    s->reqs.push_back(std::move(r));
    if (id_out) *id_out = s->reqs.size() - 1;
    return SCD_OK;
}

// Copies the graphs of an intake batch to the device: one allocation for
// all of them (each graph's indices, adj, tri, tri_deg together, 256-byte
// aligned, then every graph's two label arrays), filled through the two
// pinned staging buffers in turn, kStageBytes at a time: host threads pack
// a piece from the graphs' host arrays while the previous piece is copied
// asynchronously on the upload stream.  Returns each request's status in
// rc (the batch fails or succeeds as a whole).
static void admit_batch(scd_server *s, const std::vector<size_t> &batch, std::vector<int> &rc) {
    struct Seg {
        const char *src;
        uint64_t bytes, off;
    };
    auto al = [](uint64_t b) { return (std::max<uint64_t>(b, 4) + 255) & ~255ull; };
    std::vector<Seg> segs;
    std::vector<uint64_t> off_csr(batch.size()), off_lab(batch.size());
    uint64_t csr = 0;
    for (size_t k = 0; k < batch.size(); k++) {
        const scd_request &q = s->reqs[batch[k]].in;
        const uint64_t N = q.n, E = 2ull * q.m;
        off_csr[k] = csr;
        segs.push_back({(const char *)q.indices, 4 * (N + 1), csr});
        csr += al(4 * (N + 1));
        segs.push_back({(const char *)q.adj, 4 * E, csr});
        csr += al(4 * E);
        segs.push_back({(const char *)q.tri, 4 * N, csr});
        csr += al(4 * N);
        segs.push_back({(const char *)q.tri_deg, 4 * N, csr});
        csr += al(4 * N);
    }
    uint64_t total = csr;
    for (size_t k = 0; k < batch.size(); k++) {
        off_lab[k] = total;
        total += 2 * al(4ull * s->reqs[batch[k]].in.n);
    }
    rc.assign(batch.size(), SCD_OK);
    auto fail_all = [&](int code, const char *msg) {
        sfail(code, "%s", msg);
        rc.assign(batch.size(), code);
    };
    uint64_t got = 0;
    void *p = pool_take(s, total, &got);
    if (!p) {
        hipError_t e = hipMalloc(&p, total);
        if (e != hipSuccess) {  // pooled blocks may hold the memory: free them all and try once more
            (void)hipGetLastError();
            pool_trim(s, 0);
            e = hipMalloc(&p, total);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail_all(SCD_ENOMEM, "admission: device allocation failed");
        }
        got = total;
    }
    char *base = (char *)p;
    bool ok = true;
    size_t first = 0;  // first segment that may overlap the piece
    for (uint64_t a = 0; ok && a < csr; a += kStageBytes) {
        const uint64_t b = std::min<uint64_t>(csr, a + kStageBytes);
        const int k = s->pin_next;
        s->pin_next ^= 1;
        if (s->pin_busy[k] && hipEventSynchronize(s->pin_ev[k]) != hipSuccess) {
            ok = false;
            break;
        }
        char *buf = (char *)s->pin[k];
        // the piece's copies, cut into 256 KiB jobs shared by the threads
        struct Job {
            char *dst;
            const char *src;
            uint64_t bytes;
        };
        std::vector<Job> jobs;
        while (first < segs.size() && segs[first].off + segs[first].bytes <= a) first++;
        for (size_t i = first; i < segs.size() && segs[i].off < b; i++) {
            const Seg &g = segs[i];
            const uint64_t lo = std::max(a, g.off), hi = std::min(b, g.off + g.bytes);
            for (uint64_t x = lo; x < hi; x += 256u << 10)
                jobs.push_back({buf + (x - a), g.src + (x - g.off), std::min<uint64_t>(256u << 10, hi - x)});
        }
#pragma omp parallel for num_threads(kCopyThreads) schedule(dynamic, 4)
        for (long j = 0; j < (long)jobs.size(); j++) memcpy(jobs[j].dst, jobs[j].src, jobs[j].bytes);
        if (hipMemcpyAsync(base + a, buf, b - a, hipMemcpyHostToDevice, s->up_stream) != hipSuccess ||
            hipEventRecord(s->pin_ev[k], s->up_stream) != hipSuccess)
            ok = false;
        else
            s->pin_busy[k] = true;
    }
    if (!ok || hipStreamSynchronize(s->up_stream) != hipSuccess) {
        (void)hipStreamSynchronize(s->up_stream);
        (void)hipFree(p);
        return fail_all(SCD_EDEVICE, "admission: copy to the device failed");
    }
    Block *blk = new Block{p, got, (uint32_t)batch.size()};
    for (size_t k = 0; k < batch.size(); k++) {
        Req &r = s->reqs[batch[k]];
        const uint64_t N = r.in.n, E = 2ull * r.in.m, o = off_csr[k];
        r.blk = blk;
        r.d_ind = (uint32_t *)(base + o);
        r.d_adj = (uint32_t *)(base + o + al(4 * (N + 1)));
        r.d_tri = (int32_t *)(base + o + al(4 * (N + 1)) + al(4 * E));
        r.d_td = (int32_t *)(base + o + al(4 * (N + 1)) + al(4 * E) + al(4 * N));
        r.d_cur = (uint32_t *)(base + off_lab[k]);
        r.d_best = (uint32_t *)(base + off_lab[k] + al(4 * N));
    }
}

static scd_part part_of(const Req &r);

static void set_initial(scd_server *s, const std::vector<size_t> &batch, const std::vector<double> &w0) {
    for (size_t i = 0; i < batch.size(); i++) {
        Req &r = s->reqs[batch[i]];
        r.best = w0[i];
        r.res.initial_wcc = w0[i];
    }
}

// Initial partition and initial WCC of a batch of resident graphs in the
// preparation context (caller holds prep_mu); their labels go to their own
// buffers.
static int prepare(scd_server *s, const std::vector<size_t> &batch) {
    const uint32_t P = (uint32_t)batch.size();
    std::vector<scd_part> parts(P);
    std::vector<uint32_t *> cur(P), best(P);
    for (uint32_t i = 0; i < P; i++) {
        const Req &r = s->reqs[batch[i]];
        parts[i] = part_of(r);
        cur[i] = r.d_cur;
        best[i] = r.d_best;
    }
    std::vector<double> w0(P);
    int rc;
    if ((rc = scd_batch_upload_parts(s->prep, P, parts.data(), nullptr, nullptr)) ||
        (rc = scd_batch_initial_partition(s->prep, nullptr, nullptr)) ||
        (rc = scd_batch_compute_wcc(s->prep, w0.data())) ||
        (rc = scd_batch_export_labels(s->prep, cur.data(), best.data())))
        return engine_fail(rc, "initial partition of joining graphs");
    set_initial(s, batch, w0);
    return SCD_OK;
}

// The intake thread: takes every admitted request queued so far, copies
// each to the device (its own stream), prepares them together in the
// preparation context, and hands them to the serving loop — all of it
// overlapping the refinement of the joined graphs on the device.
static void intake(scd_server *s) {
    (void)hipSetDevice(s->o.device);
    while (true) {
        std::vector<size_t> batch;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->up_stop || !s->up_queue.empty(); });
            if (s->up_queue.empty()) return;  // stop requested, nothing left
            // bounded batches: the first graphs start refining while later
            // ones are still being copied and prepared (each batch costs one
            // preparation and one arena assembly, so batches stay large)
            uint64_t bytes = 0;
            const size_t max_graphs = s->o.intake_graphs ? s->o.intake_graphs : kIntakeGraphs;
            const uint64_t max_bytes = s->o.intake_bytes ? s->o.intake_bytes : kIntakeBytes;
            while (!s->up_queue.empty() && batch.size() < max_graphs && bytes < max_bytes) {
                bytes += s->reqs[s->up_queue.front()].bytes;
                batch.push_back(s->up_queue.front());
                s->up_queue.pop_front();
            }
        }
        std::vector<std::pair<size_t, int>> done;
        std::vector<size_t> ok;
        std::unique_lock<std::mutex> cap(s->capture_mu, std::defer_lock);
        if (s->o.graph_mode) cap.lock();
        pool_trim(s, pool_cap(s));  // hipFree: not while the main context captures
        auto t = std::chrono::steady_clock::now();
        std::vector<int> arc;
        admit_batch(s, batch, arc);
        for (size_t k = 0; k < batch.size(); k++) {
            if (arc[k]) done.emplace_back(batch[k], arc[k]);
            else ok.push_back(batch[k]);
        }
        auto t2 = std::chrono::steady_clock::now();
        s->upload_ms += std::chrono::duration<double, std::milli>(t2 - t).count();
        std::vector<size_t> raw;
        if (!ok.empty() && s->main_idle.load()) {
            raw.swap(ok);  // nothing refining: the main loop prepares it in its arena
        } else if (!ok.empty()) {
            std::lock_guard<std::mutex> pl(s->prep_mu);
            const int rc = prepare(s, ok);
            s->prep_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t2).count();
            for (size_t i : ok) done.emplace_back(i, rc);
        }
        if (cap.owns_lock()) cap.unlock();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->up_done.insert(s->up_done.end(), done.begin(), done.end());
            if (!raw.empty()) s->up_raw.push_back(std::move(raw));
        }
        s->cv_done.notify_one();
    }
}

static void stop_uploader(scd_server *s) {
    if (!s->up_thread.joinable()) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->up_stop = true;
    }
    s->cv.notify_all();
    s->up_thread.join();
}

static scd_part part_of(const Req &r) {
    scd_part p;
    p.n = r.in.n;
    p.m = r.in.m;
    p.indices = r.d_ind;
    p.adj = r.d_adj;
    p.tri = r.d_tri;
    p.tri_deg = r.d_td;
    p.clustering_coef = r.in.clustering_coef;
    return p;
}

// Initial labels of the waiting graphs, then a new main arena from the
// still-refining graphs (labels carried over) and the waiting ones.
static int join(scd_server *s) {
    int rc;
    const double t = s->now();
    for (size_t i : s->waiting) s->reqs[i].res.start_ms = t;
    // carry the refining graphs' labels out of the old arena
    if (!s->arena.empty()) {
        const size_t G = s->arena.size();
        std::vector<uint32_t *> c2(G, nullptr), b2(G, nullptr);
        for (size_t g = 0; g < G; g++)
            if (s->active[g]) {
                c2[g] = s->reqs[s->arena[g]].d_cur;
                b2[g] = s->reqs[s->arena[g]].d_best;
            }
        if ((rc = scd_batch_export_labels(s->main, c2.data(), b2.data())))
            return engine_fail(rc, "label carry-over");
    }
    std::vector<size_t> next;
    for (size_t g = 0; g < s->arena.size(); g++)
        if (s->active[g]) next.push_back(s->arena[g]);
    for (size_t i : s->waiting) next.push_back(i);
    const uint32_t G = (uint32_t)next.size();
    std::vector<scd_part> all(G);
    std::vector<const uint32_t *> ac(G), ab(G);
    for (uint32_t g = 0; g < G; g++) {
        const Req &r = s->reqs[next[g]];
        all[g] = part_of(r);
        ac[g] = r.d_cur;
        ab[g] = r.d_best;
    }
    if ((rc = scd_batch_upload_parts(s->main, G, all.data(), ac.data(), ab.data())))
        return engine_fail(rc, "arena assembly");
    s->st.assemble_ms += s->now() - t;
    for (size_t i : s->waiting) {
        s->reqs[i].state = REFINING;
        s->reqs[i].res.joined_arena_graphs = G;
    }
    s->arena.swap(next);
    s->active.assign(G, 1u);
    s->n_active = G;
    s->waiting.clear();
    s->st.joins++;
    s->st.peak_graphs = std::max(s->st.peak_graphs, G);
    return SCD_OK;
}

// One iteration of every refining graph; finished graphs hand back their
// best labels and leave.
static int step(scd_server *s) {
    const size_t G = s->arena.size();
    std::vector<double> w(G, 0.0);
    int rc = scd_batch_step(s->main, s->active.data(), w.data());
    if (rc) return engine_fail(rc, "iteration");
    s->st.iterations++;
    std::vector<uint32_t> store(G, 0u), stop(G, 0u);
    bool any_store = false;
    for (size_t g = 0; g < G; g++) {
        if (!s->active[g]) continue;
        Req &r = s->reqs[s->arena[g]];
        r.res.iterations++;
        if (w[g] > r.best) {
            store[g] = 1;
            any_store = true;
        }
        if (w[g] < r.best * (1 + kThreshold)) {
            stop[g] = 1;
        } else {
            r.best = w[g];
            if (r.res.iterations >= s->o.max_iters) stop[g] = 1;
        }
    }
    if (any_store && (rc = scd_batch_store_best(s->main, store.data())))
        return engine_fail(rc, "store best labels");
    // every graph that stopped hands back its best labels in one download
    std::vector<uint32_t> fin;
    std::vector<uint32_t *> outs;
    for (size_t g = 0; g < G; g++) {
        if (!stop[g]) continue;
        Req &r = s->reqs[s->arena[g]];
        r.labels.resize(r.in.n);
        fin.push_back((uint32_t)g);
        outs.push_back(r.labels.data());
    }
    if (!fin.empty() && (rc = scd_batch_graphs_labels(s->main, (uint32_t)fin.size(), fin.data(), 1, outs.data())))
        return engine_fail(rc, "result labels");
    for (size_t g = 0; g < G; g++) {
        if (!stop[g]) continue;
        Req &r = s->reqs[s->arena[g]];
        r.res.best_wcc = r.best;
        r.res.done_ms = s->now();
        r.state = DONE;
        s->active[g] = 0;
        s->n_active--;
        release(s, r);
    }
    return SCD_OK;
}

// Collects requests the intake thread finished (uploaded and prepared, or
// failed); waits for some when block is set.
static void drain_uploads(scd_server *s, bool block, std::vector<size_t> *raw) {
    std::vector<std::pair<size_t, int>> done;
    std::vector<std::vector<size_t>> rb;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        if (block) s->cv_done.wait(lk, [&] { return !s->up_done.empty() || !s->up_raw.empty(); });
        done.swap(s->up_done);
        rb.swap(s->up_raw);
    }
    for (auto &b : rb)
        for (size_t i : b) {
            s->uploading--;
            s->reqs[i].state = RESIDENT;
            raw->push_back(i);
        }
    for (auto &d : done) {
        Req &r = s->reqs[d.first];
        s->uploading--;
        if (d.second) {
            release(s, r);
            r.state = FAILED;
            r.res.status = d.second;
            r.res.done_ms = s->now();
        } else {
            r.state = RESIDENT;
            s->waiting.push_back(d.first);
        }
    }
}

// A batch handed over unprepared while nothing refines: assembled once, in
// the main arena, given its initial partition and initial WCC there, and
// refined on from there (the same per-graph results as the preparation
// context gives: arena composition never changes a graph's result).
static int prepare_main(scd_server *s, const std::vector<size_t> &batch) {
    const double t = s->now();
    const uint32_t P = (uint32_t)batch.size();
    std::vector<scd_part> parts(P);
    for (uint32_t i = 0; i < P; i++) parts[i] = part_of(s->reqs[batch[i]]);
    std::vector<double> w0(P);
    int rc;
    if ((rc = scd_batch_upload_parts(s->main, P, parts.data(), nullptr, nullptr)) ||
        (rc = scd_batch_initial_partition(s->main, nullptr, nullptr)) ||
        (rc = scd_batch_compute_wcc(s->main, w0.data())))
        return engine_fail(rc, "initial partition of joining graphs");
    set_initial(s, batch, w0);
    const double t2 = s->now();
    for (size_t i : batch) {
        Req &r = s->reqs[i];
        r.res.start_ms = t2;
        r.state = REFINING;
        r.res.joined_arena_graphs = P;
    }
    s->arena = batch;
    s->active.assign(P, 1u);
    s->n_active = P;
    s->st.joins++;
    s->st.peak_graphs = std::max(s->st.peak_graphs, P);
    s->main_prep_ms += t2 - t;
    s->st.busy_ms += t2 - t;
    return SCD_OK;
}

// Unprepared batches that arrived while the arena was busy after all: the
// preparation context takes them here; they wait to join like the others.
static void prepare_late(scd_server *s, const std::vector<size_t> &batch) {
    const double t = s->now();
    int rc;
    {
        std::lock_guard<std::mutex> pl(s->prep_mu);
        rc = prepare(s, batch);
    }
    s->main_prep_ms += s->now() - t;
    for (size_t i : batch) {
        Req &r = s->reqs[i];
        if (rc) {
            release(s, r);
            r.state = FAILED;
            r.res.status = rc;
            r.res.done_ms = s->now();
        } else {
            s->waiting.push_back(i);
        }
    }
}

static int serve_loop(scd_server *s, const std::vector<size_t> &order) {
    size_t qi = 0;
    std::vector<size_t> raw;  // unprepared batches handed over by the intake thread
    while (true) {
        const double now = s->now();
        // admission, first come first served, within the footprint budget;
        // the budget is reserved here, the copy runs on the uploader thread
        while (qi < order.size()) {
            Req &r = s->reqs[order[qi]];
            if (r.in.arrival_ms > now) break;
            const uint64_t fp = scd_server_footprint(r.in.n, r.in.m);
            if (fp > s->o.hbm_budget) {
                r.state = FAILED;
                r.res.status = SCD_ELIMIT;
                r.res.done_ms = now;
                qi++;
                continue;
            }
            if (s->resident + fp > s->o.hbm_budget) break;
            if (s->o.max_graphs && s->n_active + s->waiting.size() + s->uploading + raw.size() >= s->o.max_graphs)
                break;
            r.bytes = fp;
            s->resident += fp;
            s->st.peak_bytes = std::max(s->st.peak_bytes, s->resident);
            r.state = UPLOADING;
            r.res.admit_ms = now;
            s->uploading++;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->up_queue.push_back(order[qi]);
            }
            s->cv.notify_one();
            qi++;
        }
        drain_uploads(s, false, &raw);
        if (!raw.empty()) {
            if (s->n_active == 0) {
                const int rc = prepare_main(s, raw);
                if (rc) return rc;
            } else {
                prepare_late(s, raw);
            }
            raw.clear();
        }
        s->main_idle.store(s->n_active == 0 && s->waiting.empty());
        if (!s->waiting.empty()) {
            const double oldest = s->reqs[s->waiting.front()].in.arrival_ms;
            if (s->n_active == 0 || now - oldest >= s->o.join_wait_ms || 2 * s->n_active <= s->arena.size()) {
                const double t = s->now();
                const int rc = join(s);  // no capture: runs beside the intake thread
                s->st.busy_ms += s->now() - t;
                if (rc) return rc;
            }
        }
        if (s->n_active == 0) {
            if (!s->waiting.empty()) continue;
            if (s->uploading) {
                drain_uploads(s, true, &raw);
                continue;
            }
            if (qi >= order.size()) break;
            const double wait = s->reqs[order[qi]].in.arrival_ms - s->now();
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double, std::milli>(wait));
            continue;
        }
        const double t = s->now();
        int rc;
        {
            // only a step that captures excludes the intake thread; replays
            // of captured iterations run beside its uploads and preparation
            std::unique_lock<std::mutex> cap(s->capture_mu, std::defer_lock);
            if (s->o.graph_mode && scd_next_step_captures(s->main)) cap.lock();
            rc = step(s);
        }
        s->st.busy_ms += s->now() - t;
        s->st.step_ms += s->now() - t;
        if (rc) return rc;
    }
    return SCD_OK;
}

extern "C" int scd_server_run(scd_server *s) {
    if (!s) return sfail(SCD_EINVAL, "scd_server_run: null server");
    std::vector<size_t> order;
    for (size_t i = 0; i < s->reqs.size(); i++)
        if (s->reqs[i].state == QUEUED) order.push_back(i);
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return s->reqs[a].in.arrival_ms < s->reqs[b].in.arrival_ms; });
    s->t0 = std::chrono::steady_clock::now();
    s->up_stop = false;
    s->upload_ms = 0;
    s->prep_ms = 0;
    s->main_prep_ms = 0;
    s->main_idle.store(false);
    s->up_thread = std::thread(intake, s);
    const int rc = serve_loop(s, order);
    stop_uploader(s);
    s->st.admit_ms += s->upload_ms;
    s->st.prep_ms += s->prep_ms + s->main_prep_ms;
    if (rc) {  // requests left mid-way keep their device buffers until destroy
        std::vector<size_t> raw;
        drain_uploads(s, false, &raw);
        return rc;
    }
    return SCD_OK;
}

extern "C" int scd_server_result(scd_server *s, uint64_t id, uint32_t *labels_out, scd_result *out) {
    if (!s || id >= s->reqs.size()) return sfail(SCD_EINVAL, "scd_server_result: unknown request");
    const Req &r = s->reqs[id];
    if (r.state != DONE && r.state != FAILED) return sfail(SCD_ESTATE, "scd_server_result: request not finished");
    if (out) *out = r.res;
    if (labels_out && r.state == DONE) memcpy(labels_out, r.labels.data(), 4ull * r.in.n);
    return SCD_OK;
}

extern "C" int scd_server_get_stats(const scd_server *s, scd_server_stats *out) {
    if (!s || !out) return sfail(SCD_EINVAL, "scd_server_get_stats: null argument");
    *out = s->st;
    return SCD_OK;
}
