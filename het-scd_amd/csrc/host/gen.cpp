// Deterministic graph generators and the convert pipeline.
//
// The generators draw from splitmix64 only (no std::*_distribution), so a
// (generator, parameters, seed) triple names one edge sequence on every
// machine.  scd_convert reproduces the reference convert tool's output bytes
// (src/convert.cpp:466-537) with a different, parallel algorithm:
//   * edges are packed into 64-bit keys and radix sorted (reference: omp_sort
//     of pairs, src/convert.cpp:33-65, :501);
//   * per-edge triangle counts come from a degree-oriented bitmap probe — each
//     edge is owned by its higher-degree end, which marks its own list in a
//     thread-local bitmap and scans the lower-degree end's list — instead of
//     the reference's merge of forward lists (src/convert.cpp:290-319);
//   * the degree relabel is a radix sort of (degree, id) keys
//     (src/convert.cpp:375-392).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <omp.h>
#include <vector>

#include "host_common.hpp"
#include "scd_host.h"

using scd::set_host_error;
using scd::SplitMix64;

static scd_edges *edges_alloc(uint64_t count) {
    scd_edges *e = (scd_edges *)calloc(1, sizeof(scd_edges));
    if (!e) return nullptr;
    e->count = count;
    e->u = (uint32_t *)malloc(sizeof(uint32_t) * (count ? count : 1));
    e->v = (uint32_t *)malloc(sizeof(uint32_t) * (count ? count : 1));
    if (!e->u || !e->v) {
        scd_edges_free(e);
        return nullptr;
    }
    return e;
}

extern "C" void scd_edges_free(scd_edges *e) {
    if (!e) return;
    free(e->u);
    free(e->v);
    free(e);
}

extern "C" int scd_edges_from_arrays(uint64_t count, const uint32_t *u, const uint32_t *v,
                                     scd_edges **out) {
    if (!out || (count && (!u || !v))) {
        set_host_error("scd_edges_from_arrays: null argument");
        return SCD_EINVAL;
    }
    scd_edges *e = edges_alloc(count);
    if (!e) {
        set_host_error("scd_edges_from_arrays: out of host memory");
        return SCD_ENOMEM;
    }
    if (count) {
        memcpy(e->u, u, sizeof(uint32_t) * count);
        memcpy(e->v, v, sizeof(uint32_t) * count);
    }
    *out = e;
    return SCD_OK;
}

extern "C" int scd_edges_write_text(const scd_edges *e, const char *path) {
    if (!e || !path) {
        set_host_error("scd_edges_write_text: null argument");
        return SCD_EINVAL;
    }
    FILE *f = fopen(path, "w");
    if (!f) {
        set_host_error("failed to open %s for writing", path);
        return SCD_EIO;
    }
    std::vector<char> buf(1 << 20);
    setvbuf(f, buf.data(), _IOFBF, buf.size());
    bool ok = true;
    for (uint64_t i = 0; i < e->count && ok; i++) ok = fprintf(f, "%u %u\n", e->u[i], e->v[i]) > 0;
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        set_host_error("an error occured while writing %s", path);
        return SCD_EIO;
    }
    return SCD_OK;
}

// Text edge list with the reference reader's rules (src/convert.cpp:74-136):
// per line, leading white space, then an empty line or '#' is skipped;
// otherwise two decimal numbers (uint32 arithmetic, wrapping) separated by
// white space, the rest of the line ignored.  A line of 1024 or more
// characters is a read error (the reference's getline into char[1024] sets
// failbit after processing its first 1023 characters).  The first problem in
// line order is reported.  Lines are parsed in parallel over chunks cut at
// line ends.
namespace {
struct TextChunk {
    size_t begin = 0, end = 0;  // byte range, whole lines
    uint64_t first_line = 0;    // 1-based number of its first line
    std::vector<uint32_t> u, v;
    uint64_t bad_line = 0;      // first problem in the chunk (0: none)
    int bad_kind = 0;           // 1 corrupt, 2 read error
};

inline bool c_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }
inline bool c_digit(char c) { return c >= '0' && c <= '9'; }

// One line [s, e) (no '\n'): 0 skipped, 1 edge, 2 corrupt.
int parse_line(const char *s, size_t len, uint32_t *a_out, uint32_t *b_out) {
    size_t i = 0;
    auto at = [&](size_t k) -> char { return k < len ? s[k] : '\0'; };
    while (c_space(at(i))) i++;
    if (at(i) == '\0' || at(i) == '#') return 0;
    if (!c_digit(at(i))) return 2;
    uint32_t a = 0;
    while (c_digit(at(i))) a = 10u * a + (uint32_t)(at(i++) - '0');
    while (c_space(at(i))) i++;
    if (!c_digit(at(i))) return 2;
    uint32_t b = 0;
    while (c_digit(at(i))) b = 10u * b + (uint32_t)(at(i++) - '0');
    *a_out = a;
    *b_out = b;
    return 1;
}
}  // namespace

extern "C" int scd_edges_read_text(const char *path, scd_edges **out) {
    if (!path || !out) {
        set_host_error("scd_edges_read_text: null argument");
        return SCD_EINVAL;
    }
    FILE *f = fopen(path, "rb");
    if (!f) {
        set_host_error("an error occured while reading input file (cannot open %s)", path);
        return SCD_EIO;
    }
    std::vector<char> buf;
    {
        std::vector<char> blk(1 << 24);
        size_t got;
        while ((got = fread(blk.data(), 1, blk.size(), f)) > 0) buf.insert(buf.end(), blk.begin(), blk.begin() + got);
        const bool err = ferror(f) != 0;
        fclose(f);
        if (err) {
            set_host_error("an error occured while reading input file %s", path);
            return SCD_EIO;
        }
    }
    const size_t size = buf.size();
    const int T = std::max(1, std::min(omp_get_max_threads(), (int)(size >> 20) + 1));
    std::vector<TextChunk> ch(T);
    for (int t = 0; t < T; t++) {
        size_t b = size * (size_t)t / T;
        if (t > 0) {  // start after a line end
            const char *nl = (const char *)memchr(buf.data() + b - 1, '\n', size - (b - 1));
            b = nl ? (size_t)(nl - buf.data()) + 1 : size;
            b = std::max(b, ch[t - 1].begin);
        }
        ch[t].begin = b;
        if (t > 0) ch[t - 1].end = b;
    }
    ch[T - 1].end = size;
    // line numbers: lines that start in earlier chunks
    std::vector<uint64_t> nlines(T, 0);
#pragma omp parallel for num_threads(T)
    for (int t = 0; t < T; t++) {
        uint64_t c = 0;
        for (size_t k = ch[t].begin; k < ch[t].end; k++) c += buf[k] == '\n';
        nlines[t] = c;
    }
    uint64_t line = 1;
    for (int t = 0; t < T; t++) {
        ch[t].first_line = line;
        line += nlines[t];
    }
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (int t = 0; t < T; t++) {
        TextChunk &c = ch[t];
        uint64_t ln = c.first_line;
        size_t k = c.begin;
        while (k < c.end) {
            const char *nl = (const char *)memchr(buf.data() + k, '\n', c.end - k);
            const size_t e = nl ? (size_t)(nl - buf.data()) : c.end;
            const size_t len = e - k;
            uint32_t a, b;
            const int r = parse_line(buf.data() + k, std::min<size_t>(len, 1023), &a, &b);
            if (r == 2) {
                c.bad_line = ln;
                c.bad_kind = 1;
                break;
            }
            if (r == 1) {
                c.u.push_back(a);
                c.v.push_back(b);
            }
            if (len >= 1024) {
                c.bad_line = ln;
                c.bad_kind = 2;
                break;
            }
            k = e + 1;
            ln++;
        }
    }
    uint64_t total = 0;
    for (int t = 0; t < T; t++) {
        if (ch[t].bad_kind == 1) {
            set_host_error("line %llu is corrupt", (unsigned long long)ch[t].bad_line);
            return SCD_EINVAL;
        }
        if (ch[t].bad_kind == 2) {
            set_host_error("an error occured while reading input file (line %llu has 1024 or more characters)",
                           (unsigned long long)ch[t].bad_line);
            return SCD_EIO;
        }
        total += ch[t].u.size();
    }
    scd_edges *e = edges_alloc(total);
    if (!e) {
        set_host_error("scd_edges_read_text: out of host memory");
        return SCD_ENOMEM;
    }
    uint64_t o = 0;
    for (int t = 0; t < T; t++) {
        if (!ch[t].u.empty()) {
            memcpy(e->u + o, ch[t].u.data(), 4 * ch[t].u.size());
            memcpy(e->v + o, ch[t].v.data(), 4 * ch[t].v.size());
        }
        o += ch[t].u.size();
    }
    *out = e;
    return SCD_OK;
}

// ---------------------------------------------------------------------------
// Planted partition
// ---------------------------------------------------------------------------

// Number of non-edges before the next edge of a Bernoulli(p) sequence, drawn
// by inversion (a geometric variate), so a block of N candidate pairs costs
// O(#edges) draws rather than N.  Returns UINT64_MAX for "past the end".
static inline uint64_t geometric_skip(SplitMix64 &rng, double p, double log1mp) {
    if (p >= 1.0) return 0;
    const double u = 1.0 - rng.uniform();  // (0, 1]
    const double s = std::floor(std::log(u) / log1mp);
    if (!(s < 1.8e19)) return UINT64_MAX;
    return (uint64_t)s;
}

static int planted_partition(uint32_t k, const uint32_t *sizes, double p_in, double p_out,
                             uint64_t seed, scd_edges **out) {
    std::vector<uint64_t> off(k + 1, 0);
    for (uint32_t a = 0; a < k; a++) off[a + 1] = off[a] + sizes[a];
    if (off[k] > 0xFFFFFFFEull) {
        set_host_error("planted partition: %llu vertices exceed uint32 ids",
                       (unsigned long long)off[k]);
        return SCD_ELIMIT;
    }
    std::vector<uint32_t> us, vs;
    SplitMix64 rng(seed);
    for (uint32_t a = 0; a < k; a++) {
        for (uint32_t b = a; b < k; b++) {
            const double p = (a == b) ? p_in : p_out;
            if (!(p > 0.0)) continue;
            const uint64_t sa = sizes[a], sb = sizes[b];
            const uint64_t npairs = (a == b) ? sa * (sa - (sa > 0)) / 2 : sa * sb;
            if (npairs == 0) continue;
            const double log1mp = std::log1p(-p);
            // Intra-community pairs (i < j) are enumerated row by row; the
            // row of the current position advances monotonically.
            uint64_t row = 0, row_start = 0;
            uint64_t pos = 0;  // index of the last pair taken
            bool first = true;
            while (true) {
                const uint64_t skip = geometric_skip(rng, p, log1mp);
                if (skip == UINT64_MAX) break;
                // next candidate = skip (first draw) or pos + 1 + skip
                const uint64_t room = first ? npairs : npairs - pos - 1;
                if (skip >= room) break;
                pos = first ? skip : pos + 1 + skip;
                first = false;
                uint64_t i, j;
                if (a == b) {
                    while (pos >= row_start + (sa - 1 - row)) {
                        row_start += sa - 1 - row;
                        row++;
                    }
                    i = row;
                    j = row + 1 + (pos - row_start);
                } else {
                    i = pos / sb;
                    j = pos % sb;
                }
                us.push_back((uint32_t)(off[a] + i));
                vs.push_back((uint32_t)(off[b] + j));
            }
        }
    }
    scd_edges *e = edges_alloc(us.size());
    if (!e) {
        set_host_error("planted partition: out of host memory");
        return SCD_ENOMEM;
    }
    if (!us.empty()) {
        memcpy(e->u, us.data(), sizeof(uint32_t) * us.size());
        memcpy(e->v, vs.data(), sizeof(uint32_t) * vs.size());
    }
    *out = e;
    return SCD_OK;
}

extern "C" int scd_gen_planted_partition(uint32_t k, const uint32_t *sizes, double p_in,
                                         double p_out, uint64_t seed, scd_edges **out) {
    if (!out || (k && !sizes) || !(p_in >= 0.0 && p_in <= 1.0) || !(p_out >= 0.0 && p_out <= 1.0)) {
        set_host_error("scd_gen_planted_partition: invalid argument");
        return SCD_EINVAL;
    }
    return planted_partition(k, sizes, p_in, p_out, seed, out);
}

extern "C" int scd_gen_planted_random(uint64_t seed, uint32_t n_min, uint32_t n_max,
                                      uint32_t k_min, uint32_t k_max, double p_in, double p_out,
                                      uint32_t *n_out, uint32_t *k_out, scd_edges **out) {
    if (!out || n_min == 0 || n_min > n_max || k_min == 0 || k_min > k_max ||
        !(p_in >= 0.0 && p_in <= 1.0) || !(p_out >= 0.0 && p_out <= 1.0)) {
        set_host_error("scd_gen_planted_random: invalid argument");
        return SCD_EINVAL;
    }
    // The parameter stream is separate from the edge stream so that a graph's
    // shape does not depend on how many edges an earlier draw produced.
    SplitMix64 prng(scd::mix64(seed ^ 0x5bd1e9955bd1e995ull));
    const double lo = std::log((double)n_min), hi = std::log((double)n_max);
    uint64_t n = (uint64_t)std::floor(std::exp(lo + prng.uniform() * (hi - lo)));
    n = std::min<uint64_t>(std::max<uint64_t>(n, n_min), n_max);
    uint64_t k = k_min + (uint64_t)std::floor(prng.uniform() * (double)(k_max - k_min + 1));
    k = std::min<uint64_t>(std::min<uint64_t>(k, k_max), n);
    // Symmetric Dirichlet(1) proportions are normalised Exp(1) variates; each
    // community gets one vertex, the rest go by largest remainder.
    std::vector<double> w(k);
    double sum = 0.0;
    for (uint64_t i = 0; i < k; i++) {
        w[i] = -std::log(1.0 - prng.uniform());
        sum += w[i];
    }
    const uint64_t rest = n - k;
    std::vector<uint32_t> sizes(k);
    std::vector<std::pair<double, uint64_t>> frac(k);
    uint64_t given = 0;
    for (uint64_t i = 0; i < k; i++) {
        const double share = (double)rest * (w[i] / sum);
        const uint64_t fl = (uint64_t)std::floor(share);
        sizes[i] = (uint32_t)(1 + fl);
        given += fl;
        frac[i] = std::make_pair(share - (double)fl, i);
    }
    std::sort(frac.begin(), frac.end(), [](const std::pair<double, uint64_t> &x,
                                           const std::pair<double, uint64_t> &y) {
        return x.first > y.first || (x.first == y.first && x.second < y.second);
    });
    for (uint64_t i = 0; given < rest; i++, given++) sizes[frac[i % k].second]++;
    if (n_out) *n_out = (uint32_t)n;
    if (k_out) *k_out = (uint32_t)k;
    return planted_partition((uint32_t)k, sizes.data(), p_in, p_out, scd::mix64(seed), out);
}

// ---------------------------------------------------------------------------
// R-MAT (Chakrabarti, Zhan, Faloutsos, SDM 2004): each edge descends `scale`
// levels of the adjacency matrix, choosing a quadrant with probabilities
// (a, b, c, 1-a-b-c).  Edge i uses its own stream, so the list does not
// depend on the thread count.
// ---------------------------------------------------------------------------
extern "C" int scd_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                            uint64_t seed, scd_edges **out) {
    if (!out || scale == 0 || scale > 31 || edge_factor == 0 || !(a >= 0 && b >= 0 && c >= 0) ||
        !(a + b + c <= 1.0)) {
        set_host_error("scd_gen_rmat: invalid argument");
        return SCD_EINVAL;
    }
    const uint64_t count = (uint64_t)edge_factor << scale;
    if (count > (1ull << 34)) {
        set_host_error("scd_gen_rmat: %llu edges is too many", (unsigned long long)count);
        return SCD_ELIMIT;
    }
    scd_edges *e = edges_alloc(count);
    if (!e) {
        set_host_error("scd_gen_rmat: out of host memory");
        return SCD_ENOMEM;
    }
    const double ab = a + b, abc = a + b + c;
    const uint64_t base = scd::mix64(seed * 0xD1B54A32D192ED03ull + 0x9E3779B97F4A7C15ull);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)count; i++) {
        SplitMix64 rng(scd::mix64(base ^ ((uint64_t)i * 0xA24BAED4963EE407ull)));
        uint32_t src = 0, dst = 0;
        for (uint32_t l = 0; l < scale; l++) {
            const double r = rng.uniform();
            const uint32_t bit = 1u << (scale - 1 - l);
            if (r < a) {
            } else if (r < ab) {
                dst |= bit;
            } else if (r < abc) {
                src |= bit;
            } else {
                src |= bit;
                dst |= bit;
            }
        }
        e->u[i] = src;
        e->v[i] = dst;
    }
    *out = e;
    return SCD_OK;
}

// ---------------------------------------------------------------------------
// Convert
// ---------------------------------------------------------------------------

// Stable LSD radix sort of 64-bit keys with 11-bit digits; per-thread
// histograms keep each pass a single read and a single scattered write.
static void radix_sort_u64(std::vector<uint64_t> &keys, int key_bits) {
    const size_t n = keys.size();
    if (n < 2 || key_bits <= 0) return;
    std::vector<uint64_t> tmp(n);
    // inside an enclosing parallel region (one graph per thread) the nested
    // regions below run with a single thread: size the split for that
    const int nthreads = omp_in_parallel() ? 1 : std::max(1, omp_get_max_threads());
    const int B = 11, R = 1 << B;
    std::vector<size_t> hist((size_t)nthreads * R);
    for (int shift = 0; shift < key_bits; shift += B) {
        std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(nthreads)
        {
            const int t = omp_get_thread_num();
            const size_t lo = n * t / nthreads, hi = n * (t + 1) / nthreads;
            size_t *h = &hist[(size_t)t * R];
            for (size_t i = lo; i < hi; i++) h[(keys[i] >> shift) & (R - 1)]++;
        }
        size_t run = 0;
        for (int d = 0; d < R; d++)
            for (int t = 0; t < nthreads; t++) {
                const size_t c = hist[(size_t)t * R + d];
                hist[(size_t)t * R + d] = run;
                run += c;
            }
#pragma omp parallel num_threads(nthreads)
        {
            const int t = omp_get_thread_num();
            const size_t lo = n * t / nthreads, hi = n * (t + 1) / nthreads;
            size_t *h = &hist[(size_t)t * R];
            for (size_t i = lo; i < hi; i++) tmp[h[(keys[i] >> shift) & (R - 1)]++] = keys[i];
        }
        keys.swap(tmp);
    }
}

static int bits_for(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b)) b++;
    return b;
}

extern "C" int scd_convert(const scd_edges *e, scd_graph **out) {
    if (!e || !out || (e->count && (!e->u || !e->v))) {
        set_host_error("scd_convert: null argument");
        return SCD_EINVAL;
    }
    const uint64_t count = e->count;

    // 1. Reindex by first appearance of (min, max) endpoints
    //    (src/convert.cpp:120-123, :141-170).
    uint32_t max_id = 0;
    for (uint64_t i = 0; i < count; i++) max_id = std::max(max_id, std::max(e->u[i], e->v[i]));
    std::vector<uint32_t> remap(count ? (size_t)max_id + 1 : 0, UINT32_MAX);
    std::vector<uint64_t> keys(count);
    uint32_t nv = 0;
    for (uint64_t i = 0; i < count; i++) {
        const uint32_t a = std::min(e->u[i], e->v[i]), b = std::max(e->u[i], e->v[i]);
        if (remap[a] == UINT32_MAX) remap[a] = nv++;
        if (remap[b] == UINT32_MAX) remap[b] = nv++;
        const uint32_t x = std::min(remap[a], remap[b]), y = std::max(remap[a], remap[b]);
        keys[i] = ((uint64_t)x << 32) | y;
    }
    std::vector<uint32_t>().swap(remap);

    // 2. Sort, drop loops and duplicates (src/convert.cpp:252-268, :501).
    const int idbits = bits_for(nv);
    if (idbits < 32) {
        for (auto &k : keys) k = ((k >> 32) << idbits) | (k & 0xFFFFFFFFull);
        radix_sort_u64(keys, 2 * idbits);
        const uint64_t lowmask = (idbits ? ((1ull << idbits) - 1) : 0);
        for (auto &k : keys) k = ((k >> idbits) << 32) | (k & lowmask);
    } else {
        radix_sort_u64(keys, 64);
    }
    size_t mu = 0;
    for (size_t i = 0; i < keys.size(); i++) {
        const uint64_t k = keys[i];
        if ((uint32_t)(k >> 32) != (uint32_t)k && (mu == 0 || keys[mu - 1] != k)) keys[mu++] = k;
    }
    keys.resize(mu);

    // 3. Symmetric CSR of the simple graph; lists come out sorted because the
    //    keys are sorted by (lo, hi): all lower neighbours of a vertex arrive
    //    in increasing order before its higher neighbours.
    std::vector<uint64_t> off((size_t)nv + 1, 0), fill_pos;
    for (uint64_t k : keys) {
        off[(k >> 32) + 1]++;
        off[(uint32_t)k + 1]++;
    }
    for (size_t v = 0; v < nv; v++) off[v + 1] += off[v];
    std::vector<uint32_t> adj(2 * mu);
    fill_pos.assign(off.begin(), off.end() - (nv ? 1 : 0));
    for (uint64_t k : keys) adj[fill_pos[(uint32_t)k]++] = (uint32_t)(k >> 32);
    for (uint64_t k : keys) adj[fill_pos[k >> 32]++] = (uint32_t)k;
    std::vector<uint64_t>().swap(fill_pos);
    std::vector<uint64_t>().swap(keys);

    // 4. Triangles per edge = |N(x) ∩ N(y)|, stored on both directed slots.
    //    The owner of edge {x, y} is the end with the larger (degree, id); it
    //    marks N(owner) in a bitmap and scans the other end's list, so the
    //    total work is bounded by sum over edges of min degree.
    std::vector<uint32_t> tslot(2 * mu, 0);
    auto deg = [&](uint32_t v) { return off[v + 1] - off[v]; };
#pragma omp parallel
    {
        std::vector<uint64_t> bits(((size_t)nv + 63) / 64 + 1, 0);
#pragma omp for schedule(dynamic, 64)
        for (int64_t xi = 0; xi < (int64_t)nv; xi++) {
            const uint32_t x = (uint32_t)xi;
            const uint64_t dx = deg(x);
            for (uint64_t s = off[x]; s < off[x + 1]; s++) bits[adj[s] >> 6] |= 1ull << (adj[s] & 63);
            for (uint64_t s = off[x]; s < off[x + 1]; s++) {
                const uint32_t y = adj[s];
                const uint64_t dy = deg(y);
                if (dy > dx || (dy == dx && y > x)) continue;  // y owns this edge
                uint32_t cnt = 0;
                for (uint64_t t = off[y]; t < off[y + 1]; t++)
                    cnt += (uint32_t)((bits[adj[t] >> 6] >> (adj[t] & 63)) & 1ull);
                tslot[s] = cnt;
                const uint32_t *lb = std::lower_bound(&adj[off[y]], &adj[0] + off[y + 1], x);
                tslot[(size_t)(lb - &adj[0])] = cnt;
            }
            for (uint64_t s = off[x]; s < off[x + 1]; s++) bits[adj[s] >> 6] = 0;
        }
    }

    // 5. tri / tri_deg (src/convert.cpp:321-340), wrapping like int atomics.
    std::vector<int32_t> tri(nv), tdeg(nv);
#pragma omp parallel for schedule(static)
    for (int64_t xi = 0; xi < (int64_t)nv; xi++) {
        uint32_t t = 0, d = 0;
        for (uint64_t s = off[xi]; s < off[xi + 1]; s++) {
            t += tslot[s];
            d += tslot[s] > 0;
        }
        tri[xi] = (int32_t)t;
        tdeg[xi] = (int32_t)d;
    }

    // 6. Relabel by (remaining degree, id) descending; isolated vertices go
    //    last and are dropped (src/convert.cpp:354-414).
    std::vector<uint64_t> order(nv);
    for (uint32_t v = 0; v < nv; v++)
        order[v] = ((uint64_t)(uint32_t)~(uint32_t)tdeg[v] << 32) | (uint32_t)~v;
    radix_sort_u64(order, 64);
    std::vector<uint32_t> newid(nv);
    uint32_t n_new = 0;
    for (uint32_t r = 0; r < nv; r++) {
        const uint32_t v = ~(uint32_t)order[r];
        newid[v] = r;
        if (tdeg[v] > 0) n_new = r + 1;
    }
    uint64_t m2 = 0;
    for (uint32_t v = 0; v < nv; v++) m2 += (uint32_t)tdeg[v];
    if (m2 / 2 >= (1ull << 31)) {
        set_host_error("scd_convert: %llu edges exceed the binary format",
                       (unsigned long long)(m2 / 2));
        return SCD_ELIMIT;
    }
    scd_graph *g = nullptr;
    {
        std::vector<uint32_t> idx((size_t)n_new + 1, 0);
        for (uint32_t r = 0; r < n_new; r++) idx[r + 1] = idx[r] + (uint32_t)tdeg[~(uint32_t)order[r]];
        std::vector<int32_t> ntri(n_new), ntdeg(n_new);
        std::vector<uint32_t> nadj(m2);
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t r = 0; r < (int64_t)n_new; r++) {
            const uint32_t v = ~(uint32_t)order[r];
            ntri[r] = tri[v];
            ntdeg[r] = tdeg[v];
            uint32_t *dst = &nadj[idx[r]];
            uint32_t k = 0;
            for (uint64_t s = off[v]; s < off[v + 1]; s++)
                if (tslot[s] > 0) dst[k++] = newid[adj[s]];
            std::sort(dst, dst + k);
        }
        const int rc = scd_graph_from_arrays(n_new, (uint32_t)(m2 / 2), idx.data(), nadj.data(),
                                             ntri.data(), ntdeg.data(), &g);
        if (rc != SCD_OK) return rc;
    }
    *out = g;
    return SCD_OK;
}

// ---------------------------------------------------------------------------
// Arenas: several graphs as one disjoint union (vertex ids offset per graph)
// ---------------------------------------------------------------------------
static int concat_graphs(const std::vector<const scd_graph *> &gs, scd_graph **out, uint64_t *vtx_off) {
    uint64_t n = 0, m = 0;
    for (size_t i = 0; i < gs.size(); i++) {
        vtx_off[i] = n;
        n += gs[i]->n;
        m += gs[i]->m;
    }
    vtx_off[gs.size()] = n;
    if (n >= 0xFFFFFFFFull || m >= (1ull << 31)) {
        set_host_error("arena of %llu vertices and %llu edges exceeds the u32 format",
                       (unsigned long long)n, (unsigned long long)m);
        return SCD_ELIMIT;
    }
    scd_graph *g = (scd_graph *)calloc(1, sizeof(scd_graph));
    if (!g) return SCD_ENOMEM;
    g->n = (uint32_t)n;
    g->m = (uint32_t)m;
    g->indices = (uint32_t *)malloc(4 * (n + 1));
    g->adj = (uint32_t *)malloc(4 * (2 * m + 1));
    g->tri = (int32_t *)malloc(4 * (n + 1));
    g->tri_deg = (int32_t *)malloc(4 * (n + 1));
    if (!g->indices || !g->adj || !g->tri || !g->tri_deg) {
        scd_graph_free(g);
        set_host_error("arena: out of host memory");
        return SCD_ENOMEM;
    }
    std::vector<uint64_t> eoff(gs.size() + 1, 0);
    for (size_t i = 0; i < gs.size(); i++) eoff[i + 1] = eoff[i] + 2ull * gs[i]->m;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t ii = 0; ii < (int64_t)gs.size(); ii++) {
        const scd_graph *s = gs[ii];
        const uint32_t vo = (uint32_t)vtx_off[ii], eo = (uint32_t)eoff[ii];
        for (uint32_t v = 0; v < s->n; v++) {
            g->indices[vo + v] = eo + s->indices[v];
            g->tri[vo + v] = s->tri[v];
            g->tri_deg[vo + v] = s->tri_deg[v];
        }
        for (uint64_t k = 0; k < 2ull * s->m; k++) g->adj[eo + k] = vo + s->adj[k];
    }
    g->indices[n] = (uint32_t)(2 * m);
    *out = g;
    return SCD_OK;
}

extern "C" int scd_graph_concat(uint32_t count, const scd_graph *const *graphs, scd_graph **out,
                                uint64_t *vtx_off) {
    if (!out || !vtx_off || (count && !graphs)) {
        set_host_error("scd_graph_concat: null argument");
        return SCD_EINVAL;
    }
    std::vector<const scd_graph *> gs(graphs, graphs + count);
    return concat_graphs(gs, out, vtx_off);
}

extern "C" int scd_gen_planted_batch(uint64_t seed_base, uint32_t count, uint32_t n_min, uint32_t n_max,
                                     uint32_t k_min, uint32_t k_max, double p_in, double p_out,
                                     scd_graph **out, uint64_t *vtx_off, double *cc, uint64_t *seeds) {
    if (!out || !vtx_off || count == 0) {
        set_host_error("scd_gen_planted_batch: invalid argument");
        return SCD_EINVAL;
    }
    std::vector<scd_graph *> gs(count, nullptr);
    std::vector<int> rcs(count, SCD_OK);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t ii = 0; ii < (int64_t)count; ii++) {
        // a graph that the convert pipeline empties (no triangle) is drawn
        // again with the next seed of its own sequence
        for (uint64_t attempt = 0; attempt < 64; attempt++) {
            const uint64_t seed = seed_base + (uint64_t)ii + attempt * (uint64_t)count;
            scd_edges *e = nullptr;
            int rc = scd_gen_planted_random(seed, n_min, n_max, k_min, k_max, p_in, p_out, nullptr, nullptr, &e);
            scd_graph *g = nullptr;
            if (rc == SCD_OK) rc = scd_convert(e, &g);
            scd_edges_free(e);
            if (rc != SCD_OK) {
                rcs[ii] = rc;
                break;
            }
            if (g->n > 0) {
                gs[ii] = g;
                if (seeds) seeds[ii] = seed;
                break;
            }
            scd_graph_free(g);
        }
        if (!gs[ii] && rcs[ii] == SCD_OK) rcs[ii] = SCD_EINVAL;
    }
    int rc = SCD_OK;
    for (uint32_t i = 0; i < count; i++)
        if (rcs[i] != SCD_OK) rc = rcs[i];
    if (rc == SCD_OK) {
        if (cc)
            for (uint32_t i = 0; i < count; i++) cc[i] = scd_graph_clustering_coef(gs[i]);
        rc = concat_graphs(std::vector<const scd_graph *>(gs.begin(), gs.end()), out, vtx_off);
    } else {
        set_host_error("scd_gen_planted_batch: generating a graph failed (%d)", rc);
    }
    for (auto g : gs) scd_graph_free(g);
    return rc;
}
