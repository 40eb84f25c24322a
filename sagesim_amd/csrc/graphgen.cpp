// Seeded graph generators and CSR construction for the per-tick agent step.
//
// This is the ingestion step before the path: the reference builds neighbour
// lists one `connect_agents` call at a time (space.py:107-125) and turns them
// into CSR with per-element Python loops (internal_utils.py:111-164,
// model.py:31-110).  Here every generator writes an edge list with a
// deterministic, seeded stream and one CSR builder turns it into
// int64 row_ptr / int32 col_idx with rows sorted ascending and duplicate
// edges and self loops removed.  The same binary produces the same graph on
// any host, whatever the thread count.
//
// Host-only C++ (g++), exported with a C ABI; no GPU is needed.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>
#include <omp.h>

namespace {

// splitmix64: seeds streams and hashes (block, seed) pairs.
inline uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// xoshiro256** (Blackman & Vigna), one instance per deterministic stream.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed, uint64_t stream = 0) {
        uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull);
        for (auto &w : s) w = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform double in [0, 1) with 53 random bits
    inline double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    // uniform integer in [0, n) by 128-bit multiply (Lemire), n > 0
    inline uint64_t below(uint64_t n) {
        return (uint64_t)(((unsigned __int128)next() * (unsigned __int128)n) >> 64);
    }
};

// Geometric skip for Bernoulli(p) trials: number of failures before the next
// success (Batagelj & Brandes 2005).  p in (0, 1).
inline uint64_t geo_skip(Rng &r, double log1mp) {
    double u = r.uniform();
    return (uint64_t)std::floor(std::log1p(-u) / log1mp);
}

// Edge list as interleaved int32 endpoint pairs (u0, v0, u1, v1, ...).
using Pairs = std::vector<int32_t>;

}  // namespace

struct sg_graph {
    int64_t n = 0;             // rows
    int64_t n_cols = 0;        // column id space (>= n when ghost columns exist)
    std::vector<int64_t> row_ptr;
    std::vector<int32_t> col_idx;
    std::vector<int64_t> extra; // generator-specific (SBM: global index of each ghost column)
};

static thread_local std::string g_err;

// Build an undirected (both directions) or directed CSR from ne endpoint
// pairs: rows sorted ascending, self loops and duplicates dropped.  Rows are
// filled in parallel with atomic cursors; the fill order inside a row depends
// on thread timing, but every row is sorted afterwards, so the result does
// not.  Few large temporaries: first-touch of fresh pages is the main cost on
// hosts where page faults are expensive.
static sg_graph *csr_from_pairs(int64_t n_rows, int64_t n_cols, const int32_t *pr, int64_t ne, bool directed) {
    auto *g = new sg_graph();
    g->n = n_rows;
    g->n_cols = n_cols;
    std::vector<int64_t> &rp = g->row_ptr;
    rp.assign(n_rows + 1, 0);
    int64_t *cnt = rp.data() + 1;
    // one thread (a builder of a pool): plain increments; a locked add on a
    // random cache line costs about as much as the miss itself
    const bool serial = omp_get_max_threads() == 1;
    if (serial) {
        for (int64_t e = 0; e < ne; ++e) {
            const int32_t u = pr[2 * e], v = pr[2 * e + 1];
            if (u == v) continue;
            if (u < n_rows) ++cnt[u];
            if (!directed && v < n_rows) ++cnt[v];
        }
    } else {
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < ne; ++e) {
            const int32_t u = pr[2 * e], v = pr[2 * e + 1];
            if (u == v) continue;
            if (u < n_rows) __atomic_fetch_add(&cnt[u], 1, __ATOMIC_RELAXED);
            if (!directed && v < n_rows) __atomic_fetch_add(&cnt[v], 1, __ATOMIC_RELAXED);
        }
    }
    for (int64_t i = 0; i < n_rows; ++i) rp[i + 1] += rp[i];
    std::vector<int64_t> cur(rp.begin(), rp.end() - 1);
    std::vector<int32_t> &col = g->col_idx;
    col.resize(rp[n_rows]);
    if (serial) {
        for (int64_t e = 0; e < ne; ++e) {
            const int32_t u = pr[2 * e], v = pr[2 * e + 1];
            if (u == v) continue;
            if (u < n_rows) col[cur[u]++] = v;
            if (!directed && v < n_rows) col[cur[v]++] = u;
        }
    } else {
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < ne; ++e) {
            const int32_t u = pr[2 * e], v = pr[2 * e + 1];
            if (u == v) continue;
            if (u < n_rows) col[__atomic_fetch_add(&cur[u], 1, __ATOMIC_RELAXED)] = v;
            if (!directed && v < n_rows) col[__atomic_fetch_add(&cur[v], 1, __ATOMIC_RELAXED)] = u;
        }
    }
    // sort + unique every row (cur[i] := unique length), then compact left in place
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < n_rows; ++i) {
        auto b = col.begin() + rp[i], e = col.begin() + rp[i + 1];
        std::sort(b, e);
        cur[i] = std::unique(b, e) - b;
    }
    int64_t w = 0;
    for (int64_t i = 0; i < n_rows; ++i) {
        const int64_t src = rp[i], len = cur[i];
        if (src != w) std::memmove(col.data() + w, col.data() + src, sizeof(int32_t) * (size_t)len);
        rp[i] = w;
        w += len;
    }
    rp[n_rows] = w;
    col.resize(w);
    return g;
}

// Concatenate per-stream edge lists in stream order.
static Pairs concat(std::vector<Pairs> &parts) {
    size_t tot = 0;
    for (auto &p : parts) tot += p.size();
    Pairs all;
    all.reserve(tot);
    for (auto &p : parts) {
        all.insert(all.end(), p.begin(), p.end());
        Pairs().swap(p);
    }
    return all;
}

extern "C" {

const char *sg_last_error(void) { return g_err.c_str(); }

// OpenMP threads of the calling thread's later generator calls (lets a pool
// of builder threads avoid oversubscribing the host).
void sg_set_num_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

int64_t sg_n(const sg_graph *g) { return g->n; }
int64_t sg_n_cols(const sg_graph *g) { return g->n_cols; }
int64_t sg_nnz(const sg_graph *g) { return (int64_t)g->col_idx.size(); }
int64_t sg_n_extra(const sg_graph *g) { return (int64_t)g->extra.size(); }
const int64_t *sg_row_ptr(const sg_graph *g) { return g->row_ptr.data(); }
const int32_t *sg_col_idx(const sg_graph *g) { return g->col_idx.data(); }
const int64_t *sg_extra(const sg_graph *g) { return g->extra.data(); }
void sg_free(sg_graph *g) { delete g; }

// CSR from explicit edges (NetworkSpace bulk ingestion).
sg_graph *sg_from_edges(int64_t n, const int64_t *u, const int64_t *v, int64_t ne, int directed) {
    if (n < 0 || n > INT32_MAX) { g_err = "sg_from_edges: n exceeds int32 column range"; return nullptr; }
    Pairs pr((size_t)ne * 2);
    for (int64_t e = 0; e < ne; ++e) {
        if (u[e] < 0 || u[e] >= n || v[e] < 0 || v[e] >= n) {
            g_err = "sg_from_edges: endpoint out of range";
            return nullptr;
        }
        pr[2 * e] = (int32_t)u[e];
        pr[2 * e + 1] = (int32_t)v[e];
    }
    return csr_from_pairs(n, n, pr.data(), ne, directed != 0);
}

// Erdos-Renyi G(n, p), undirected.  Rows are cut into blocks of 4096 with one
// stream each, so the edge set does not depend on the thread count.
sg_graph *sg_er(int64_t n, double p, uint64_t seed) {
    if (n < 0 || n > INT32_MAX || !(p >= 0.0 && p <= 1.0)) { g_err = "sg_er: bad n or p"; return nullptr; }
    const int64_t B = 4096;
    int64_t nb = (n + B - 1) / B;
    std::vector<Pairs> parts(nb);
    if (p > 0.0) {
        double l1p = std::log1p(-std::min(p, 1.0 - 1e-16));
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t b = 0; b < nb; ++b) {
            Rng r(seed, (uint64_t)b + 1);
            auto &pl = parts[b];
            for (int64_t i = b * B; i < std::min(n, (b + 1) * B); ++i) {
                // candidates j in (i, n)
                int64_t j = i + 1;
                while (true) {
                    j += (p >= 1.0) ? 0 : (int64_t)geo_skip(r, l1p);
                    if (j >= n) break;
                    pl.push_back((int32_t)i);
                    pl.push_back((int32_t)j);
                    ++j;
                }
            }
        }
    }
    Pairs all = concat(parts);
    return csr_from_pairs(n, n, all.data(), (int64_t)all.size() / 2, false);
}

// Barabasi-Albert preferential attachment: start from a star (node 0 joined
// to 1..m); each later node v draws m distinct targets, each with
// probability proportional to degree, from the list that holds every edge
// endpoint once; duplicates are redrawn.  That endpoint list, read in pairs,
// is also the edge list.
sg_graph *sg_ba(int64_t n, int32_t m, uint64_t seed) {
    if (m < 1 || n < m + 1 || n > INT32_MAX) { g_err = "sg_ba: need 1 <= m < n <= INT32_MAX"; return nullptr; }
    Rng r(seed, 0xBA);
    const int64_t ne = (int64_t)m + (n - m - 1) * (int64_t)m;
    Pairs rep((size_t)ne * 2);
    int64_t len = 0;
    for (int32_t i = 1; i <= m; ++i) { rep[len++] = 0; rep[len++] = i; }
    int32_t tg[64];
    std::vector<int32_t> tgv;
    int32_t *t = m <= 64 ? tg : (tgv.resize(m), tgv.data());
    // The draws come from a window of the stream's next outputs, taken in the
    // stream's order (the graph is that of drawing one at a time).  Node
    // v + D will take, barring redraws, the outputs D * m further on, and its
    // list length is 2 m D more than now: its endpoints are prefetched D
    // nodes ahead, which hides the cache misses of the random reads.
    const int D = 4;
    uint64_t qn = 1;
    while (qn < (uint64_t)(D + 2) * (uint64_t)m + 1) qn <<= 1;
    const uint64_t qm = qn - 1;
    std::vector<uint64_t> q(qn);
    uint64_t qh = 0, qt = 0;                   // outputs consumed / generated
    auto pick = [](uint64_t x, uint64_t k) { return (uint64_t)(((unsigned __int128)x * (unsigned __int128)k) >> 64); };
    for (int64_t v = m + 1; v < n; ++v) {
        while (qt < qh + (uint64_t)(D + 1) * (uint64_t)m) q[qt++ & qm] = r.next();
        if (v + D < n) {
            const uint64_t len_d = (uint64_t)len + 2ull * (uint64_t)m * D;
            for (int32_t j = 0; j < m; ++j)
                __builtin_prefetch(&rep[pick(q[(qh + (uint64_t)D * m + j) & qm], len_d)]);
        }
        int cnt = 0;
        while (cnt < m) {
            if (qh == qt) q[qt++ & qm] = r.next();
            const int32_t c = rep[pick(q[qh++ & qm], (uint64_t)len)];
            bool dup = false;
            for (int k = 0; k < cnt; ++k) dup |= t[k] == c;
            if (!dup) t[cnt++] = c;
        }
        for (int k = 0; k < m; ++k) { rep[len++] = (int32_t)v; rep[len++] = t[k]; }
    }
    return csr_from_pairs(n, n, rep.data(), ne, false);
}

// Watts-Strogatz: ring lattice with k/2 neighbours on each side; each lattice
// edge (u, u+j) is rewired with probability p to (u, w), w uniform over the
// other nodes.  Edges that coincide after rewiring are merged.
sg_graph *sg_ws(int64_t n, int32_t k, double p, uint64_t seed) {
    if (k < 2 || (k & 1) || n <= k || n > INT32_MAX) { g_err = "sg_ws: need even k >= 2 and n > k"; return nullptr; }
    const int32_t h = k / 2;
    Pairs pr((size_t)n * h * 2);
    const int64_t B = 4096;
    int64_t nb = (n + B - 1) / B;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < nb; ++b) {
        Rng r(seed, (uint64_t)b + 0x57000000ull);
        for (int64_t uu = b * B; uu < std::min(n, (b + 1) * B); ++uu)
            for (int32_t j = 1; j <= h; ++j) {
                int64_t w = (uu + j) % n;
                if (p > 0.0 && r.uniform() < p) {
                    int64_t c = (int64_t)r.below((uint64_t)(n - 1));
                    w = c >= uu ? c + 1 : c;
                }
                const size_t e = (size_t)uu * h + (j - 1);
                pr[2 * e] = (int32_t)uu;
                pr[2 * e + 1] = (int32_t)w;
            }
    }
    return csr_from_pairs(n, n, pr.data(), n * h, false);
}

// One part of a stochastic block model, as one partition of a multi-part
// run would hold it.  n_blocks blocks of block_size agents, blocks split into
// n_parts contiguous parts; intra-block edge probability gives mean degree
// deg_in, inter-block probability gives mean degree deg_out.  Rows are the
// local agents of `part` (global ids part*bpp*block_size ...); columns are
// local slots [0, n_local) followed by ghost slots, one per distinct remote
// endpoint, in ascending global order.  extra[j] = global index of ghost j.
sg_graph *sg_sbm_part(int32_t n_blocks, int64_t block_size, double deg_in, double deg_out,
                      int32_t n_parts, int32_t part, uint64_t seed) {
    if (n_blocks < 1 || n_parts < 1 || n_blocks % n_parts || part < 0 || part >= n_parts || block_size < 2) {
        g_err = "sg_sbm_part: bad block/part layout";
        return nullptr;
    }
    const int64_t N = (int64_t)n_blocks * block_size;
    const int32_t bpp = n_blocks / n_parts;
    const int64_t lo = (int64_t)part * bpp * block_size, n_local = (int64_t)bpp * block_size;
    const double p_in = deg_in / (double)(block_size - 1);
    const double p_out = n_blocks > 1 ? deg_out / (double)(N - block_size) : 0.0;
    if (p_in > 1.0 || p_out > 1.0) { g_err = "sg_sbm_part: probability above 1"; return nullptr; }
    if (n_local > INT32_MAX) { g_err = "sg_sbm_part: part exceeds int32 slots"; return nullptr; }
    // Every block pair (a, b), a <= b, with a or b local, owns a stream keyed
    // by (a, b): the same pair draws the same edges in every part's run.
    struct Pair { int32_t a, b; };
    std::vector<Pair> pairs;
    for (int32_t a = 0; a < n_blocks; ++a)
        for (int32_t b = a; b < n_blocks; ++b)
            if (a / bpp == part || b / bpp == part) pairs.push_back({a, b});
    // per pair: edges as (local-or-remote global id) pairs, int64
    std::vector<std::vector<int64_t>> parts(pairs.size());
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t q = 0; q < pairs.size(); ++q) {
        const int32_t a = pairs[q].a, b = pairs[q].b;
        const double p = a == b ? p_in : p_out;
        if (p <= 0.0) continue;
        Rng r(seed, ((uint64_t)a << 32) | (uint64_t)b);
        const double l1p = std::log1p(-std::min(p, 1.0 - 1e-16));
        auto &pl = parts[q];
        const int64_t ba = (int64_t)a * block_size, bb = (int64_t)b * block_size;
        if (a == b) {
            // upper triangle of the block, linearised row by row
            for (int64_t i = 0; i < block_size; ++i) {
                int64_t j = i + 1;
                while (true) {
                    j += (int64_t)geo_skip(r, l1p);
                    if (j >= block_size) break;
                    pl.push_back(ba + i);
                    pl.push_back(ba + j);
                    ++j;
                }
            }
        } else {
            const uint64_t total = (uint64_t)block_size * (uint64_t)block_size;
            uint64_t x = 0;
            while (true) {
                x += geo_skip(r, l1p);
                if (x >= total) break;
                pl.push_back(ba + (int64_t)(x / block_size));
                pl.push_back(bb + (int64_t)(x % block_size));
                ++x;
            }
        }
    }
    // remote endpoints -> ghost slots in ascending global order
    std::vector<int64_t> remote;
    for (auto &pl : parts)
        for (int64_t gid : pl)
            if (gid < lo || gid >= lo + n_local) remote.push_back(gid);
    std::sort(remote.begin(), remote.end());
    remote.erase(std::unique(remote.begin(), remote.end()), remote.end());
    const int64_t n_cols = n_local + (int64_t)remote.size();
    if (n_cols > INT32_MAX) { g_err = "sg_sbm_part: local + ghost slots exceed int32"; return nullptr; }
    size_t tot = 0;
    for (auto &pl : parts) tot += pl.size();
    Pairs pr;
    pr.reserve(tot);
    for (auto &pl : parts) {
        for (int64_t gid : pl)
            pr.push_back((int32_t)(gid >= lo && gid < lo + n_local
                                       ? gid - lo
                                       : n_local + (std::lower_bound(remote.begin(), remote.end(), gid) - remote.begin())));
        std::vector<int64_t>().swap(pl);
    }
    // rows only for local agents; a ghost-ghost edge cannot occur (one end is local)
    sg_graph *g = csr_from_pairs(n_local, n_cols, pr.data(), (int64_t)pr.size() / 2, false);
    g->extra = std::move(remote);
    return g;
}

// Relabel the n rows of a CSR by descending degree (ties by ascending old
// index), keeping the order inside every row.  Columns >= n (ghost columns)
// keep their value.  perm[new] = old, inv[old] = new.  Output arrays are
// caller-allocated: out_row_ptr[n + 1], out_col_idx[nnz].
int sg_degree_sort_csr(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, int64_t *out_row_ptr,
                       int32_t *out_col_idx, int32_t *perm, int32_t *inv) {
    if (n < 0 || n > INT32_MAX) { g_err = "sg_degree_sort_csr: n out of range"; return -1; }
    // counting sort by degree (descending), stable in the old index
    int64_t maxd = 0;
    for (int64_t i = 0; i < n; ++i) maxd = std::max(maxd, row_ptr[i + 1] - row_ptr[i]);
    std::vector<int64_t> cnt(maxd + 2, 0);
    for (int64_t i = 0; i < n; ++i) cnt[maxd - (row_ptr[i + 1] - row_ptr[i])]++;
    int64_t acc = 0;
    for (auto &c : cnt) { int64_t t = c; c = acc; acc += t; }
    for (int64_t i = 0; i < n; ++i) {
        int64_t slot = cnt[maxd - (row_ptr[i + 1] - row_ptr[i])]++;
        perm[slot] = (int32_t)i;
        inv[i] = (int32_t)slot;
    }
    out_row_ptr[0] = 0;
    for (int64_t i = 0; i < n; ++i) out_row_ptr[i + 1] = out_row_ptr[i] + (row_ptr[perm[i] + 1] - row_ptr[perm[i]]);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t i = 0; i < n; ++i) {
        int64_t src = row_ptr[perm[i]], dst = out_row_ptr[i], len = out_row_ptr[i + 1] - dst;
        for (int64_t e = 0; e < len; ++e) {
            int32_t c = col_idx[src + e];
            out_col_idx[dst + e] = c < n ? inv[c] : c;
        }
    }
    return 0;
}

// Bijective scramble of [0, 2^62) used to give agents random-looking int64
// ids: odd multiplier and xor-shifts, all invertible modulo 2^62.
void sg_scramble_ids(const int64_t *in, int64_t *out, int64_t n, uint64_t seed) {
    const uint64_t MASK = (1ull << 62) - 1;
    uint64_t s = seed;
    uint64_t k1 = splitmix64(s) | 1ull, k2 = splitmix64(s) | 1ull, x0 = splitmix64(s) & MASK;
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) {
        uint64_t x = ((uint64_t)in[i] ^ x0) & MASK;
        x = (x * k1) & MASK;
        x ^= x >> 29;
        x = (x * k2) & MASK;
        x ^= x >> 31;
        out[i] = (int64_t)x;
    }
}

}  // extern "C"
