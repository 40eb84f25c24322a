// libsagestep runtime: contexts, replicas, tick loops, ghosts (C ABI of
// include/sagestep.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/sagestep.h"
#include "ss_runtime.h"

static thread_local std::string g_err;

#define SS_FAIL(ctx, code, ...)                                        \
    do {                                                               \
        char _b[512];                                                  \
        std::snprintf(_b, sizeof _b, __VA_ARGS__);                     \
        if (ctx) static_cast<ss_ctx *>(ctx)->err = _b; else g_err = _b; \
        return code;                                                   \
    } while (0)

#define SS_HIP(ctx, expr)                                                               \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            SS_FAIL(ctx, _e == hipErrorOutOfMemory ? SS_ERR_OOM : SS_ERR_HIP, "%s: %s (%s:%d)", \
                    #expr, hipGetErrorString(_e), __FILE__, __LINE__);                  \
    } while (0)

namespace {

// Length of a per-agent array padded to whole groups of 64 agents.
inline int64_t pad_agents(int64_t n) { return (n + 63) / 64 * 64; }

template <class T>
hipError_t dalloc(ss_ctx *ctx, ss_replica *r, T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    size_t b = sizeof(T) * count;
    hipError_t e = hipMalloc((void **)p, b);
    if (e == hipSuccess) {
        if (r) r->bytes += b;
        ctx->bytes += b;
    }
    return e;
}

void dfree(void *p) {
    if (p) (void)hipFree(p);
}

// inf_T[k] = ceil(p_k * 2^23), p_k = 1 - q_k, q_k = q_{k-1} * (1 - beta),
// for k = 0 .. min(max_deg, first k with p_k * 2^23 == 2^23).
std::vector<uint32_t> threshold_table(double beta, int64_t max_deg) {
    std::vector<uint32_t> T;
    T.push_back(0u);
    double q = 1.0;
    const double b = 1.0 - beta;
    for (int64_t k = 1; k <= max_deg; ++k) {
        q = q * b;
        const double p = 1.0 - q;
        const double x = p * 8388608.0;
        const uint32_t v = (uint32_t)std::ceil(x);
        T.push_back(v);
        if (v >= 8388608u) break;
    }
    return T;
}

uint32_t prob_threshold(double p) {
    if (!(p > 0.0)) return 0u;
    if (p >= 1.0) return 8388608u;
    return (uint32_t)std::ceil(p * 8388608.0);
}

bool valid_rep(ss_ctx *ctx, int32_t rep) {
    return ctx && rep >= 0 && rep < (int32_t)ctx->reps.size() && ctx->reps[rep].live;
}

int check_model(ss_ctx *ctx, const ss_model *m) {
    if (!m) SS_FAIL(ctx, SS_ERR_INVALID, "model is NULL");
    if (m->n_states < 1 || m->n_states > 4) SS_FAIL(ctx, SS_ERR_INVALID, "n_states must be 1..4");
    auto in = [&](int32_t s) { return s >= 0 && s < m->n_states; };
    if (!in(m->s_code) || !in(m->inf_code) || !in(m->s_target))
        SS_FAIL(ctx, SS_ERR_INVALID, "model state codes out of range");
    if (!(m->beta >= 0.0 && m->beta <= 1.0)) SS_FAIL(ctx, SS_ERR_INVALID, "beta must be in [0, 1]");
    for (int s = 0; s < 4; ++s) {
        if (!(m->trans_p[s] >= 0.0 && m->trans_p[s] <= 1.0))
            SS_FAIL(ctx, SS_ERR_INVALID, "trans_p[%d] must be in [0, 1]", s);
        if (s < m->n_states && m->trans_p[s] > 0.0 && !in(m->trans_next[s]))
            SS_FAIL(ctx, SS_ERR_INVALID, "trans_next[%d] out of range", s);
    }
    return SS_OK;
}

int apply_model(ss_ctx *ctx, ss_replica &r, const ss_model *m) {
    int rc = check_model(ctx, m);
    if (rc) return rc;
    r.model = *m;
    std::vector<uint32_t> T = threshold_table(m->beta, r.max_deg);
    dfree(r.inf_T);
    r.inf_T = nullptr;
    SS_HIP(ctx, dalloc(ctx, &r, &r.inf_T, T.size()));
    SS_HIP(ctx, hipMemcpy(r.inf_T, T.data(), T.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    r.dev.inf_T = r.inf_T;
    r.dev.inf_T_last = (int32_t)T.size() - 1;
    for (int s = 0; s < 4; ++s) {
        r.dev.trans_T[s] = s < m->n_states && s != m->s_code ? prob_threshold(m->trans_p[s]) : 0u;
        r.dev.trans_next[s] = (int8_t)m->trans_next[s];
    }
    r.dev.s_code = (int8_t)m->s_code;
    r.dev.inf_code = (int8_t)m->inf_code;
    r.dev.s_target = (int8_t)m->s_target;
    r.dev.n_states = (int8_t)m->n_states;
    // an agent that is not susceptible at tick t is not at t + 1 either when
    // infection leaves s_code and no transition enters it: the pushed counts
    // of such agents would never be read
    bool back = m->s_target == m->s_code;
    for (int s = 0; s < 4; ++s) back |= r.dev.trans_T[s] != 0u && r.dev.trans_next[s] == r.dev.s_code;
    r.dev.push_sus_only = back ? 0u : 1u;
    return SS_OK;
}

void free_replica(ss_replica &r) {
    dfree(r.row_ptr); dfree(r.col_idx); dfree(r.rng_lo); dfree(r.rng_hi); dfree(r.deg8);
    dfree(r.state[0]); dfree(r.state[1]); dfree(r.init_state);
    dfree(r.ibits[0]); dfree(r.ibits[1]);
    dfree(r.history); dfree(r.init_history); dfree(r.inf_T);
    dfree(r.cnt[0]); dfree(r.cnt[1]); dfree(r.dec); dfree(r.acc);
    dfree(r.map_keys); dfree(r.map_vals); dfree(r.halo_slots);
    for (auto &d : r.halo_dev) dfree(d);
    for (auto &d : r.halo_grp) dfree(d);
    dfree((void *)r.halo_ptrs);
    dfree(r.halo_rec);
    r = ss_replica();
}

// rebuild the infectious bitmap of buffer `par` over words [w0, w1)
int rebuild_bits(ss_ctx *ctx, ss_replica &r, int par, int64_t w0, int64_t w1) {
    SS_HIP(ctx, ss_launch_bits(r.state[par], r.ibits[par], w0, w1, r.n_total, r.dev.inf_code, ctx->stream));
    return SS_OK;
}

int64_t heavy_prefix(const ss_replica &r, int32_t thr) {
    for (size_t q = 0; q < r.heavy_prefix_deg.size(); ++q)
        if (r.heavy_prefix_deg[q] == thr) return r.heavy_prefix_len[q];
    return -1;
}

// Direction state of the replica's current tick t: statistics of t in slot
// t % 3 (the other slots cleared), the push decision for t and, when it
// pushes, the infectious-neighbour counts of t from the infectious bitmap.
int init_acc(ss_ctx *ctx, ss_replica &r) {
    SS_HIP(ctx, hipMemsetAsync(r.acc, 0, sizeof(unsigned long long) * SS_ACC_WORDS, ctx->stream));
    const int slot = (int)((uint64_t)r.tick % 3u);
    SS_HIP(ctx, ss_launch_acc_init(r.state[r.tick & 1], r.row_ptr, r.n_local, r.dev.s_code, r.dev.inf_code, r.acc,
                                   slot, ctx->num_cus * 8, ctx->stream));
    if (r.cnt[0]) {
        const int par = (int)(r.tick & 1);
        const uint32_t num = r.dev.push_ok ? ctx->push_num : 0u;
        for (int b = 0; b < 2; ++b)
            SS_HIP(ctx, hipMemsetAsync(r.cnt[b], 0, sizeof(uint16_t) * (size_t)pad_agents(r.n_local), ctx->stream));
        SS_HIP(ctx, ss_launch_count_init(r.row_ptr, r.col_idx, r.ibits[par], r.n_local, heavy_prefix(r, 64),
                                         r.cnt[par], r.dec + par, r.acc + slot * (SS_ACC_SLOTS * 2), num,
                                         ctx->num_cus * 8, ctx->stream));
    }
    return SS_OK;
}

void drop_graphs(ss_ctx *ctx) {
    for (auto &kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
    ctx->graphs.clear();
}


// Tiles of every live replica: one HUB tile per row of degree >= hub_deg at the
// front, HEAVY tiles of SS_HEAVY_TILE agents up to the first 64-aligned slot
// past the rows of degree >= heavy_deg, LIGHT tiles for the rest (light tiles
// start 64-aligned: no bitmap word is shared by two light tiles, the hub and
// heavy tiles set single bits).  A light tile holds 256 * apl agents; with 16
// agents per lane it holds 4 * 1024, fewer where its rows are longer than the
// replica's average light row: a tile of rows of degree d > dbar holds dbar / d
// of that (about as many CSR entries as an average tile), in steps of 256
// agents down to 256 and of 64 below, so that the tiles at the long-row end
// cost about as much as the others.  Blocks stride over the list (see below).
int build_tiles(ss_ctx *ctx) {
    std::vector<ss_tile> hubs, heavy, light;
    int bpc = 0;
    SS_HIP(ctx, ss_occupancy_tick(ctx->apl, &bpc));
    if (bpc < 1) bpc = 1;
    const int g = ctx->grid_blocks > 0 ? ctx->grid_blocks : ctx->num_cus * bpc;
    const int64_t cap = (int64_t)(ctx->apl == 16 ? 4 * 1024 : SS_BLOCK * ctx->apl);
    struct Pos { int64_t hub, heavy, light, lblocks; };
    std::vector<Pos> pos(ctx->reps.size(), Pos{0, 0, 0, 0});
    int64_t n_lblocks = 0;                       // tag slots of light agents, all replicas
    for (int32_t ri = 0; ri < (int32_t)ctx->reps.size(); ++ri) {
        ss_replica &r = ctx->reps[ri];
        if (!r.live || r.n_total == 0) continue;
        const int64_t H = heavy_prefix(r, ctx->heavy_deg), Hh = heavy_prefix(r, ctx->hub_deg);
        if (H < 0 || Hh < 0) SS_FAIL(ctx, SS_ERR_INVALID, "degree threshold without a prefix table");
        const int64_t nt = r.n_total;
        const int64_t H64 = std::min<int64_t>((H + 63) / 64 * 64, nt);
        // hub count rounded up to 8: the agents of a heavy wavefront (4) then
        // always share one bitmap word
        const int64_t nh = std::min((Hh + 7) / 8 * 8, H64);
        pos[ri] = Pos{(int64_t)hubs.size(), (int64_t)heavy.size(), (int64_t)light.size(), n_lblocks};
        for (int64_t b = 0; b < nh; ++b) hubs.push_back({ri, SS_TILE_HUB, b, b + 1, 0, 0});
        for (int64_t b = nh; b < H64; b += SS_HEAVY_TILE)
            heavy.push_back({ri, SS_TILE_HEAVY, b, std::min<int64_t>(b + SS_HEAVY_TILE, H64), 0, 0});
        // CSR entries before agent i (i a multiple of 64, or n_total)
        auto E = [&](int64_t i) { return i >= nt ? r.edge64.back() : r.edge64[(size_t)(i / 64)]; };
        const double dbar = H64 < nt ? (double)(E(nt) - E(H64)) / (double)(nt - H64) : 0.0;
        for (int64_t b = H64; b < nt;) {
            const int64_t g1 = std::min<int64_t>(b + 64, nt);
            const double d = (double)(E(g1) - E(b)) / (double)(g1 - b);
            int64_t a = cap;
            if (ctx->apl == 16 && dbar > 0.0 && d > dbar) a = (int64_t)((double)cap * dbar / d);
            // whole groups of 4 x 64 agents: the 4 wavefronts of a wide tile get
            // equal chunks of whole bitmap words; a wide tile of long rows may
            // hold 64, 128 or 192 agents (chunks of 16, 32 or 48)
            const int64_t step = ctx->apl == 16 && a >= 256 ? 256 : 64;
            a = std::max<int64_t>(64, std::min(a, cap) / step * step);
            const int64_t e = std::min(b + a, nt);
            light.push_back({ri, SS_TILE_LIGHT, b, e, 0, 0});
            b = e;
        }
        n_lblocks += ((nt - H64) + (1 << SS_LTAG_SHIFT) - 1) >> SS_LTAG_SHIFT;
        r.dev.n_hub = nh;
        r.dev.h64 = H64;
        r.dev.light_cap = (int32_t)cap;
    }
    const int64_t nhub = (int64_t)hubs.size(), nheavy = (int64_t)heavy.size(), n_block = nhub + nheavy;
    // tag slots: one per hub or heavy tile (its place in the list before the
    // rounds below are reordered); light agents by blocks after them
    for (auto &t : hubs) t.tag_lo = t.tag_hi = (int32_t)(pos[t.rep].hub + t.begin);
    for (auto &t : heavy)
        t.tag_lo = t.tag_hi = (int32_t)(nhub + pos[t.rep].heavy + (t.begin - ctx->reps[t.rep].dev.n_hub) / SS_HEAVY_TILE);
    for (auto &t : light) {
        const int64_t l0 = n_block + pos[t.rep].lblocks, h64 = ctx->reps[t.rep].dev.h64;
        t.tag_lo = (int32_t)(l0 + ((t.begin - h64) >> SS_LTAG_SHIFT));
        t.tag_hi = (int32_t)(l0 + ((t.end - 1 - h64) >> SS_LTAG_SHIFT));
    }
    std::vector<ss_tile> all;
    all.reserve(hubs.size() + heavy.size() + light.size());
    all.insert(all.end(), hubs.begin(), hubs.end());
    all.insert(all.end(), heavy.begin(), heavy.end());
    all.insert(all.end(), light.begin(), light.end());
    // block b runs tiles b, b + G, b + 2G, ... of the list; the list runs from
    // the longest rows to the shortest, so every other round of G tiles is
    // reversed: no block gets the most expensive tile of every round
    const int64_t G = std::max<int64_t>(1, std::min<int64_t>(g, (int64_t)all.size()));
    for (int64_t r0 = G; r0 < (int64_t)all.size(); r0 += 2 * G)
        std::reverse(all.begin() + r0, all.begin() + std::min<int64_t>(r0 + G, (int64_t)all.size()));
    const int64_t n_slots = n_block + n_lblocks;
    if ((int64_t)all.size() > INT32_MAX / 2 || n_slots > INT32_MAX / 2) SS_FAIL(ctx, SS_ERR_RANGE, "too many tiles");
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if ((int64_t)all.size() > ctx->tiles_cap) {
        dfree(ctx->d_tiles);
        ctx->d_tiles = nullptr;
        int64_t nc = std::max<int64_t>((int64_t)all.size() * 5 / 4, 1024);
        SS_HIP(ctx, dalloc(ctx, nullptr, &ctx->d_tiles, (size_t)nc));
        ctx->tiles_cap = nc;
    }
    if (n_slots > ctx->slots_cap) {
        dfree(ctx->d_tact[0]);
        dfree(ctx->d_tact[1]);
        ctx->d_tact[0] = ctx->d_tact[1] = nullptr;
        int64_t nc = std::max<int64_t>(n_slots * 5 / 4, 1024);
        SS_HIP(ctx, dalloc(ctx, nullptr, &ctx->d_tact[0], (size_t)nc));
        SS_HIP(ctx, dalloc(ctx, nullptr, &ctx->d_tact[1], (size_t)nc));
        ctx->slots_cap = nc;
    }
    // tile tags start cleared: a replica that was pushing with tile tags runs
    // every tile at its next tick (its agent tags stay valid)
    SS_HIP(ctx, hipMemsetAsync(ctx->d_tact[0], 0, (size_t)ctx->slots_cap, ctx->stream));
    SS_HIP(ctx, hipMemsetAsync(ctx->d_tact[1], 0, (size_t)ctx->slots_cap, ctx->stream));
    for (auto &r : ctx->reps)
        if (r.live && r.dec) SS_HIP(ctx, ss_launch_dec_demote(r.dec, ctx->stream));
    if (!all.empty())
        SS_HIP(ctx, hipMemcpyAsync(ctx->d_tiles, all.data(), all.size() * sizeof(ss_tile), hipMemcpyHostToDevice,
                                   ctx->stream));
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->n_tiles = (int)all.size();
    for (int32_t ri = 0; ri < (int32_t)ctx->reps.size(); ++ri) {
        ss_replica &r = ctx->reps[ri];
        r.dev.tact[0] = ctx->d_tact[0];
        r.dev.tact[1] = ctx->d_tact[1];
        r.dev.hub0 = (int32_t)pos[ri].hub;
        r.dev.heavy0 = (int32_t)(nhub + pos[ri].heavy);
        r.dev.light0 = (int32_t)(n_block + pos[ri].lblocks);
    }
    ctx->grid = std::max(1, std::min(g, std::max(ctx->n_tiles, 1)));
    // the block of a replica's first tile writes its per-tick flags
    for (auto &r : ctx->reps) r.dev.lead_tile = -1;
    for (int32_t q = (int32_t)all.size() - 1; q >= 0; --q) ctx->reps[all[q].rep].dev.lead_tile = q;
    drop_graphs(ctx);
    ctx->tiles_dirty = false;
    return SS_OK;
}

int ensure_rep_table(ss_ctx *ctx) {
    int need = (int)ctx->reps.size();
    if (need <= ctx->reps_cap) return SS_OK;
    int nc = std::max(need * 2, 16);
    ss_rep_dev *d = nullptr, *h = nullptr;
    SS_HIP(ctx, dalloc(ctx, nullptr, &d, (size_t)nc));
    SS_HIP(ctx, hipHostMalloc((void **)&h, sizeof(ss_rep_dev) * nc, hipHostMallocDefault));
    std::memset(h, 0, sizeof(ss_rep_dev) * nc);
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    dfree(ctx->d_reps);
    if (ctx->h_reps) (void)hipHostFree(ctx->h_reps);
    ctx->d_reps = d;
    ctx->h_reps = h;
    ctx->reps_cap = nc;
    drop_graphs(ctx);  // captured launches hold the old table pointer
    return SS_OK;
}

// copy descriptors with tick window [tick, tick + ticks) clipped to each limit
int push_descriptors(ss_ctx *ctx, int64_t ticks) {
    int rc = ensure_rep_table(ctx);
    if (rc) return rc;
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));  // the mirror may still be in flight
    for (size_t i = 0; i < ctx->reps.size(); ++i) {
        ss_replica &r = ctx->reps[i];
        ss_rep_dev d = r.dev;
        if (r.live) {
            d.tick0 = r.tick;
            int64_t end = ticks > INT64_MAX - r.tick ? INT64_MAX : r.tick + ticks;
            d.tick_end = std::min(end, r.tick_limit);
            d.push_num = ctx->push_num;
        } else {
            d.tick0 = 0;
            d.tick_end = 0;
        }
        ctx->h_reps[i] = d;
    }
    SS_HIP(ctx, hipMemcpyAsync(ctx->d_reps, ctx->h_reps, sizeof(ss_rep_dev) * ctx->reps.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    return SS_OK;
}

// after a loop of `ticks` launches every live replica sits at
// min(tick + ticks, tick_limit), or where it was if already past its limit
void advance_ticks(ss_ctx *ctx, int64_t ticks) {
    for (auto &r : ctx->reps)
        if (r.live) {
            int64_t end = ticks > INT64_MAX - r.tick ? INT64_MAX : r.tick + ticks;
            end = std::min(end, r.tick_limit);
            r.tick = std::max(r.tick, end);
        }
}

// launch step k of a tick loop
hipError_t tick_launch(ss_ctx *ctx, int64_t k) {
    return ss_launch_tick(ctx->d_reps, ctx->d_tiles, ctx->n_tiles, (int)k, ctx->apl, ctx->grid,
                          ctx->counting ? ctx->d_work : nullptr, ctx->halo_fused, ctx->stream);
}

int launch_loop(ss_ctx *ctx, int64_t ticks) {
    if (ticks > INT32_MAX) SS_FAIL(ctx, SS_ERR_RANGE, "ticks per call must fit int32");
    if (!ctx->use_graph) {
        for (int64_t k = 0; k < ticks; ++k) SS_HIP(ctx, tick_launch(ctx, k));
        return SS_OK;
    }
    auto it = ctx->graphs.find(ticks);
    if (it == ctx->graphs.end()) {
        hipGraph_t g = nullptr;
        SS_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        hipError_t le = hipSuccess;
        for (int64_t k = 0; k < ticks && le == hipSuccess; ++k) le = tick_launch(ctx, k);
        hipError_t ee = hipStreamEndCapture(ctx->stream, &g);
        SS_HIP(ctx, le);
        SS_HIP(ctx, ee);
        hipGraphExec_t ex = nullptr;
        hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        SS_HIP(ctx, ie);
        it = ctx->graphs.emplace(ticks, ex).first;
    }
    SS_HIP(ctx, hipGraphLaunch(it->second, ctx->stream));
    return SS_OK;
}

}  // namespace

extern "C" {

const char *ss_last_error(const ss_ctx *ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

static hipError_t create_halo_events(ss_ctx *ctx) {
    hipError_t e = hipSuccess;
    for (int i = 0; i < SS_HALO_RING && e == hipSuccess; ++i) {
        e = hipEventCreateWithFlags(&ctx->halo_ev[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->halo_used[i], hipEventDisableTiming);
    }
    for (int b = 0; b < 2 && e == hipSuccess; ++b) {
        e = hipEventCreateWithFlags(&ctx->grp_done[b], hipEventDisableTiming);
        for (int i = 0; i < SS_HALO_SPANS && e == hipSuccess; ++i)
            e = hipEventCreateWithFlags(&ctx->span_ev[b][i], hipEventDisableTiming);
    }
    return e;
}

int ss_create(int device, ss_ctx **out) {
    if (!out) SS_FAIL(nullptr, SS_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) SS_FAIL(nullptr, SS_ERR_HIP, "no HIP device: %s", hipGetErrorString(e));
    if (device < 0 || device >= ndev) SS_FAIL(nullptr, SS_ERR_INVALID, "device %d out of range", device);
    auto *ctx = new ss_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    if ((e = hipSetDevice(device)) != hipSuccess || (e = hipGetDeviceProperties(&prop, device)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&ctx->t_start)) != hipSuccess || (e = hipEventCreate(&ctx->t_stop)) != hipSuccess ||
        (e = hipMalloc((void **)&ctx->d_work, SS_WORK_N * sizeof(unsigned long long))) != hipSuccess ||
        (e = hipMalloc((void **)&ctx->d_stats, 8 * sizeof(unsigned long long))) != hipSuccess ||
        (e = hipMemset(ctx->d_work, 0, SS_WORK_N * sizeof(unsigned long long))) != hipSuccess ||
        (e = hipMalloc((void **)&ctx->d_counters, 64 * sizeof(int32_t))) != hipSuccess ||
        (e = create_halo_events(ctx)) != hipSuccess) {
        delete ctx;
        SS_FAIL(nullptr, SS_ERR_HIP, "context setup: %s", hipGetErrorString(e));
    }
    ctx->num_cus = prop.multiProcessorCount;
    ctx->total_mem = (int64_t)prop.totalGlobalMem;
    *out = ctx;
    return SS_OK;
}

int ss_destroy(ss_ctx *ctx) {
    if (!ctx) return SS_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->copy_stream);
    drop_graphs(ctx);
    for (auto &r : ctx->reps) free_replica(r);
    dfree(ctx->d_reps);
    dfree(ctx->d_tiles);
    dfree(ctx->d_tact[0]);
    dfree(ctx->d_tact[1]);
    dfree(ctx->d_counters);
    dfree(ctx->d_work);
    dfree(ctx->d_stats);
    if (ctx->t_start) (void)hipEventDestroy(ctx->t_start);
    if (ctx->t_stop) (void)hipEventDestroy(ctx->t_stop);
    if (ctx->h_reps) (void)hipHostFree(ctx->h_reps);
    for (int i = 0; i < SS_HALO_RING; ++i) {
        if (ctx->h_halo_pin[i]) (void)hipHostFree(ctx->h_halo_pin[i]);
        if (ctx->halo_ev[i]) (void)hipEventDestroy(ctx->halo_ev[i]);
        if (ctx->halo_used[i]) (void)hipEventDestroy(ctx->halo_used[i]);
    }
    for (int b = 0; b < 2; ++b) {
        if (ctx->grp_done[b]) (void)hipEventDestroy(ctx->grp_done[b]);
        for (int i = 0; i < SS_HALO_SPANS; ++i)
            if (ctx->span_ev[b][i]) (void)hipEventDestroy(ctx->span_ev[b][i]);
    }
    (void)hipStreamDestroy(ctx->stream);
    (void)hipStreamDestroy(ctx->copy_stream);
    delete ctx;
    return SS_OK;
}

int ss_sync(ss_ctx *ctx) {
    if (!ctx) SS_FAIL(nullptr, SS_ERR_INVALID, "ctx is NULL");
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SS_OK;
}

int ss_device_info(ss_ctx *ctx, int32_t *num_cus, int64_t *total_mem, int64_t *ctx_bytes) {
    if (!ctx) SS_FAIL(nullptr, SS_ERR_INVALID, "ctx is NULL");
    if (num_cus) *num_cus = ctx->num_cus;
    if (total_mem) *total_mem = ctx->total_mem;
    if (ctx_bytes) *ctx_bytes = (int64_t)ctx->bytes;
    return SS_OK;
}

int64_t ss_replica_footprint(int64_t n_local, int64_t n_ghost, int64_t nnz, int32_t with_history, int64_t max_deg) {
    int64_t n_total = n_local + n_ghost;
    int64_t words = (n_total + 63) / 64;
    int64_t b = 8 * (n_total + 1) + 4 * nnz + 6 * pad_agents(n_local) + 2 * pad_agents(n_total) + n_total +
                2 * 8 * words + 4 * (std::min<int64_t>(max_deg, 1 << 24) + 1) + 8 * SS_ACC_WORDS +
                (n_ghost == 0 ? 4 * pad_agents(n_local) + 8 : 0);
    if (with_history) b += 2 * 4 * n_local * SS_HISTORY_WIDTH;
    return b;
}

int ss_replica_create(ss_ctx *ctx, const ss_replica_spec *sp, int32_t *rep_out) {
    if (!ctx || !sp || !rep_out) SS_FAIL(ctx, SS_ERR_INVALID, "NULL argument");
    (void)hipSetDevice(ctx->device);
    const int64_t nl = sp->n_local, ng = sp->n_ghost, nnz = sp->nnz, nt = nl + ng;
    if (nl < 0 || ng < 0 || nnz < 0 || nt > INT32_MAX) SS_FAIL(ctx, SS_ERR_RANGE, "slot counts out of range");
    if ((nl > 0 && (!sp->row_ptr || !sp->rng_key_id)) || (nt > 0 && !sp->state) || (nnz > 0 && !sp->col_idx))
        SS_FAIL(ctx, SS_ERR_INVALID, "missing arrays");
    // shape checks on the host before anything reaches a kernel
    int64_t max_deg = 0;
    if (nl > 0) {
        if (sp->row_ptr[0] != 0 || sp->row_ptr[nl] != nnz) SS_FAIL(ctx, SS_ERR_INVALID, "row_ptr[0] must be 0 and row_ptr[n_local] == nnz");
        for (int64_t i = 0; i < nl; ++i) {
            int64_t d = sp->row_ptr[i + 1] - sp->row_ptr[i];
            if (d < 0) SS_FAIL(ctx, SS_ERR_INVALID, "row_ptr decreases at row %lld", (long long)i);
            max_deg = std::max(max_deg, d);
        }
    } else if (nnz != 0) {
        SS_FAIL(ctx, SS_ERR_INVALID, "nnz must be 0 without local rows");
    }
    // col_idx is range-checked on the device right after its upload (below),
    // before any kernel indexes with it.
    int rc = check_model(ctx, &sp->model);
    if (rc) return rc;
    bool need_hi = false;
    for (int64_t i = 0; i < nl; ++i)
        if ((uint64_t)sp->rng_key_id[i] >> 32) { need_hi = true; break; }

    ss_replica r;
    r.n_local = nl; r.n_ghost = ng; r.n_total = nt; r.nnz = nnz;
    r.n_words = (nt + 63) / 64;
    r.max_deg = max_deg;
    // CSR entries before agents 0, 64, 128, ... and before n_total (light tile sizes)
    r.edge64.resize((size_t)((nt + 63) / 64 + 1));
    for (int64_t j = 0; j * 64 < nt; ++j) r.edge64[(size_t)j] = nl > 0 ? sp->row_ptr[std::min(j * 64, nl)] : 0;
    r.edge64.back() = nnz;
    // heavy prefix lengths for the tiling thresholds the autotuner may pick
    for (int32_t thr : SS_THRESHOLDS) {
        int64_t H = 0;
        while (H < nl && sp->row_ptr[H + 1] - sp->row_ptr[H] >= thr) ++H;
        r.heavy_prefix_deg.push_back(thr);
        r.heavy_prefix_len.push_back(H);
    }
    auto fail = [&](int code) {
        ctx->bytes -= r.bytes;
        free_replica(r);
        return code;
    };
#define R_HIP(expr)                                                                  \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) {                                                      \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(_e);            \
            return fail(_e == hipErrorOutOfMemory ? SS_ERR_OOM : SS_ERR_HIP);        \
        }                                                                            \
    } while (0)
    R_HIP(dalloc(ctx, &r, &r.row_ptr, (size_t)nt + 1));
    // whole 16-byte groups: a lane's scan reads its row through aligned 16-byte loads
    R_HIP(dalloc(ctx, &r, &r.col_idx, (size_t)(nnz + 3) / 4 * 4));
    if (nnz & 3) R_HIP(hipMemset(r.col_idx + nnz, 0, sizeof(int32_t) * (size_t)(4 - (nnz & 3))));
    // per-agent arrays the light tiles read as 16-agent vectors are padded
    R_HIP(dalloc(ctx, &r, &r.rng_lo, (size_t)pad_agents(nl)));
    R_HIP(dalloc(ctx, &r, &r.deg8, (size_t)pad_agents(nl)));
    if (need_hi) R_HIP(dalloc(ctx, &r, &r.rng_hi, (size_t)nl));
    for (int b = 0; b < 2; ++b) {
        R_HIP(dalloc(ctx, &r, &r.state[b], (size_t)pad_agents(nt)));
        R_HIP(dalloc(ctx, &r, &r.ibits[b], (size_t)r.n_words));
    }
    R_HIP(dalloc(ctx, &r, &r.init_state, (size_t)nt));
    R_HIP(dalloc(ctx, &r, &r.acc, (size_t)SS_ACC_WORDS));
    if (ng == 0 && nl > 0) {
        R_HIP(dalloc(ctx, &r, &r.cnt[0], (size_t)pad_agents(nl)));
        R_HIP(dalloc(ctx, &r, &r.cnt[1], (size_t)pad_agents(nl)));
        R_HIP(dalloc(ctx, &r, &r.dec, (size_t)2));
        R_HIP(hipMemset(r.dec, 0, 2 * sizeof(uint32_t)));
    }
    if (sp->history) {
        R_HIP(dalloc(ctx, &r, &r.history, (size_t)nl * SS_HISTORY_WIDTH));
        R_HIP(dalloc(ctx, &r, &r.init_history, (size_t)nl * SS_HISTORY_WIDTH));
    }
    // uploads (synchronous: host arrays are borrowed only for this call)
    {
        std::vector<int64_t> rp((size_t)nt + 1);
        if (nl > 0) std::memcpy(rp.data(), sp->row_ptr, sizeof(int64_t) * (nl + 1));
        else rp[0] = 0;
        for (int64_t i = nl + 1; i <= nt; ++i) rp[i] = nnz;  // ghost rows are empty
        R_HIP(hipMemcpy(r.row_ptr, rp.data(), sizeof(int64_t) * rp.size(), hipMemcpyHostToDevice));
    }
    if (nnz) {
        R_HIP(hipMemcpy(r.col_idx, sp->col_idx, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
        unsigned long long first_bad = ~0ull;
        R_HIP(hipMemcpy(ctx->d_stats, &first_bad, sizeof first_bad, hipMemcpyHostToDevice));
        R_HIP(ss_launch_check_cols(r.col_idx, nnz, nt, ctx->d_stats, ctx->num_cus * 8, ctx->stream));
        R_HIP(hipStreamSynchronize(ctx->stream));
        R_HIP(hipMemcpy(&first_bad, ctx->d_stats, sizeof first_bad, hipMemcpyDeviceToHost));
        if (first_bad != ~0ull) {
            char b[200];
            std::snprintf(b, sizeof b, "col_idx[%llu] = %d outside [0, %lld)", first_bad,
                          sp->col_idx[first_bad], (long long)nt);
            ctx->err = b;
            return fail(SS_ERR_RANGE);
        }
    }
    {
        std::vector<uint32_t> lo((size_t)std::max<int64_t>(nl, 1)), hi;
        if (need_hi) hi.resize((size_t)nl);
        uint64_t kh = 1469598103934665603ull;
        for (int64_t i = 0; i < nl; ++i) {
            lo[i] = (uint32_t)sp->rng_key_id[i];
            if (need_hi) hi[i] = (uint32_t)((uint64_t)sp->rng_key_id[i] >> 32);
            kh = (kh ^ (uint64_t)sp->rng_key_id[i]) * 1099511628211ull;
        }
        r.key_hash = kh;
        if (nl) R_HIP(hipMemcpy(r.rng_lo, lo.data(), sizeof(uint32_t) * nl, hipMemcpyHostToDevice));
        std::vector<uint8_t> d8((size_t)std::max<int64_t>(nl, 1));
        for (int64_t i = 0; i < nl; ++i)
            d8[i] = (uint8_t)std::min<int64_t>(sp->row_ptr[i + 1] - sp->row_ptr[i], 255);
        if (nl) R_HIP(hipMemcpy(r.deg8, d8.data(), nl, hipMemcpyHostToDevice));
        if (need_hi) R_HIP(hipMemcpy(r.rng_hi, hi.data(), sizeof(uint32_t) * nl, hipMemcpyHostToDevice));
    }
    if (nt) R_HIP(hipMemcpy(r.init_state, sp->state, (size_t)nt, hipMemcpyHostToDevice));
    // the push direction needs every entry's mirror (a neighbour's row then
    // lists the agent) and no ghosts (ghost rows are empty)
    r.symmetric = nnz == 0;
    if (nnz > 0) {
        unsigned long long h[4] = {0, 0, 0, 0};
        R_HIP(hipMemcpy(ctx->d_stats, h, sizeof h, hipMemcpyHostToDevice));
        R_HIP(ss_launch_symcheck(r.row_ptr, r.col_idx, nl, ctx->d_stats, ctx->num_cus * 8, ctx->stream));
        R_HIP(hipStreamSynchronize(ctx->stream));
        R_HIP(hipMemcpy(h, ctx->d_stats, sizeof h, hipMemcpyDeviceToHost));
        r.symmetric = h[0] == h[2] && h[1] == h[3];
        r.csr_hash[0] = h[0];
        r.csr_hash[1] = h[1];
    }
    if (sp->history)
        R_HIP(hipMemcpy(r.init_history, sp->history, sizeof(int32_t) * nl * SS_HISTORY_WIDTH, hipMemcpyHostToDevice));
    r.dev = ss_rep_dev{};
    r.dev.row_ptr = r.row_ptr;
    r.dev.col_idx = r.col_idx;
    r.dev.rng_lo = r.rng_lo;
    r.dev.rng_hi = r.rng_hi;
    r.dev.deg8 = r.deg8;
    r.dev.state[0] = r.state[0];
    r.dev.state[1] = r.state[1];
    r.dev.ibits[0] = r.ibits[0];
    r.dev.ibits[1] = r.ibits[1];
    r.dev.history = r.history;
    r.dev.cnt[0] = r.cnt[0];
    r.dev.cnt[1] = r.cnt[1];
    r.dev.dec = r.dec;
    r.dev.acc = r.acc;
    // 16-bit counts: every count stays below 65536
    r.dev.push_ok = r.symmetric && r.cnt[0] && max_deg < 65535 ? 1u : 0u;
    r.dev.lead_tile = -1;
    r.dev.seed_lo = (uint32_t)sp->seed;
    r.dev.seed_hi = (uint32_t)(sp->seed >> 32);
    r.dev.n_local = nl;
    r.dev.n_total = nt;
    rc = apply_model(ctx, r, &sp->model);
    if (rc) return fail(rc);
    r.tick = r.tick_init = sp->tick0;
    r.tick_limit = sp->tick_limit > 0 ? sp->tick_limit : INT64_MAX;
    r.live = true;
    // slot reuse keeps replica ids small for long-running serving loops
    int32_t id = -1;
    for (int32_t i = 0; i < (int32_t)ctx->reps.size(); ++i)
        if (!ctx->reps[i].live) { id = i; break; }
    if (id < 0) { id = (int32_t)ctx->reps.size(); ctx->reps.emplace_back(); }
    ctx->reps[id] = r;
    ctx->tiles_dirty = true;
    *rep_out = id;
#undef R_HIP
    return ss_replica_reset(ctx, id);
}

int ss_replica_destroy(ss_ctx *ctx, int32_t rep) {
    if (!valid_rep(ctx, rep)) SS_FAIL(ctx, SS_ERR_INVALID, "no live replica %d", rep);
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SS_HIP(ctx, hipStreamSynchronize(ctx->copy_stream));
    ctx->bytes -= ctx->reps[rep].bytes;
    free_replica(ctx->reps[rep]);
    ctx->tiles_dirty = true;
    return SS_OK;
}

int ss_replica_reset(ss_ctx *ctx, int32_t rep) {
    if (!valid_rep(ctx, rep)) SS_FAIL(ctx, SS_ERR_INVALID, "no live replica %d", rep);
    ss_replica &r = ctx->reps[rep];
    r.tick = r.tick_init;
    const int par = (int)(r.tick & 1);
    if (r.n_total) {
        SS_HIP(ctx, hipMemcpyAsync(r.state[par], r.init_state, (size_t)r.n_total, hipMemcpyDeviceToDevice, ctx->stream));
        SS_HIP(ctx, hipMemcpyAsync(r.state[par ^ 1], r.init_state, (size_t)r.n_total, hipMemcpyDeviceToDevice, ctx->stream));
    }
    if (r.history)
        SS_HIP(ctx, hipMemcpyAsync(r.history, r.init_history, sizeof(int32_t) * r.n_local * SS_HISTORY_WIDTH,
                                   hipMemcpyDeviceToDevice, ctx->stream));
    int rc = rebuild_bits(ctx, r, par, 0, r.n_words);
    if (rc) return rc;
    if ((rc = rebuild_bits(ctx, r, par ^ 1, 0, r.n_words))) return rc;
    return init_acc(ctx, r);
}

int ss_replica_tick(ss_ctx *ctx, int32_t rep, int64_t *tick) {
    if (!valid_rep(ctx, rep) || !tick) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    *tick = ctx->reps[rep].tick;
    return SS_OK;
}

int ss_replica_set_tick(ss_ctx *ctx, int32_t rep, int64_t tick, int64_t tick_limit) {
    if (!valid_rep(ctx, rep) || tick < 0) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or tick");
    ss_replica &r = ctx->reps[rep];
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if ((tick & 1) != (r.tick & 1) && r.n_total) {
        // keep "tick t reads state[t & 1]": move the live buffer
        const int from = (int)(r.tick & 1);
        SS_HIP(ctx, hipMemcpyAsync(r.state[from ^ 1], r.state[from], (size_t)r.n_total, hipMemcpyDeviceToDevice, ctx->stream));
        int rc = rebuild_bits(ctx, r, from ^ 1, 0, r.n_words);
        if (rc) return rc;
    }
    r.tick = tick;
    r.tick_limit = tick_limit > 0 ? tick_limit : INT64_MAX;
    return init_acc(ctx, r);
}

int ss_replica_bytes(ss_ctx *ctx, int32_t rep, int64_t *bytes) {
    if (!valid_rep(ctx, rep) || !bytes) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    *bytes = (int64_t)ctx->reps[rep].bytes;
    return SS_OK;
}

int ss_set_model(ss_ctx *ctx, int32_t rep, const ss_model *m) {
    if (!valid_rep(ctx, rep)) SS_FAIL(ctx, SS_ERR_INVALID, "no live replica %d", rep);
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ss_replica &r = ctx->reps[rep];
    int rc = apply_model(ctx, r, m);
    if (rc) return rc;
    const int par = (int)(r.tick & 1);
    if ((rc = rebuild_bits(ctx, r, par, 0, r.n_words))) return rc;
    return init_acc(ctx, r);
}

int ss_set_seed(ss_ctx *ctx, int32_t rep, uint64_t seed) {
    if (!valid_rep(ctx, rep)) SS_FAIL(ctx, SS_ERR_INVALID, "no live replica %d", rep);
    ctx->reps[rep].dev.seed_lo = (uint32_t)seed;
    ctx->reps[rep].dev.seed_hi = (uint32_t)(seed >> 32);
    return SS_OK;
}

int ss_state_download(ss_ctx *ctx, int32_t rep, int8_t *out, int64_t n) {
    if (!valid_rep(ctx, rep) || (!out && n)) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    ss_replica &r = ctx->reps[rep];
    if (n != r.n_total) SS_FAIL(ctx, SS_ERR_INVALID, "n = %lld, replica has %lld slots", (long long)n, (long long)r.n_total);
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (n) SS_HIP(ctx, hipMemcpy(out, r.state[r.tick & 1], (size_t)n, hipMemcpyDeviceToHost));
    return SS_OK;
}

int ss_state_upload(ss_ctx *ctx, int32_t rep, const int8_t *in, int64_t n) {
    if (!valid_rep(ctx, rep) || (!in && n)) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    ss_replica &r = ctx->reps[rep];
    if (n != r.n_total) SS_FAIL(ctx, SS_ERR_INVALID, "n = %lld, replica has %lld slots", (long long)n, (long long)r.n_total);
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int par = (int)(r.tick & 1);
    if (n) {
        SS_HIP(ctx, hipMemcpy(r.state[par], in, (size_t)n, hipMemcpyHostToDevice));
        SS_HIP(ctx, hipMemcpy(r.state[par ^ 1], in, (size_t)n, hipMemcpyHostToDevice));
    }
    int rc = rebuild_bits(ctx, r, par, 0, r.n_words);
    if (rc) return rc;
    if ((rc = rebuild_bits(ctx, r, par ^ 1, 0, r.n_words))) return rc;
    return init_acc(ctx, r);
}

int ss_history_download(ss_ctx *ctx, int32_t rep, int32_t *out, int64_t n_rows) {
    if (!valid_rep(ctx, rep) || !out) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    ss_replica &r = ctx->reps[rep];
    if (!r.history) SS_FAIL(ctx, SS_ERR_STATE, "replica has no history property");
    if (n_rows != r.n_local) SS_FAIL(ctx, SS_ERR_INVALID, "n_rows must equal n_local");
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SS_HIP(ctx, hipMemcpy(out, r.history, sizeof(int32_t) * n_rows * SS_HISTORY_WIDTH, hipMemcpyDeviceToHost));
    return SS_OK;
}

// ---- checkpoint: replaces the pickle of Model.save / Model.load
// (model.py:1289-1310) with the replica's state, bit for bit ----
// int64 words: magic, version, n_local, n_ghost, nnz, csr_hash[2], key_hash,
// tick, tick_limit, seed, has_history; then the model, the live state buffer
// and the history rows, each padded to 8 bytes.
static const int64_t kRepMagic = 0x53534147455252ll;   // "SSAGERR"
static const int64_t kRepWords = 12;

static int64_t rep_ckpt_bytes(const ss_replica &r) {
    return kRepWords * 8 + (int64_t)((sizeof(ss_model) + 7) / 8 * 8) + (r.n_total + 7) / 8 * 8 +
           (r.history ? 4 * r.n_local * SS_HISTORY_WIDTH : 0);
}

int ss_replica_save(ss_ctx *ctx, int32_t rep, void *buf, int64_t cap, int64_t *size) {
    if (!valid_rep(ctx, rep) || !size || cap < 0 || (cap && !buf)) SS_FAIL(ctx, SS_ERR_INVALID, "bad arguments");
    const ss_replica &r = ctx->reps[rep];
    const int64_t need = rep_ckpt_bytes(r);
    *size = need;
    if (!buf) return SS_OK;
    if (cap < need) SS_FAIL(ctx, SS_ERR_INVALID, "buffer of %lld bytes, checkpoint needs %lld", (long long)cap,
                            (long long)need);
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::memset(buf, 0, (size_t)need);
    int64_t *h = static_cast<int64_t *>(buf);
    const uint64_t seed = (uint64_t)r.dev.seed_lo | ((uint64_t)r.dev.seed_hi << 32);
    const int64_t w[kRepWords] = {kRepMagic, 1, r.n_local, r.n_ghost, r.nnz, (int64_t)r.csr_hash[0],
                                  (int64_t)r.csr_hash[1], (int64_t)r.key_hash, r.tick,
                                  r.tick_limit == INT64_MAX ? 0 : r.tick_limit, (int64_t)seed, r.history ? 1 : 0};
    std::memcpy(h, w, sizeof w);
    char *c = static_cast<char *>(buf) + kRepWords * 8;
    std::memcpy(c, &r.model, sizeof(ss_model));
    c += (sizeof(ss_model) + 7) / 8 * 8;
    if (r.n_total) SS_HIP(ctx, hipMemcpy(c, r.state[r.tick & 1], (size_t)r.n_total, hipMemcpyDeviceToHost));
    c += (r.n_total + 7) / 8 * 8;
    if (r.history)
        SS_HIP(ctx, hipMemcpy(c, r.history, sizeof(int32_t) * r.n_local * SS_HISTORY_WIDTH, hipMemcpyDeviceToHost));
    return SS_OK;
}

int ss_replica_restore(ss_ctx *ctx, int32_t rep, const void *buf, int64_t size, int64_t *tick) {
    if (!valid_rep(ctx, rep) || !buf || size < kRepWords * 8) SS_FAIL(ctx, SS_ERR_INVALID, "bad arguments");
    ss_replica &r = ctx->reps[rep];
    const int64_t *h = static_cast<const int64_t *>(buf);
    if (h[0] != kRepMagic || h[1] != 1) SS_FAIL(ctx, SS_ERR_INVALID, "not a replica checkpoint");
    if (h[2] != r.n_local || h[3] != r.n_ghost || h[4] != r.nnz || h[5] != (int64_t)r.csr_hash[0] ||
        h[6] != (int64_t)r.csr_hash[1])
        SS_FAIL(ctx, SS_ERR_INVALID, "checkpoint of another network (agents, ghosts or CSR entries differ)");
    if (h[7] != (int64_t)r.key_hash) SS_FAIL(ctx, SS_ERR_INVALID, "checkpoint of other RNG key ids");
    if ((h[11] != 0) != (r.history != nullptr)) SS_FAIL(ctx, SS_ERR_INVALID, "history present in only one of them");
    if (size != rep_ckpt_bytes(r)) SS_FAIL(ctx, SS_ERR_INVALID, "checkpoint has %lld bytes, expected %lld",
                                           (long long)size, (long long)rep_ckpt_bytes(r));
    const char *c = static_cast<const char *>(buf) + kRepWords * 8;
    ss_model m;
    std::memcpy(&m, c, sizeof m);
    c += (sizeof(ss_model) + 7) / 8 * 8;
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int rc = apply_model(ctx, r, &m);
    if (rc) return rc;
    r.dev.seed_lo = (uint32_t)(uint64_t)h[10];
    r.dev.seed_hi = (uint32_t)((uint64_t)h[10] >> 32);
    r.tick = h[8];
    r.tick_limit = h[9] > 0 ? h[9] : INT64_MAX;
    if (r.history)
        SS_HIP(ctx, hipMemcpy(r.history, c + (r.n_total + 7) / 8 * 8, sizeof(int32_t) * r.n_local * SS_HISTORY_WIDTH,
                              hipMemcpyHostToDevice));
    if (tick) *tick = r.tick;
    // both state buffers, the bitmaps and the direction state of the restored tick
    return ss_state_upload(ctx, rep, reinterpret_cast<const int8_t *>(c), r.n_total);
}

int ss_set_tiling(ss_ctx *ctx, int32_t hub_degree, int32_t heavy_degree, int32_t apl, int32_t grid_blocks,
                  int32_t use_graph) {
    if (!ctx) SS_FAIL(nullptr, SS_ERR_INVALID, "ctx is NULL");
    if (apl != 1 && apl != 2 && apl != 4 && apl != 8 && apl != 16)
        SS_FAIL(ctx, SS_ERR_INVALID, "agents_per_lane must be 1, 2, 4, 8 or 16");
    bool hub_ok = false, heavy_ok = false;
    for (int32_t v : SS_THRESHOLDS) {
        hub_ok |= v == hub_degree;
        heavy_ok |= v == heavy_degree;
    }
    if (!hub_ok || !heavy_ok)
        SS_FAIL(ctx, SS_ERR_INVALID, "degree thresholds must be one of 8,16,24,32,48,64,96,128,192,256,512,"
                                     "1024,2048,4096,16384,2^30");
    if (hub_degree < heavy_degree) SS_FAIL(ctx, SS_ERR_INVALID, "hub_degree must be >= heavy_degree");
    if (grid_blocks < 0) SS_FAIL(ctx, SS_ERR_INVALID, "grid_blocks must be >= 0");
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->hub_deg = hub_degree;
    ctx->heavy_deg = heavy_degree;
    ctx->apl = apl;
    ctx->grid_blocks = grid_blocks;
    ctx->use_graph = use_graph ? 1 : 0;
    ctx->tiles_dirty = true;
    return SS_OK;
}

int ss_launch_shape(ss_ctx *ctx, int32_t *n_tiles, int32_t *grid) {
    if (!ctx) SS_FAIL(nullptr, SS_ERR_INVALID, "ctx is NULL");
    if (ctx->tiles_dirty) {
        int rc = build_tiles(ctx);
        if (rc) return rc;
    }
    if (n_tiles) *n_tiles = ctx->n_tiles;
    if (grid) *grid = ctx->grid;
    return SS_OK;
}

int ss_simulate(ss_ctx *ctx, int64_t ticks) {
    if (!ctx) SS_FAIL(nullptr, SS_ERR_INVALID, "ctx is NULL");
    if (ticks < 0) SS_FAIL(ctx, SS_ERR_INVALID, "ticks must be >= 0");
    if (ticks == 0) return SS_OK;
    (void)hipSetDevice(ctx->device);
    int rc;
    if (ctx->tiles_dirty && (rc = build_tiles(ctx))) return rc;
    if ((rc = push_descriptors(ctx, ticks))) return rc;
    if ((rc = launch_loop(ctx, ticks))) return rc;
    advance_ticks(ctx, ticks);
    return SS_OK;
}

// ---------------- ghosts ----------------

static int check_map_entries(ss_ctx *ctx, const ss_replica &r, const int64_t *ids, const int32_t *slots,
                             int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        if (ids[i] == -1 || ids[i] == -2) SS_FAIL(ctx, SS_ERR_RANGE, "ids -1 and -2 are reserved");
        if (slots && (slots[i] < 0 || slots[i] >= r.n_total))
            SS_FAIL(ctx, SS_ERR_RANGE, "slot %d outside the replica", slots[i]);
    }
    return SS_OK;
}

// device copies of a batch of map entries (freed by the destructor)
struct MapBatch {
    int64_t *ids = nullptr;
    int32_t *vals = nullptr;
    ~MapBatch() { dfree(ids); dfree(vals); }
    hipError_t stage(const int64_t *h_ids, const int32_t *h_vals, int64_t n) {
        hipError_t e = hipMalloc((void **)&ids, sizeof(int64_t) * (size_t)n);
        if (e == hipSuccess && h_vals) e = hipMalloc((void **)&vals, sizeof(int32_t) * (size_t)n);
        if (e == hipSuccess) e = hipMemcpy(ids, h_ids, sizeof(int64_t) * (size_t)n, hipMemcpyHostToDevice);
        if (e == hipSuccess && h_vals) e = hipMemcpy(vals, h_vals, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice);
        return e;
    }
};

static int alloc_map(ss_ctx *ctx, ss_replica &r, int64_t cap) {
    if (cap != r.map_cap || !r.map_keys) {
        if (r.map_keys) {
            const size_t b = (sizeof(unsigned long long) + sizeof(int32_t)) * (size_t)r.map_cap;
            r.bytes -= b;
            ctx->bytes -= b;
        }
        dfree(r.map_keys); dfree(r.map_vals);
        r.map_keys = nullptr; r.map_vals = nullptr; r.map_cap = 0;
        SS_HIP(ctx, dalloc(ctx, &r, &r.map_keys, (size_t)cap));
        SS_HIP(ctx, dalloc(ctx, &r, &r.map_vals, (size_t)cap));
        r.map_cap = cap;
    }
    SS_HIP(ctx, ss_launch_map_clear(r.map_keys, r.map_vals, cap, ctx->stream));
    r.map_size = 0;
    return SS_OK;
}

// entries of a batch in input order on one wavefront; out[0] = entries
// added or removed, out[1] = entries that found no slot
static int map_seq(ss_ctx *ctx, ss_replica &r, int op, const int64_t *ids, const int32_t *vals, int64_t n,
                   int32_t out[2]) {
    out[0] = out[1] = 0;
    if (n == 0) return SS_OK;
    MapBatch b;
    hipError_t e = b.stage(ids, vals, n);
    if (e == hipSuccess) e = hipMemsetAsync(ctx->d_counters, 0, 2 * sizeof(int32_t), ctx->stream);
    if (e == hipSuccess)
        e = ss_launch_map_seq(r.map_keys, r.map_vals, r.map_cap, op, b.ids, b.vals, n, ctx->d_counters, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(out, ctx->d_counters, 2 * sizeof(int32_t), hipMemcpyDeviceToHost);
    SS_HIP(ctx, e);
    return SS_OK;
}

int ss_map_build(ss_ctx *ctx, int32_t rep, const int64_t *ids, const int32_t *slots, int64_t n, double lf) {
    if (!valid_rep(ctx, rep) || n < 0 || (n && (!ids || !slots))) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    if (!(lf > 0.0 && lf <= 0.9)) SS_FAIL(ctx, SS_ERR_INVALID, "load_factor must be in (0, 0.9]");
    if (n >= INT32_MAX) SS_FAIL(ctx, SS_ERR_INVALID, "at most 2^31 - 2 entries");
    ss_replica &r = ctx->reps[rep];
    int rc = check_map_entries(ctx, r, ids, slots, n);
    if (rc) return rc;
    const int64_t cap = std::max<int64_t>(64, (int64_t)std::ceil((double)n / lf));
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if ((rc = alloc_map(ctx, r, cap))) return rc;
    if (n > 0) {
        MapBatch b;
        int64_t *probe = nullptr;
        int32_t *claim = nullptr;
        hipError_t e = b.stage(ids, slots, n);
        if (e == hipSuccess) e = hipMalloc((void **)&probe, sizeof(int64_t) * (size_t)n);
        if (e == hipSuccess) e = hipMalloc((void **)&claim, sizeof(int32_t) * (size_t)cap);
        if (e == hipSuccess) e = ss_launch_map_build_init(b.ids, n, cap, probe, claim, ctx->stream);
        // rounds in batches of four between reads of the count left unplaced;
        // every round places at least one entry, so n + cap rounds bound it
        const int R = 4;
        int64_t rounds = 0;
        int32_t left = 1;
        while (e == hipSuccess && left > 0 && rounds <= n + cap) {
            e = hipMemsetAsync(ctx->d_counters, 0, R * sizeof(int32_t), ctx->stream);
            for (int j = 0; j < R && e == hipSuccess; ++j)
                e = ss_launch_map_build_round(r.map_keys, r.map_vals, cap, b.ids, b.vals, n, probe, claim,
                                              ctx->d_counters + j, ctx->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(&left, ctx->d_counters + R - 1, sizeof(int32_t),
                                                    hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            rounds += R;
        }
        dfree(probe);
        dfree(claim);
        SS_HIP(ctx, e);
        if (left > 0) SS_FAIL(ctx, SS_ERR_FULL, "map build left %d entries unplaced", left);
    }
    r.map_size = n;    // gpu_kernels.py:137 counts every entry, repeated ids included
    return SS_OK;
}

int ss_map_insert(ss_ctx *ctx, int32_t rep, const int64_t *ids, const int32_t *slots, int64_t n) {
    if (!valid_rep(ctx, rep) || n < 0 || (n && (!ids || !slots))) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    ss_replica &r = ctx->reps[rep];
    if (!r.map_keys) SS_FAIL(ctx, SS_ERR_STATE, "map not built");
    int rc = check_map_entries(ctx, r, ids, slots, n);
    if (rc) return rc;
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int32_t out[2];
    if ((rc = map_seq(ctx, r, SS_MAP_INSERT, ids, slots, n, out))) return rc;
    r.map_size += out[0];
    if (out[1]) SS_FAIL(ctx, SS_ERR_FULL, "%d inserts found no free slot (capacity %lld; see ss_map_resize)",
                        out[1], (long long)r.map_cap);
    return SS_OK;
}

int ss_map_remove(ss_ctx *ctx, int32_t rep, const int64_t *ids, int64_t n, int64_t *n_removed) {
    if (!valid_rep(ctx, rep) || n < 0 || (n && !ids)) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    ss_replica &r = ctx->reps[rep];
    if (!r.map_keys) SS_FAIL(ctx, SS_ERR_STATE, "map not built");
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int32_t out[2];
    int rc = map_seq(ctx, r, SS_MAP_REMOVE, ids, nullptr, n, out);
    if (rc) return rc;
    r.map_size -= out[0];
    if (n_removed) *n_removed = out[0];
    return SS_OK;
}

int ss_map_resize(ss_ctx *ctx, int32_t rep, int64_t new_cap) {
    if (!valid_rep(ctx, rep)) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica");
    ss_replica &r = ctx->reps[rep];
    if (!r.map_keys) SS_FAIL(ctx, SS_ERR_STATE, "map not built");
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // live entries in slot order (gpu_kernels.py:232-234)
    std::vector<int64_t> keys((size_t)r.map_cap);
    std::vector<int32_t> vals((size_t)r.map_cap);
    SS_HIP(ctx, hipMemcpy(keys.data(), r.map_keys, sizeof(int64_t) * keys.size(), hipMemcpyDeviceToHost));
    SS_HIP(ctx, hipMemcpy(vals.data(), r.map_vals, sizeof(int32_t) * vals.size(), hipMemcpyDeviceToHost));
    size_t live = 0;
    for (size_t s = 0; s < keys.size(); ++s)
        if (keys[s] != -1 && keys[s] != -2) {
            keys[live] = keys[s];
            vals[live] = vals[s];
            ++live;
        }
    if (new_cap < 1 || new_cap < (int64_t)live || new_cap >= INT32_MAX)
        SS_FAIL(ctx, SS_ERR_INVALID, "capacity %lld cannot hold %zu entries", (long long)new_cap, live);
    int rc = alloc_map(ctx, r, new_cap);
    if (rc) return rc;
    int32_t out[2];
    if ((rc = map_seq(ctx, r, SS_MAP_RAW, keys.data(), vals.data(), (int64_t)live, out))) return rc;
    r.map_size = out[0];
    if (out[1]) SS_FAIL(ctx, SS_ERR_FULL, "%d entries found no slot", out[1]);
    return SS_OK;
}

int ss_map_info(ss_ctx *ctx, int32_t rep, int64_t *capacity, int64_t *size) {
    if (!valid_rep(ctx, rep)) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica");
    const ss_replica &r = ctx->reps[rep];
    if (capacity) *capacity = r.map_keys ? r.map_cap : 0;
    if (size) *size = r.map_keys ? r.map_size : 0;
    return SS_OK;
}

int ss_map_download(ss_ctx *ctx, int32_t rep, int64_t *keys, int32_t *vals, int64_t capacity) {
    if (!valid_rep(ctx, rep) || !keys || !vals) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    const ss_replica &r = ctx->reps[rep];
    if (!r.map_keys) SS_FAIL(ctx, SS_ERR_STATE, "map not built");
    if (capacity != r.map_cap) SS_FAIL(ctx, SS_ERR_INVALID, "capacity %lld, map has %lld", (long long)capacity,
                                       (long long)r.map_cap);
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SS_HIP(ctx, hipMemcpy(keys, r.map_keys, sizeof(int64_t) * (size_t)capacity, hipMemcpyDeviceToHost));
    SS_HIP(ctx, hipMemcpy(vals, r.map_vals, sizeof(int32_t) * (size_t)capacity, hipMemcpyDeviceToHost));
    return SS_OK;
}

static int lookup_dev(ss_ctx *ctx, ss_replica &r, const int64_t *ids, int32_t *d_out, int64_t n, int64_t *n_missing) {
    int64_t *d_ids = nullptr;
    SS_HIP(ctx, hipMalloc((void **)&d_ids, sizeof(int64_t) * n));
    int32_t zero = 0, miss = 0;
    hipError_t e = hipMemcpy(d_ids, ids, sizeof(int64_t) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ctx->d_counters, &zero, sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = ss_launch_map_lookup(r.map_keys, r.map_vals, r.map_cap, d_ids, d_out, n, ctx->d_counters, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(&miss, ctx->d_counters, sizeof(int32_t), hipMemcpyDeviceToHost);
    dfree(d_ids);
    SS_HIP(ctx, e);
    if (n_missing) *n_missing = miss;
    return SS_OK;
}

int ss_map_lookup(ss_ctx *ctx, int32_t rep, const int64_t *ids, int32_t *out, int64_t n, int64_t *n_missing) {
    if (!valid_rep(ctx, rep) || (n && (!ids || !out))) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    ss_replica &r = ctx->reps[rep];
    if (!r.map_keys) SS_FAIL(ctx, SS_ERR_STATE, "map not built");
    if (n == 0) { if (n_missing) *n_missing = 0; return SS_OK; }
    int32_t *d_out = nullptr;
    SS_HIP(ctx, hipMalloc((void **)&d_out, sizeof(int32_t) * n));
    int rc = lookup_dev(ctx, r, ids, d_out, n, n_missing);
    if (rc == SS_OK) {
        hipError_t e = hipMemcpy(out, d_out, sizeof(int32_t) * n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) { dfree(d_out); SS_HIP(ctx, e); }
    }
    dfree(d_out);
    return rc;
}

int ss_halo_bind(ss_ctx *ctx, int32_t rep, const int64_t *ids, int64_t n, int64_t *n_missing) {
    if (!valid_rep(ctx, rep) || (n && !ids)) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    ss_replica &r = ctx->reps[rep];
    if (!r.map_keys) SS_FAIL(ctx, SS_ERR_STATE, "map not built");
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    dfree(r.halo_slots); dfree(r.halo_rec);
    for (auto &d : r.halo_dev) { dfree(d); d = nullptr; }
    for (auto &d : r.halo_grp) { dfree(d); d = nullptr; }
    r.halo_grp_bytes = 0;
    r.halo_slots = nullptr; r.halo_rec = nullptr;
    r.n_halo = 0;
    if (n == 0) return SS_OK;
    SS_HIP(ctx, dalloc(ctx, &r, &r.halo_slots, (size_t)n));
    for (auto &d : r.halo_dev) SS_HIP(ctx, dalloc(ctx, &r, &d, (size_t)n));
    int64_t miss = 0;
    int rc = lookup_dev(ctx, r, ids, r.halo_slots, n, &miss);
    if (rc) return rc;
    // the refresh runs in slot order: record of every slot from the first
    // bound one (or the first ghost) on; a slot bound twice takes the later record
    std::vector<int32_t> slots((size_t)n);
    SS_HIP(ctx, hipMemcpy(slots.data(), r.halo_slots, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost));
    int64_t lo = r.n_local;
    for (int32_t s : slots)
        if (s >= 0) lo = std::min<int64_t>(lo, s);
    r.halo_base = lo / 64 * 64;
    // padded to whole groups of sixteen slots (the refresh reads four records
    // at once, a lane of the tick kernel's fused refresh sixteen)
    std::vector<int32_t> rec((size_t)((r.n_total - r.halo_base + 15) / 16 * 16), -1);
    for (int64_t h = 0; h < n; ++h)
        if (slots[(size_t)h] >= 0) rec[(size_t)(slots[(size_t)h] - r.halo_base)] = (int32_t)h;
    SS_HIP(ctx, dalloc(ctx, &r, &r.halo_rec, rec.size()));
    SS_HIP(ctx, hipMemcpy(r.halo_rec, rec.data(), sizeof(int32_t) * rec.size(), hipMemcpyHostToDevice));
    if (n_missing) *n_missing = miss;
    r.n_halo = n;
    return SS_OK;
}

static int refresh_ghost_bits(ss_ctx *ctx, ss_replica &r) {
    const int par = (int)(r.tick & 1);
    return rebuild_bits(ctx, r, par, r.n_local / 64, r.n_words);
}

static int halo_unpack(ss_ctx *ctx, int32_t rep, const void *vals, int64_t n, bool packed) {
    if (!valid_rep(ctx, rep) || (n && !vals)) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    ss_replica &r = ctx->reps[rep];
    if (n != r.n_halo) SS_FAIL(ctx, SS_ERR_INVALID, "n = %lld, bound halo has %lld records", (long long)n, (long long)r.n_halo);
    if (n == 0) return SS_OK;
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SS_HIP(ctx, hipMemcpy(r.halo_dev[0], vals, (size_t)(packed ? (n + 3) / 4 : n), hipMemcpyHostToDevice));
    const int par = (int)(r.tick & 1);
    SS_HIP(ctx, ss_launch_halo_gather(r.state[0], r.state[1], r.state[par], r.ibits[par], r.halo_rec, r.halo_dev[0],
                                      r.halo_base, r.n_total, r.dev.inf_code, packed, ctx->stream));
    // the words below the refreshed range that hold ghosts
    if (r.halo_base > r.n_local / 64 * 64) return rebuild_bits(ctx, r, par, r.n_local / 64, r.halo_base / 64);
    return SS_OK;
}

int ss_halo_unpack(ss_ctx *ctx, int32_t rep, const int8_t *vals, int64_t n) {
    return halo_unpack(ctx, rep, vals, n, false);
}

int ss_halo_unpack_2bit(ss_ctx *ctx, int32_t rep, const uint8_t *packed, int64_t n) {
    return halo_unpack(ctx, rep, packed, n, true);
}

int ss_halo_unpack_ids(ss_ctx *ctx, int32_t rep, const int64_t *ids, const int8_t *vals, int64_t n, int64_t *n_missing) {
    if (!valid_rep(ctx, rep) || (n && (!ids || !vals))) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    ss_replica &r = ctx->reps[rep];
    if (!r.map_keys) SS_FAIL(ctx, SS_ERR_STATE, "map not built");
    if (n == 0) { if (n_missing) *n_missing = 0; return SS_OK; }
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int64_t *d_ids = nullptr;
    int8_t *d_vals = nullptr;
    SS_HIP(ctx, hipMalloc((void **)&d_ids, sizeof(int64_t) * n));
    hipError_t e = hipMalloc((void **)&d_vals, (size_t)n);
    if (e != hipSuccess) { dfree(d_ids); SS_HIP(ctx, e); }
    int32_t zero = 0, miss = 0;
    e = hipMemcpy(d_ids, ids, sizeof(int64_t) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_vals, vals, (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ctx->d_counters, &zero, sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = ss_launch_halo_unpack_ids(r.map_keys, r.map_vals, r.map_cap, d_ids, d_vals, n, r.state[0], r.state[1], ctx->d_counters, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(&miss, ctx->d_counters, sizeof(int32_t), hipMemcpyDeviceToHost);
    dfree(d_ids);
    dfree(d_vals);
    SS_HIP(ctx, e);
    if (n_missing) *n_missing = miss;
    return refresh_ghost_bits(ctx, r);
}

static int halo_pack(ss_ctx *ctx, int32_t rep, const int32_t *slots, int64_t n, void *out, bool packed) {
    if (!valid_rep(ctx, rep) || (n && (!slots || !out))) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    ss_replica &r = ctx->reps[rep];
    for (int64_t i = 0; i < n; ++i)
        if (slots[i] < 0 || slots[i] >= r.n_total) SS_FAIL(ctx, SS_ERR_RANGE, "slot %d outside the replica", slots[i]);
    if (n == 0) return SS_OK;
    const size_t nb = (size_t)(packed ? (n + 3) / 4 : n);
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int32_t *d_slots = nullptr;
    int8_t *d_out = nullptr;
    SS_HIP(ctx, hipMalloc((void **)&d_slots, sizeof(int32_t) * n));
    hipError_t e = hipMalloc((void **)&d_out, nb);
    if (e != hipSuccess) { dfree(d_slots); SS_HIP(ctx, e); }
    e = hipMemcpy(d_slots, slots, sizeof(int32_t) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = packed ? ss_launch_gather2(r.state[r.tick & 1], d_slots, (uint8_t *)d_out, n, ctx->stream)
                   : ss_launch_gather(r.state[r.tick & 1], d_slots, d_out, n, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d_out, nb, hipMemcpyDeviceToHost);
    dfree(d_slots);
    dfree(d_out);
    SS_HIP(ctx, e);
    return SS_OK;
}

int ss_halo_pack(ss_ctx *ctx, int32_t rep, const int32_t *slots, int64_t n, int8_t *out) {
    return halo_pack(ctx, rep, slots, n, out, false);
}

int ss_halo_pack_2bit(ss_ctx *ctx, int32_t rep, const int32_t *slots, int64_t n, uint8_t *out) {
    return halo_pack(ctx, rep, slots, n, out, true);
}

// The ghost refresh of launch step k of a halo run from the staged halo
// bytes, and the bitmap words shared by local agents and ghosts.
static hipError_t halo_refresh(ss_ctx *ctx, ss_replica &r, int64_t k, const int8_t *staged, bool packed) {
    const int par = (int)((r.tick + k) & 1);
    hipError_t e = ss_launch_halo_gather(r.state[0], r.state[1], r.state[par], r.ibits[par], r.halo_rec, staged,
                                         r.halo_base, r.n_total, r.dev.inf_code, packed, ctx->stream);
    const int64_t w0 = r.n_local / 64;
    if (e == hipSuccess && r.halo_base > w0 * 64)
        e = ss_launch_bits(r.state[par], r.ibits[par], w0, r.halo_base / 64, r.n_total, r.dev.inf_code, ctx->stream);
    return e;
}

// One tick of a halo run: the refresh, then the tick.
static hipError_t halo_tick(ss_ctx *ctx, ss_replica &r, int64_t k, const int8_t *staged, bool packed) {
    hipError_t e = halo_refresh(ctx, r, k, staged, packed);
    if (e == hipSuccess) e = tick_launch(ctx, k);
    return e;
}

static int simulate_halo_groups(ss_ctx *ctx, ss_replica &r, int64_t ticks, const int8_t *vals, int64_t n,
                                bool packed, int64_t per, bool fuse);

// Halo run from page-locked host memory, in groups of at most SS_HALO_GROUP
// ticks (and SS_HALO_GROUP_BYTES of halo).  A group's halo goes to one of two
// device buffers in copies issued before any of its kernels, over spans of
// 1, 1, 2, 4, ... ticks: the first tick waits for one tick's bytes, and the
// copy engine, faster than the ticks consume, stays ahead after that.  The
// host issues a tick's kernels without waiting for anything; it waits once a
// group, for the ticks of the group before last to have read the buffer it is
// about to overwrite.  The first tick of a group has the refresh kernel; with
// 16-agent wide tiles (and no ghost slot in a hub or heavy tile), every other
// tick's refresh is fused into the tick before it: that tick writes the
// ghosts of the next buffer from the next step's halo (halo_next in the
// replica descriptor), so a group of G ticks launches one refresh, not G.
static int simulate_halo_pinned(ss_ctx *ctx, ss_replica &r, int64_t ticks, const int8_t *vals, int64_t n,
                                bool packed) {
    const int64_t gmax = std::max<int64_t>(1, std::min<int64_t>(SS_HALO_GROUP, SS_HALO_GROUP_BYTES / n));
    const int64_t want = std::min<int64_t>(gmax, ticks) * n;
    if (r.halo_grp_bytes < want) {
        SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (auto &d : r.halo_grp) { dfree(d); d = nullptr; }
        r.halo_grp_bytes = 0;
        for (auto &d : r.halo_grp) SS_HIP(ctx, dalloc(ctx, &r, &d, (size_t)want));
        r.halo_grp_bytes = want;
    }
    const int64_t per = r.halo_grp_bytes / n;         // ticks a group buffer holds
    // the caller's buffer is borrowed for this call only: every way out of it,
    // errors included, waits for the copies from it
    struct CopyWait {
        hipStream_t s;
        ~CopyWait() { (void)hipStreamSynchronize(s); }
    } copy_wait{ctx->copy_stream};
    // the call's first tick is copied before the rest of its set-up, which
    // its transfer then overlaps
    if (ticks > 0) {
        const int b0 = (int)(ctx->grp_pos & 1);
        SS_HIP(ctx, hipEventSynchronize(ctx->grp_done[b0]));
        SS_HIP(ctx, hipMemcpyAsync(r.halo_grp[b0], vals, (size_t)n, hipMemcpyHostToDevice, ctx->copy_stream));
        SS_HIP(ctx, hipEventRecord(ctx->span_ev[b0][0], ctx->copy_stream));
    }
    const bool fuse = ctx->apl == 16 && r.dev.h64 <= r.n_local && ticks > 1;
    std::vector<const uint8_t *> hp;
    if (fuse) {
        // the staged halo of every step of the call; null where a group starts
        hp.assign((size_t)ticks, nullptr);
        int64_t gp = ctx->grp_pos;
        for (int64_t k0 = 0; k0 < ticks; k0 += per, ++gp) {
            const int8_t *buf = r.halo_grp[gp & 1];
            for (int64_t j = 1; j < std::min(per, ticks - k0); ++j) hp[(size_t)(k0 + j)] = (const uint8_t *)(buf + j * n);
        }
        if (r.halo_ptrs_cap < ticks) {
            SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
            dfree((void *)r.halo_ptrs);
            r.halo_ptrs = nullptr;
            r.halo_ptrs_cap = 0;
            SS_HIP(ctx, dalloc(ctx, &r, &r.halo_ptrs, (size_t)ticks));
            r.halo_ptrs_cap = ticks;
        }
        r.dev.halo_next = r.halo_ptrs;
        r.dev.halo_n = (int32_t)ticks;
        r.dev.halo_rec = r.halo_rec;
        r.dev.halo_base = r.halo_base;
        r.dev.halo_packed = packed ? 1u : 0u;
    }
    int rc = push_descriptors(ctx, ticks);              // the compute stream is idle after it
    r.dev.halo_next = nullptr;                          // later calls upload the descriptor without them
    r.dev.halo_n = 0;
    if (rc) return rc;
    if (fuse)
        SS_HIP(ctx, hipMemcpy((void *)r.halo_ptrs, hp.data(), sizeof(const uint8_t *) * hp.size(),
                              hipMemcpyHostToDevice));
    ctx->halo_fused = fuse;
    rc = simulate_halo_groups(ctx, r, ticks, vals, n, packed, per, fuse);
    ctx->halo_fused = false;
    if (rc) return rc;
    SS_HIP(ctx, hipStreamSynchronize(ctx->copy_stream));
    advance_ticks(ctx, ticks);
    return SS_OK;
}

// The groups of a page-locked halo run (see simulate_halo_pinned).
static int simulate_halo_groups(ss_ctx *ctx, ss_replica &r, int64_t ticks, const int8_t *vals, int64_t n,
                                bool packed, int64_t per, bool fuse) {
    for (int64_t k0 = 0; k0 < ticks;) {
        const int64_t G = std::min(per, ticks - k0);
        const int b = (int)(ctx->grp_pos++ & 1);
        int8_t *buf = r.halo_grp[b];
        int64_t span_end[SS_HALO_SPANS];
        int ns = 0;
        int64_t a = 0;
        if (k0 == 0) {                                   // the first span went ahead (simulate_halo_pinned)
            span_end[ns++] = a = 1;
        } else {
            SS_HIP(ctx, hipEventSynchronize(ctx->grp_done[b]));
        }
        while (a < G) {
            const int64_t e = ns + 1 == SS_HALO_SPANS ? G : std::min(G, a == 0 ? 1 : 2 * a);
            SS_HIP(ctx, hipMemcpyAsync(buf + a * n, vals + (size_t)(k0 + a) * (size_t)n, (size_t)((e - a) * n),
                                       hipMemcpyHostToDevice, ctx->copy_stream));
            SS_HIP(ctx, hipEventRecord(ctx->span_ev[b][ns], ctx->copy_stream));
            span_end[ns++] = e;
            a = e;
        }
        // the compute stream waits for the spans holding ticks up to x, in order
        int s = 0;
        auto need = [&](int64_t x) -> hipError_t {
            hipError_t he = hipSuccess;
            while (he == hipSuccess && s < ns && (s == 0 ? 0 : span_end[s - 1]) <= x)
                he = hipStreamWaitEvent(ctx->stream, ctx->span_ev[b][s++], 0);
            return he;
        };
        for (int64_t j = 0; j < G; ++j) {
            if (j == 0 || !fuse) {
                SS_HIP(ctx, need(j));
                SS_HIP(ctx, halo_refresh(ctx, r, k0 + j, buf + j * n, packed));
            }
            if (fuse && j + 1 < G) SS_HIP(ctx, need(j + 1));   // this tick writes the next one's ghosts
            SS_HIP(ctx, tick_launch(ctx, k0 + j));
        }
        SS_HIP(ctx, hipEventRecord(ctx->grp_done[b], ctx->stream));
        k0 += G;
    }
    return SS_OK;
}

static int simulate_halo(ss_ctx *ctx, int32_t rep, int64_t ticks, const int8_t *vals, bool packed) {
    if (!valid_rep(ctx, rep) || (ticks > 0 && !vals)) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    if (ticks < 0 || ticks > INT32_MAX) SS_FAIL(ctx, SS_ERR_INVALID, "ticks out of range");
    ss_replica &r = ctx->reps[rep];
    // bytes of one tick's halo
    const int64_t n = packed ? (r.n_halo + 3) / 4 : r.n_halo;
    if (n == 0) return ss_simulate(ctx, ticks);
    int rc;
    if (ctx->tiles_dirty && (rc = build_tiles(ctx))) return rc;
    // values in page-locked memory are copied from where they are
    bool pinned = false;
    {
        hipPointerAttribute_t at;
        const hipError_t e = hipPointerGetAttributes(&at, vals);
        pinned = e == hipSuccess && at.type == hipMemoryTypeHost;
        if (e != hipSuccess) (void)hipGetLastError();   // pageable memory: not an error here
    }
    if (pinned) return simulate_halo_pinned(ctx, r, ticks, vals, n, packed);
    if ((rc = push_descriptors(ctx, ticks))) return rc;
    // pageable memory: staged tick by tick through a ring of page-locked buffers
    if (ctx->h_halo_cap < n) {
        SS_HIP(ctx, hipStreamSynchronize(ctx->copy_stream));
        for (auto &h : ctx->h_halo_pin) {
            if (h) (void)hipHostFree(h);
            h = nullptr;
        }
        ctx->h_halo_cap = 0;
        for (auto &h : ctx->h_halo_pin) SS_HIP(ctx, hipHostMalloc((void **)&h, (size_t)n, hipHostMallocDefault));
        ctx->h_halo_cap = n;
    }
    for (int64_t k = 0; k < ticks; ++k) {
        // ring slot b (device buffer, staging buffer, events) is free once the
        // tick that last read it has run: the host waits for that, so the
        // copy stream never waits on the compute stream (a copy queued behind
        // such a wait holds the calling thread until the wait resolves)
        const int b = (int)(ctx->halo_pos++ % SS_HALO_RING);
        SS_HIP(ctx, hipEventSynchronize(ctx->halo_used[b]));
        std::memcpy(ctx->h_halo_pin[b], vals + (size_t)k * (size_t)n, (size_t)n);
        // the copy overlaps the kernels of the ticks before
        SS_HIP(ctx, hipMemcpyAsync(r.halo_dev[b], ctx->h_halo_pin[b], (size_t)n, hipMemcpyHostToDevice,
                                   ctx->copy_stream));
        SS_HIP(ctx, hipEventRecord(ctx->halo_ev[b], ctx->copy_stream));
        SS_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->halo_ev[b], 0));
        SS_HIP(ctx, halo_tick(ctx, r, k, r.halo_dev[b], packed));
        SS_HIP(ctx, hipEventRecord(ctx->halo_used[b], ctx->stream));
    }
    advance_ticks(ctx, ticks);
    return SS_OK;
}

int ss_simulate_halo(ss_ctx *ctx, int32_t rep, int64_t ticks, const int8_t *vals) {
    return simulate_halo(ctx, rep, ticks, vals, false);
}

int ss_simulate_halo_2bit(ss_ctx *ctx, int32_t rep, int64_t ticks, const uint8_t *packed) {
    return simulate_halo(ctx, rep, ticks, (const int8_t *)packed, true);
}

int ss_host_alloc(ss_ctx *ctx, int64_t bytes, void **out) {
    if (!ctx || !out || bytes < 0) SS_FAIL(ctx, SS_ERR_INVALID, "bad argument");
    *out = nullptr;
    (void)hipSetDevice(ctx->device);
    SS_HIP(ctx, hipHostMalloc(out, (size_t)std::max<int64_t>(bytes, 1), hipHostMallocDefault));
    return SS_OK;
}

int ss_host_free(ss_ctx *ctx, void *p) {
    if (!ctx) SS_FAIL(nullptr, SS_ERR_INVALID, "ctx is NULL");
    if (p) SS_HIP(ctx, hipHostFree(p));
    return SS_OK;
}

// ---------------- measurement ----------------

int ss_set_counting(ss_ctx *ctx, int32_t on) {
    if (!ctx) SS_FAIL(nullptr, SS_ERR_INVALID, "ctx is NULL");
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->counting != (on != 0)) drop_graphs(ctx);  // the counter pointer is a captured argument
    ctx->counting = on != 0;
    return SS_OK;
}

int ss_read_counters(ss_ctx *ctx, int64_t *out) {
    if (!ctx || !out) SS_FAIL(ctx, SS_ERR_INVALID, "NULL argument");
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    unsigned long long h[SS_WORK_N];
    SS_HIP(ctx, hipMemcpy(h, ctx->d_work, sizeof h, hipMemcpyDeviceToHost));
    SS_HIP(ctx, hipMemset(ctx->d_work, 0, sizeof h));
    for (int i = 0; i < SS_WORK_N; ++i) out[i] = (int64_t)h[i];
    return SS_OK;
}

int ss_set_push(ss_ctx *ctx, double ratio) {
    if (!ctx) SS_FAIL(nullptr, SS_ERR_INVALID, "ctx is NULL");
    if (!(ratio >= 0.0) || ratio > 1.0e8) SS_FAIL(ctx, SS_ERR_INVALID, "ratio must be in [0, 1e8]");
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->push_num = (uint32_t)std::lround(ratio * 16.0);
    // the current tick's decision and tags follow the new rule
    for (auto &r : ctx->reps)
        if (r.live) {
            int rc = init_acc(ctx, r);
            if (rc) return rc;
        }
    return SS_OK;
}

int ss_replica_push_ok(ss_ctx *ctx, int32_t rep, int32_t *out) {
    if (!valid_rep(ctx, rep) || !out) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    *out = ctx->reps[rep].dev.push_ok ? 1 : 0;
    return SS_OK;
}

int ss_census(ss_ctx *ctx, int32_t rep, int64_t *out) {
    if (!valid_rep(ctx, rep) || !out) SS_FAIL(ctx, SS_ERR_INVALID, "bad replica or NULL");
    ss_replica &r = ctx->reps[rep];
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SS_HIP(ctx, hipMemset(ctx->d_stats, 0, 8 * sizeof(unsigned long long)));
    SS_HIP(ctx, ss_launch_stats(r.state[r.tick & 1], r.row_ptr, r.n_local, r.dev.s_code, ctx->d_stats,
                                ctx->num_cus * 8, ctx->stream));
    unsigned long long h[8];
    SS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SS_HIP(ctx, hipMemcpy(h, ctx->d_stats, sizeof h, hipMemcpyDeviceToHost));
    for (int i = 0; i < 5; ++i) out[i] = (int64_t)h[i];
    return SS_OK;
}

int ss_copy_probe(ss_ctx *ctx, int64_t bytes, int32_t repeats, float *ms) {
    if (!ctx || !ms || bytes <= 0 || bytes % 16 || repeats < 1) SS_FAIL(ctx, SS_ERR_INVALID, "bad probe arguments");
    void *src = nullptr, *dst = nullptr;
    SS_HIP(ctx, hipMalloc(&src, (size_t)bytes));
    hipError_t e = hipMalloc(&dst, (size_t)bytes);
    if (e != hipSuccess) { dfree(src); SS_HIP(ctx, e); }
    const int grid = ctx->num_cus * 8;
    e = ss_launch_fill_u32((uint32_t *)src, bytes / 4, 0x9E3779B9u, grid, ctx->stream);
    if (e == hipSuccess) e = ss_launch_copy(src, dst, bytes, grid, ctx->stream);  // warm
    if (e == hipSuccess) e = hipEventRecord(ctx->t_start, ctx->stream);
    for (int r = 0; r < repeats && e == hipSuccess; ++r) e = ss_launch_copy(src, dst, bytes, grid, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->t_stop, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->t_stop);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, ctx->t_start, ctx->t_stop);
    if (e == hipSuccess) *ms /= (float)repeats;
    dfree(src);
    dfree(dst);
    SS_HIP(ctx, e);
    return SS_OK;
}

int ss_gather_probe(ss_ctx *ctx, int64_t span_bytes, int64_t n_gathers, int32_t repeats, float *ms) {
    if (!ctx || !ms || span_bytes < 4 || span_bytes % 4 || n_gathers < 1 || repeats < 1)
        SS_FAIL(ctx, SS_ERR_INVALID, "bad probe arguments");
    const int grid = ctx->num_cus * 8;
    uint32_t *src = nullptr, *out = nullptr;
    SS_HIP(ctx, hipMalloc((void **)&src, (size_t)span_bytes));
    hipError_t e = hipMalloc((void **)&out, sizeof(uint32_t) * (size_t)grid * 256);
    if (e != hipSuccess) { dfree(src); SS_HIP(ctx, e); }
    const int64_t words = span_bytes / 4;
    e = ss_launch_fill_u32(src, words, 0x9E3779B9u, grid, ctx->stream);
    if (e == hipSuccess) e = ss_launch_gather_probe(src, words, n_gathers, out, grid, ctx->stream);   // warm
    if (e == hipSuccess) e = hipEventRecord(ctx->t_start, ctx->stream);
    for (int r = 0; r < repeats && e == hipSuccess; ++r)
        e = ss_launch_gather_probe(src, words, n_gathers, out, grid, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->t_stop, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->t_stop);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, ctx->t_start, ctx->t_stop);
    if (e == hipSuccess) *ms /= (float)repeats;
    dfree(src);
    dfree(out);
    SS_HIP(ctx, e);
    return SS_OK;
}

int ss_timer_start(ss_ctx *ctx) {
    if (!ctx) SS_FAIL(nullptr, SS_ERR_INVALID, "ctx is NULL");
    SS_HIP(ctx, hipEventRecord(ctx->t_start, ctx->stream));
    return SS_OK;
}

int ss_timer_stop(ss_ctx *ctx, float *ms) {
    if (!ctx || !ms) SS_FAIL(ctx, SS_ERR_INVALID, "NULL argument");
    SS_HIP(ctx, hipEventRecord(ctx->t_stop, ctx->stream));
    SS_HIP(ctx, hipEventSynchronize(ctx->t_stop));
    SS_HIP(ctx, hipEventElapsedTime(ms, ctx->t_start, ctx->t_stop));
    return SS_OK;
}

}  // extern "C"
