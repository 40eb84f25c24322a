// Fused per-tick step kernels for the compartment (SIR/SEIR) workloads on
// gfx950 and the kernels that set up their direction state (the bitmap,
// hash-map, halo and probe kernels around them are in aux_kernels.hip).
//
// Replaces the generated persistent `stepfunc` of the reference
// (model.py:2421-2454, tick body model.py:2302-2408) for the compartment
// step of tests/test_simulation.py:24-43 in its int8 / 1-(1-beta)^k form:
//   * one launch per tick (the kernel boundary is the grid barrier of
//     model.py:1956-1974); launches of a tick loop are captured in a hipGraph;
//   * ping-pong int8 state: tick t reads state[t&1] and writes every slot of
//     state[(t+1)&1] (carry-over), so no write-back pass (model.py:2378-2389)
//     and no host copy (model.py:1563-1570) is needed;
//   * an infectious bitmap (1 bit per slot, rebuilt by the same launch) is
//     what neighbour reads touch: 1.25 MB for 10M agents stays in L2;
//   * S agents draw first and skip the neighbour scan when even an all-
//     infectious neighbourhood could not infect them (m >= T[deg]); the scan
//     stops as soon as the count seen so far infects (T is non-decreasing).
//     Both shortcuts give the same state as the full count;
//   * push direction (symmetric CSR, no ghosts): when the infectious agents
//     have few edges compared with the number of susceptible agents, tick t
//     tags every neighbour of every agent it leaves infectious with tag(t+1)
//     in a second tag array; at tick t+1 an untagged susceptible agent has no
//     infectious neighbour, stays susceptible and needs no draw.  The rule
//     reads statistics of tick t that tick t-1 accumulated, and the decision
//     is recorded for tick t+1, so one kernel per tick carries both halves.
#include <algorithm>

#include "ss_device.h"
#include "ss_runtime.h"

namespace {

// Pointers carry the global address space so the loads are global_load, not
// flat; references to the block's LDS record carry the LDS address space.
#define SS_G __attribute__((address_space(1)))
#define SS_L __attribute__((address_space(3)))
// Entries of the threshold table cached in LDS; rows of light tiles have
// fewer neighbours than the heavy threshold, so their lookups stay in LDS.
#define SS_TLDS 64

// Everything one tile needs from its replica, as wave-uniform scalars (no
// arrays: a per-lane index into a local array would send it to scratch).
struct TickView {
    const SS_G int64_t *row_ptr;
    const SS_G int32_t *col_idx;
    const SS_G uint32_t *rng_lo;
    const SS_G uint32_t *rng_hi;
    const SS_G int8_t *cur;
    SS_G int8_t *nxt;
    const SS_G uint64_t *cb;
    SS_G uint64_t *nb;
    const SS_G uint8_t *deg8;
    SS_G uint16_t *cnt;          // infectious-neighbour counts of this tick, null when this tick pulls
    int32_t last;
    int8_t s_code, inf_code, s_target;
    uint32_t seed_lo, seed_hi;
    uint32_t t_lo, t_hi;
    int64_t n_local;
    uint8_t tag;                 // tile tag of this tick
    uint8_t push_next;           // this tick writes the next tick's counts
};

// The replica's rarely used fields, once per block in LDS (kept out of the
// scalar registers that the hot fields above occupy).
struct Cold {
    uint32_t T[SS_TLDS];         // inf_T[0 .. SS_TLDS)
    uint32_t tT[4];              // spontaneous-transition thresholds by state
    int32_t tN[4];               // ...and target states
    const uint32_t *inf_T;
    int32_t *history;
    uint16_t *pcnt;              // infectious-neighbour counts for the next tick
    uint8_t *tnext;              // tile tags for the next tick
    int64_t n_hub, h64;          // agent -> tile mapping of the replica
    int32_t hub0, heavy0, light0;          // tag slots (see ss_rep_dev)
    uint8_t ptag;                // tile tag of the next tick
    const int8_t *pcur;          // states of this tick when counts go only to susceptible agents, else null
    int8_t s_code;
    const uint8_t *hnext;        // fused halo refresh: the next step's halo, or null
    const int32_t *hrec;         // record of slot hbase + g
    int64_t hbase;
    uint32_t hpacked;            // 2-bit codes
};

typedef SS_L Cold LCold;        // the record as it sits in LDS

// Pushing to the next tick's counts: the fields, copied to registers once
// per batch of rows.
struct Pusher {
    uint16_t *pcnt;
    SS_G uint8_t *tnext;
    const SS_G int8_t *cur;      // non-null: push only to agents susceptible now
    int64_t n_hub, h64;
    int32_t hub0, heavy0, light0;
    uint8_t ptag;
    int8_t s_code;
    __device__ __forceinline__ explicit Pusher(const LCold *C)
        : pcnt(C->pcnt), tnext((SS_G uint8_t *)C->tnext), cur((const SS_G int8_t *)C->pcur), n_hub(C->n_hub),
          h64(C->h64), hub0(C->hub0), heavy0(C->heavy0), light0(C->light0), ptag(C->ptag), s_code(C->s_code) {}
    // tag slot of agent v (see ss_rep_dev)
    __device__ __forceinline__ int64_t slot_of(int64_t v) const {
        if (v < n_hub) return hub0 + v;
        if (v < h64) return heavy0 + ((v - n_hub) >> SS_HEAVY_SHIFT);
        return light0 + ((v - h64) >> SS_LTAG_SHIFT);
    }
    // one more infectious neighbour for agent c next tick; its tile runs then
    __device__ __forceinline__ void one(int32_t c) const {
        ss_count_add(pcnt, c);
        tnext[slot_of(c)] = ptag;
    }
    // entries [b, e) of col, by `width` threads from `me` on, four per thread
    // and round.  The entries of the next round are loaded before this round's
    // atomics and tag stores are issued: a wait for those loads then does not
    // wait for the atomics (memory operations count down in issue order).
    // With `cur`, the states of the round's targets are read first and only
    // susceptible targets are pushed (agents pushed to most often, hubs, are
    // infected early and would otherwise take every one of their rows'
    // atomics on one word).
    __device__ __forceinline__ void range(const SS_G int32_t *col, int64_t b, int64_t e, int me, int width) const {
        int64_t j = b + me;
        if (j >= e) return;
        int32_t c[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) c[q] = j + q * width < e ? col[j + q * width] : -1;
        for (;;) {
            const int64_t jn = j + 4 * (int64_t)width;
            int32_t cn[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) cn[q] = jn + q * width < e ? col[jn + q * width] : -1;
            bool go[4];
            if (cur) {
                int8_t st[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) st[q] = c[q] >= 0 ? cur[c[q]] : (int8_t)0;
#pragma unroll
                for (int q = 0; q < 4; ++q) go[q] = c[q] >= 0 && st[q] == s_code;
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) go[q] = c[q] >= 0;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (go[q]) one(c[q]);
            if (jn >= e) break;
            j = jn;
#pragma unroll
            for (int q = 0; q < 4; ++q) c[q] = cn[q];
        }
    }
};

// Changes of the direction statistics made by this tick, per lane: of the
// sum of degrees of infectious agents and of the number of susceptible ones.
struct Dir {
    int64_t ideg = 0;
    int32_t sus = 0;
};

// The change an agent's transition s -> ns makes to the statistics.
__device__ __forceinline__ void dir_change(Dir &d, const TickView &V, int8_t s, int8_t ns, int64_t deg) {
    if (ns == s) return;
    if (s == V.s_code) d.sus -= 1;
    if (ns == V.inf_code) d.ideg += deg;
    if (s == V.inf_code) d.ideg -= deg;
}

// Statistics slot of tick t.
__device__ __forceinline__ unsigned long long *acc_slot(const ss_rep_dev *R, int64_t t) {
    return R->acc + ((uint64_t)t % 3u) * (SS_ACC_SLOTS * 2);
}

// Statistics of tick t: (sum of degrees of infectious agents, #susceptible).
__device__ __forceinline__ void dir_totals(const ss_rep_dev *R, int64_t t, unsigned long long &d,
                                           unsigned long long &c) {
    const unsigned long long *a = acc_slot(R, t);
    d = 0;
    c = 0;
#pragma unroll
    for (int q = 0; q < SS_ACC_SLOTS; ++q) {
        d += a[2 * q];
        c += a[2 * q + 1];
    }
}

// Whether the tags for tick t + 1 are to be written while running tick t:
// the rule applied to the statistics of tick t.
__device__ __forceinline__ bool push_rule(const ss_rep_dev *R, int64_t t) {
    if (!R->push_ok || R->push_num == 0u || !R->cnt[0]) return false;
    unsigned long long d, c;
    dir_totals(R, t, d, c);
    return d * 16ull <= (unsigned long long)R->push_num * c;
}

// Push the rows of the lanes flagged in `rows` (one after another, each by
// the whole wavefront, lanes striding its entries); b, e are each lane's row.
__device__ __forceinline__ uint32_t push_rows(const TickView &V, const LCold *C, uint64_t rows, int64_t b,
                                              int64_t e, int lane) {
    uint32_t n = 0;
    if (!rows) return 0;
    const Pusher P(C);
    while (rows) {
        const int l = __ffsll((unsigned long long)rows) - 1;
        rows &= rows - 1;
        const int64_t bl = ((int64_t)__shfl((int32_t)(b >> 32), l) << 32) | (uint32_t)__shfl((int32_t)b, l);
        const int64_t el = ((int64_t)__shfl((int32_t)(e >> 32), l) << 32) | (uint32_t)__shfl((int32_t)e, l);
        P.range(V.col_idx, bl, el, lane, 64);
        n += (uint32_t)(el - bl);
    }
    return n;
}



// Work done by one lane over the whole launch, flushed to the launch counters
// when they are on: [0] susceptible agents visited, [1] draws, [2] neighbour
// scans started, [3] CSR entries read by scans, [4] counts pushed for the next
// tick, [6] agents of the tiles that ran ([5] counts replica-ticks that read
// pushed counts).
struct Work {
    uint32_t s_agents = 0, draws = 0, scans = 0, edges = 0, pushed = 0, swept = 0;
};

__device__ __forceinline__ uint32_t draw_key(const TickView V, uint32_t klo, uint32_t khi) {
    return ss_philox4x32_w0(V.t_lo, V.t_hi, klo, khi, V.seed_lo, V.seed_hi) & SS_U23_MASK;
}

__device__ __forceinline__ uint32_t draw(const TickView V, int64_t i) {
    const uint32_t hi = V.rng_hi ? V.rng_hi[i] : 0u;
    return draw_key(V, V.rng_lo[i], hi);
}

// spontaneous-transition threshold (0: none) and target of a state
__device__ __forceinline__ uint32_t trans_thr(const LCold *C, int8_t s) {
    return (uint8_t)s < 4u ? C->tT[(uint8_t)s] : 0u;
}
__device__ __forceinline__ int8_t trans_next(const LCold *C, int8_t s) {
    return (int8_t)C->tN[(uint8_t)s & 3u];
}

// T[min(k, last)], from LDS for small k
__device__ __forceinline__ uint32_t thr(const TickView V, const LCold *C, int64_t k) {
    const int32_t kk = (int32_t)(k < V.last ? k : V.last);
    return kk < SS_TLDS ? C->T[kk] : ((const SS_G uint32_t *)C->inf_T)[kk];
}

__device__ __forceinline__ int8_t transition(const TickView V, const LCold *C, int64_t i, int8_t s, Work &w) {
    const uint32_t t = trans_thr(C, s);
    if (t == 0u) return s;
    ++w.draws;
    return draw(V, i) < t ? trans_next(C, s) : s;
}

// 16-byte global loads and stores (p 16-byte aligned)
typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 ld4(const SS_G void *p) {
    const v4u32 x = *(const SS_G v4u32 *)p;
    return make_uint4(x.x, x.y, x.z, x.w);
}
__device__ __forceinline__ void st4(SS_G void *p, const uint4 v) {
    v4u32 x;
    x.x = v.x; x.y = v.y; x.z = v.z; x.w = v.w;
    *(SS_G v4u32 *)p = x;
}

// Scan of one row by one lane, eight entries per round.  The entries come
// from at most three aligned 16-byte loads (col_idx is padded to whole
// 16-byte groups) and only entries inside the row look up the bitmap, so a
// round costs the memory pipeline a few lines per lane instead of one per
// entry.  Returns whether the count of infectious neighbours reaches the
// agent's draw.
__device__ __forceinline__ bool scan_lane(const TickView V, const LCold *C, int64_t b, int64_t e, uint32_t m,
                                          int32_t &first, Work &w) {
    const SS_G int32_t *__restrict__ col = V.col_idx;
    const SS_G uint64_t *__restrict__ cb = V.cb;
    int32_t k = 0;
    for (int64_t j = b; j < e; j += 8) {
        const int64_t a = j & ~(int64_t)3;
        const int64_t lim = e < j + 8 ? e : j + 8;
        const uint4 x0 = ld4(col + a);
        const uint4 x1 = a + 4 < lim ? ld4(col + a + 4) : make_uint4(0u, 0u, 0u, 0u);
        const uint4 x2 = a + 8 < lim ? ld4(col + a + 8) : make_uint4(0u, 0u, 0u, 0u);
        // entries j .. j + 7 are words (j & 3) .. (j & 3) + 7 of x0 x1 x2,
        // selected with named values (an indexed array would go to scratch)
        const bool s2 = (j & 2) != 0, s1 = (j & 1) != 0;
        const uint32_t h0 = s2 ? x0.z : x0.x, h1 = s2 ? x0.w : x0.y, h2 = s2 ? x1.x : x0.z;
        const uint32_t h3 = s2 ? x1.y : x0.w, h4 = s2 ? x1.z : x1.x, h5 = s2 ? x1.w : x1.y;
        const uint32_t h6 = s2 ? x2.x : x1.z, h7 = s2 ? x2.y : x1.w, h8 = s2 ? x2.z : x2.x;
        int32_t c[8];
        c[0] = (int32_t)(s1 ? h1 : h0);
        c[1] = (int32_t)(s1 ? h2 : h1);
        c[2] = (int32_t)(s1 ? h3 : h2);
        c[3] = (int32_t)(s1 ? h4 : h3);
        c[4] = (int32_t)(s1 ? h5 : h4);
        c[5] = (int32_t)(s1 ? h6 : h5);
        c[6] = (int32_t)(s1 ? h7 : h6);
        c[7] = (int32_t)(s1 ? h8 : h7);
        uint64_t wd[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) wd[q] = j + q < e ? cb[c[q] >> 6] : 0ull;
        w.edges += (uint32_t)(e - j < 8 ? e - j : 8);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if ((wd[q] >> (c[q] & 63)) & 1ull) {
                if (first < 0) first = c[q];
                ++k;
                if (m < thr(V, C, k)) return true;
            }
        }
    }
    return false;
}

// Scan of one row by one wavefront, 256 entries per round.
__device__ __forceinline__ bool scan_wave(const TickView V, const LCold *C, int64_t b, int64_t e, uint32_t m,
                                          int lane, int32_t &first, Work &w) {
    const SS_G int32_t *__restrict__ col = V.col_idx;
    int32_t k = 0;
    for (int64_t base = b; base < e; base += 256) {
        int32_t c[4];
        bool h[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t j = base + q * 64 + lane;
            c[q] = j < e ? col[j] : 0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t j = base + q * 64 + lane;
            h[q] = j < e && ((V.cb[c[q] >> 6] >> (c[q] & 63)) & 1ull);
        }
        if (lane == 0) w.edges += (uint32_t)(e - base < 256 ? e - base : 256);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint64_t bal = __ballot(h[q]);
            if (bal) {
                if (first < 0) first = __shfl(c[q], __ffsll((unsigned long long)bal) - 1);
                k += __popcll(bal);
                if (m < thr(V, C, k)) return true;
            }
        }
    }
    return false;
}

// Scans of the rows of the lanes flagged in `rows` (each lane's own b, e, m),
// one after another by the whole wavefront; the flagged lanes receive inf and
// first.  Called by the whole wavefront.
__device__ __forceinline__ void scan_rows(const TickView &V, const LCold *C, uint64_t rows, int64_t b, int64_t e,
                                          uint32_t m, int lane, bool &inf, int32_t &first, Work &w) {
    while (rows) {                                   // wave-uniform
        const int l = __ffsll((unsigned long long)rows) - 1;
        rows &= rows - 1;
        const int64_t bl = ((int64_t)__shfl((int32_t)(b >> 32), l) << 32) | (uint32_t)__shfl((int32_t)b, l);
        const int64_t el = ((int64_t)__shfl((int32_t)(e >> 32), l) << 32) | (uint32_t)__shfl((int32_t)e, l);
        const uint32_t ml = (uint32_t)__shfl((int32_t)m, l);
        int32_t fl = -1;
        const bool il = scan_wave(V, C, bl, el, ml, lane, fl, w);
        if (lane == l) {
            inf = il;
            first = fl;
        }
    }
}

// Neighbour scans of a light tile, gathered in LDS so that the lanes of a
// wavefront scan concurrently instead of each lane walking its own agents in
// turn.  Each wavefront owns the queue section of its own agents.
#define SS_LIGHT_MAX (SS_BLOCK * 8)
struct ScanQueue {
    uint32_t m[SS_LIGHT_MAX];      // the agent's draw
    uint16_t slot[SS_LIGHT_MAX];   // agent - tile begin
    uint8_t hit[SS_LIGHT_MAX];     // scan result by slot
};

// The value, hidden from the optimiser: lane-derived values computed from it
// inside a tile function are not hoisted out of the tile loop, where they
// would stay live (or be spilled) across every kind of tile.
__device__ __forceinline__ int opaque(int x) {
    asm volatile("" : "+v"(x));
    return x;
}

// Orders a wavefront's LDS writes before its later LDS reads by other lanes
// (LDS operations of one wavefront complete in order).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LIGHT tile: A agents per lane (agents g0 + a*64 + lane).  Round trip 1
// reads the states (and, on a push tick, the infectious-neighbour counts,
// which are cleared as they are read), round trip 2 the degree and RNG key of
// the agents that draw.  Pull tick: rows whose draw passes the full-degree
// bound go to the block's scan queue, where each thread takes one row
// (row_ptr, then its entries, then their bitmap words).  Push tick: the count
// decides at once; only the last-infector history still scans (the queue
// entry is flagged SS_QHIST).  Every wave reaches the block barriers, whether
// or not it has agents.
#define SS_QHIST 0x8000u
template <int A>
__device__ __forceinline__ void light_tile(const TickView V, const LCold *C, int64_t begin, int64_t end,
                                           int32_t tag_lo, int32_t tag_hi, Work &w, Dir &dir, ScanQueue &Q) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t g0 = begin + (int64_t)wv * 64 * A;
    const int sec = wv * 64 * A;                     // this wavefront's queue section
    const uint64_t below = (1ull << lane) - 1ull;    // lanes before this one
    int nq = 0;                                      // wave-uniform queue length
    int8_t s[A], ns[A];
    uint32_t cand = 0u;                              // bit a: agent a is a susceptible candidate
    uint32_t key[A], khi[A], kc[A];                  // kc: infectious neighbours (push tick)
    if (V.cnt) {
#pragma unroll
        for (int a = 0; a < A; ++a) {
            const int64_t i = g0 + a * 64 + lane;
            s[a] = i < end ? V.cur[i] : (int8_t)0;
            kc[a] = i < end && i < V.n_local ? V.cnt[i] : 0u;
        }
#pragma unroll
        for (int a = 0; a < A; ++a) {
            const int64_t i = g0 + a * 64 + lane;
            if (kc[a] != 0u) V.cnt[i] = 0;            // ready for the tick after next
            cand |= (s[a] == V.s_code && kc[a] != 0u ? 1u : 0u) << a;
        }
    } else {
#pragma unroll
        for (int a = 0; a < A; ++a) {
            const int64_t i = g0 + a * 64 + lane;
            s[a] = i < end ? V.cur[i] : (int8_t)0;
            kc[a] = 0u;
            cand |= (s[a] == V.s_code ? 1u : 0u) << a;
        }
    }
    uint32_t deg[A];
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const int64_t i = g0 + a * 64 + lane;
        deg[a] = key[a] = khi[a] = 0u;
        if (i < end && i < V.n_local) {
            const bool sc = (cand >> a) & 1u;
            const bool tr = trans_thr(C, s[a]) != 0u;
            if (sc || tr || s[a] == V.inf_code) deg[a] = V.deg8[i];
            if (sc || tr) {
                key[a] = V.rng_lo[i];
                if (V.rng_hi) khi[a] = V.rng_hi[i];
            }
        }
    }
    uint32_t scanned = 0;                            // bit a: agent a went to the queue (pull)
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const int64_t i = g0 + a * 64 + lane;
        ns[a] = s[a];
        bool enq = false, hist = false;
        uint32_t m = 0u;
        if (i < end && i < V.n_local) {
            if (s[a] == V.s_code) {
                ++w.s_agents;
                if (deg[a] > 0 && ((cand >> a) & 1u)) {
                    ++w.draws;
                    m = draw_key(V, key[a], khi[a]);
                    if (V.cnt) {
                        if (m < thr(V, C, (int64_t)kc[a])) {
                            ns[a] = V.s_target;
                            dir_change(dir, V, s[a], ns[a], deg[a] == 255u ? (int64_t)(V.row_ptr[i + 1] - V.row_ptr[i])
                                                                             : (int64_t)deg[a]);
                            // the history's first infectious neighbour (m = 0 stops the scan there)
                            enq = hist = C->history != nullptr;
                            m = 0u;
                        }
                    } else if (m < thr(V, C, deg[a] == 255u ? (int64_t)V.last : (int64_t)deg[a])) {
                        // a capped degree (255) bounds with the last table entry
                        enq = true;
                        scanned |= 1u << a;
                        ++w.scans;
                    }
                }
            } else {
                const uint32_t t = trans_thr(C, s[a]);
                if (t != 0u) {
                    ++w.draws;
                    if (draw_key(V, key[a], khi[a]) < t) ns[a] = trans_next(C, s[a]);
                }
                dir_change(dir, V, s[a], ns[a], deg[a] == 255u ? (int64_t)(V.row_ptr[i + 1] - V.row_ptr[i])
                                                               : (int64_t)deg[a]);
            }
        }
        const uint64_t qm = __ballot(enq);
        if (enq) {
            const int q = sec + nq + __popcll(qm & below);
            Q.m[q] = m;
            Q.slot[q] = (uint16_t)((uint32_t)(i - begin) | (hist ? SS_QHIST : 0u));
        }
        nq += __popcll(qm);
    }
    if (nq > 0) {                                    // wave-uniform
        wave_sync();
        for (int q = lane; q < nq; q += 64) {
            int32_t first = -1;
            const uint32_t sl = Q.slot[sec + q];
            const bool hist_only = (sl & SS_QHIST) != 0u;
            const int64_t i = begin + (sl & ~SS_QHIST);
            const int64_t b = V.row_ptr[i], e = V.row_ptr[i + 1];
            // with m = 0 the scan stops at the first infectious neighbour
            const bool inf = scan_lane(V, C, b, e, Q.m[sec + q], first, w);
            if (!hist_only) Q.hit[sl] = inf ? 1 : 0;
            if (inf && C->history) ss_hist_append(C->history + i * SS_HIST_W, first);
            if (inf && !hist_only) dir_change(dir, V, V.s_code, V.s_target, e - b);
        }
        wave_sync();
    }
    uint32_t pmask = 0u;                             // bit a: agent a is infectious next tick
    bool busy = false;                               // the tile must run next tick
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const int64_t ga = g0 + a * 64;
        if (ga >= end) break;                        // wave-uniform; no barrier follows
        const int64_t i = ga + lane;
        if (((scanned >> a) & 1u) && Q.hit[i - begin]) ns[a] = V.s_target;
        if (i < end) V.nxt[i] = ns[a];
        const bool inf = i < end && ns[a] == V.inf_code;
        const uint64_t word = __ballot(inf);
        if (lane == 0) V.nb[ga >> 6] = word;
        pmask |= (inf && i < V.n_local ? 1u : 0u) << a;
        busy |= i < end && i < V.n_local && (ns[a] != s[a] || ns[a] == V.inf_code || trans_thr(C, ns[a]) != 0u);
    }
    // next tick's tile tag: a state changed (the other buffer is stale), an
    // agent will draw or is infectious (its row is pushed every tick);
    // candidates get theirs from the rows pushed below
    if (V.push_next && __ballot(busy) && lane == 0)
        for (int64_t q = tag_lo; q <= tag_hi; ++q) ((SS_G uint8_t *)C->tnext)[q] = C->ptag;
    if (V.push_next && __ballot(pmask != 0u)) {
#pragma unroll 1
        for (int a = 0; a < A; ++a) {
            const uint64_t rows = __ballot((pmask >> a) & 1u);
            if (!rows) continue;
            const int64_t i = g0 + a * 64 + lane;
            int64_t b = 0, e = 0;
            if ((pmask >> a) & 1u) {
                b = V.row_ptr[i];
                e = V.row_ptr[i + 1];
            }
            const uint32_t n = push_rows(V, C, rows, b, e, lane);
            if (lane == 0) w.pushed += n;
        }
    }
}

// Bits of agents in the heavy region are set and cleared one agent at a time
// (their words are shared by tiles that run concurrently).
__device__ __forceinline__ void put_bits(SS_G uint64_t *nb, int64_t word, uint64_t set, uint64_t clr) {
    if (set) atomicOr((unsigned long long *)&nb[word], (unsigned long long)set);
    if (clr) atomicAnd((unsigned long long *)&nb[word], ~(unsigned long long)clr);
}

// Wide LIGHT tile: each wavefront owns a chunk of up to 1024 consecutive agents
// and lane l the 16 agents [16 l, 16 l + 16) of it, so states, counts, degrees and
// RNG keys arrive as 16-byte vector loads, one round trip for the chunk.  The
// agents that draw (susceptible candidates, agents with a spontaneous
// transition) are packed into the wavefront's LDS queue, so the draws run on
// full wavefronts whatever the mix of states; the rows that must be scanned
// are packed again (in place) and scanned one per lane, long rows by the
// whole wavefront.  Results return through a byte per chunk slot.
#define SS_WIDE_A 16
#define SS_WIDE_CHUNK (64 * SS_WIDE_A)
#define SS_WIDE_SHORT 512        // rows longer than this are scanned by the whole wavefront
struct alignas(16) WideQueue {
    // draw entry: x = RNG key (pull) or infectious-neighbour count (push);
    //   y = slot | transition << 10 | (state & 3) << 11 | deg8 << 16 (deg8 only on pull)
    // scan entry: x = draw; y = slot | history-only << 10 | deg8 << 16
    uint2 e[SS_BLOCK / 64][SS_WIDE_CHUNK];
    uint8_t res[SS_BLOCK / 64][SS_WIDE_CHUNK];   // next state by chunk slot (queued agents only)
};

// 0xFF in byte b of the result for every set bit b of n (n < 16).
__device__ __forceinline__ uint32_t byte_mask4(uint32_t n) {
    return ((n * 0x00204081u) & 0x01010101u) * 0xFFu;
}

// Bit b of the result set where byte b of x equals the byte whose copies fill y.
__device__ __forceinline__ uint32_t eq_bytes4(uint32_t x, uint32_t y) {
    const uint32_t t = x ^ y;
    const uint32_t z = ~(((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t) & 0x80808080u;   // 0x80 where a byte of t is 0
    return (((z >> 7) * 0x00204081u) >> 21) & 0xFu;
}

__device__ __forceinline__ uint32_t u4_word(const uint4 v, int q) {
    return q == 0 ? v.x : q == 1 ? v.y : q == 2 ? v.z : v.w;
}

// Degree of local agent i from its 8-bit copy (255: read the row bounds).
__device__ __forceinline__ int64_t deg_of(const TickView &V, int64_t i, uint32_t d8) {
    return d8 == 255u ? V.row_ptr[i + 1] - V.row_ptr[i] : (int64_t)d8;
}

template <bool HALO>
__device__ __forceinline__ void light_wide(const TickView V, const LCold *C, int64_t begin, int64_t end, Work &w,
                                           Dir &dir, WideQueue &Q) {
    const int lane = opaque((int)threadIdx.x & 63), wv = threadIdx.x >> 6;
    // the wavefront's chunk: a quarter of the tile rounded up to 64 agents (at
    // most 1024), or to 16 in tiles of fewer than 256 agents (long rows), whose
    // wavefronts then share bitmap words
    const int64_t len4 = (end - begin + 3) / 4;
    const int64_t chunk = end - begin < 256 ? (len4 + 15) & ~(int64_t)15 : (len4 + 63) & ~(int64_t)63;
    const int64_t c0 = begin + (int64_t)wv * chunk;
    if (c0 >= end) return;                                        // wave-uniform: no block barrier below
    if (end > c0 + chunk) end = c0 + chunk;
    uint2 *const E = Q.e[wv];
    uint8_t *const res = Q.res[wv];
    const uint64_t below = (1ull << lane) - 1ull;
    const int64_t i0 = c0 + 16 * lane;
    const int nv = (int)(end - i0 < 0 ? 0 : end - i0 > 16 ? 16 : end - i0);           // agents of the lane
    const int nl = (int)(V.n_local - i0 < 0 ? 0 : V.n_local - i0 > nv ? nv : V.n_local - i0);   // local rows
    const uint32_t vmask = (1u << nv) - 1u, lmask = (1u << nl) - 1u;
    const bool pull = V.cnt == nullptr;
    // round trip 1: states; counts (push) or RNG keys and degrees (pull).  The
    // arrays are padded to whole 16-agent groups.
    uint4 sv = make_uint4(0u, 0u, 0u, 0u);
    if (nv > 0) sv = ld4(V.cur + i0);
    uint32_t a[16];
    uint32_t dg[4];
#pragma unroll
    for (int q = 0; q < 16; ++q) a[q] = 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) dg[q] = 0u;
    if (nl > 0) {
        if (!pull) {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const uint4 x = ld4(V.cnt + i0 + 8 * q);
                const uint32_t wd[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                for (int h = 0; h < 8; ++h) a[8 * q + h] = (wd[h >> 1] >> (16 * (h & 1))) & 0xFFFFu;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint4 x = ld4(V.rng_lo + i0 + 4 * q);
                a[4 * q] = x.x; a[4 * q + 1] = x.y; a[4 * q + 2] = x.z; a[4 * q + 3] = x.w;
            }
            {
                const uint4 x = ld4(V.deg8 + i0);
                dg[0] = x.x; dg[1] = x.y; dg[2] = x.z; dg[3] = x.w;
            }
        }
    }
    if (!pull && nl > 0) {
        // counts are cleared as they are read, ready for the tick after next
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            uint32_t any = 0u;
#pragma unroll
            for (int h = 0; h < 8; ++h) any |= a[8 * q + h];
            if (any != 0u) st4(V.cnt + i0 + 8 * q, make_uint4(0u, 0u, 0u, (uint32_t)opaque(0)));
        }
    }
    // states with a spontaneous transition, as a bit set over state codes 0..3
    uint32_t tbits = 0u;
#pragma unroll
    for (int s = 0; s < 4; ++s) tbits |= (C->tT[s] != 0u ? 1u : 0u) << s;
    uint32_t smask = 0u, tmask = 0u, cmask = 0u;     // susceptible, transition, candidate
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint32_t s = (u4_word(sv, j >> 2) >> (8 * (j & 3))) & 0xFFu;
        const bool is_s = (int8_t)s == V.s_code;
        smask |= (is_s ? 1u : 0u) << j;
        tmask |= (!is_s && s < 4u && ((tbits >> s) & 1u) ? 1u : 0u) << j;
        const uint32_t nz = pull ? ((dg[j >> 2] >> (8 * (j & 3))) & 0xFFu) : a[j];
        cmask |= (nz != 0u ? 1u : 0u) << j;
    }
    smask &= lmask;
    tmask &= lmask;
    const uint32_t enq = (smask & cmask) | tmask;
    w.s_agents += (uint32_t)__popc(smask);
    // the lane's queue position: exclusive prefix over lanes of popc(enq) (<= 16)
    const uint32_t ne = (uint32_t)__popc(enq);
    uint32_t pos = 0u, nq = 0u;
#pragma unroll
    for (int b = 0; b < 5; ++b) {
        const uint64_t B = __ballot((ne >> b) & 1u);
        pos += (uint32_t)__popcll(B & below) << b;
        nq += (uint32_t)__popcll(B) << b;
    }
    if (enq) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            if ((enq >> j) & 1u) {
                const uint32_t s = (u4_word(sv, j >> 2) >> (8 * (j & 3))) & 3u;
                const uint32_t d = pull ? ((dg[j >> 2] >> (8 * (j & 3))) & 0xFFu) : 0u;
                E[pos] = make_uint2(a[j], (uint32_t)(16 * lane + j) | ((tmask >> j) & 1u) << 10 | s << 11 | d << 16);
                ++pos;
            }
        }
    }
    bool busy = false;                               // the tile must run next tick
    uint32_t nsq = 0u;                               // wave-uniform scan-queue length
    if (nq > 0u) {                                   // wave-uniform
        wave_sync();
        // draws, a full wavefront of queued agents per round; the rows to scan
        // are packed over the entries already read
        for (uint32_t r0 = 0; r0 < nq; r0 += 64) {
            const uint32_t q = r0 + (uint32_t)lane;
            bool sc = false;
            uint2 se = make_uint2(0u, 0u);
            if (q < nq) {
                const uint2 en = E[q];
                const uint32_t slot = en.y & 1023u;
                const bool trans = ((en.y >> 10) & 1u) != 0u;
                const int64_t i = c0 + slot;
                const uint32_t key = pull ? en.x : V.rng_lo[i];
                const uint32_t khi = V.rng_hi ? V.rng_hi[i] : 0u;
                ++w.draws;
                const uint32_t m = draw_key(V, key, khi);
                int8_t s, ns;
                if (trans) {
                    s = (int8_t)((en.y >> 11) & 3u);
                    ns = m < C->tT[s] ? trans_next(C, s) : s;
                } else {
                    s = ns = V.s_code;
                    if (!pull) {
                        if (m < thr(V, C, (int64_t)en.x)) {
                            ns = V.s_target;
                            if (C->history) {
                                // the history's first infectious neighbour (draw 0 stops the scan there)
                                sc = true;
                                se = make_uint2(0u, slot | 1u << 10 | (uint32_t)V.deg8[i] << 16);
                            }
                        }
                    } else {
                        const uint32_t d = en.y >> 16;
                        // a capped degree (255) bounds with the last table entry
                        if (m < thr(V, C, d == 255u ? (int64_t)V.last : (int64_t)d)) {
                            sc = true;
                            se = make_uint2(m, slot | d << 16);
                            ++w.scans;
                        }
                    }
                }
                if (ns != s)
                    dir_change(dir, V, s, ns, deg_of(V, i, pull ? en.y >> 16 : (uint32_t)V.deg8[i]));
                busy |= ns != s || ns == V.inf_code || trans_thr(C, ns) != 0u;
                res[slot] = (uint8_t)ns;
            }
            const uint64_t sm = __ballot(sc);
            if (sc) E[nsq + (uint32_t)__popcll(sm & below)] = se;
            nsq += (uint32_t)__popcll(sm);
        }
    }
    if (nsq > 0u) {                                  // wave-uniform
        wave_sync();
        for (uint32_t r0 = 0; r0 < nsq; r0 += 64) {
            const uint32_t q = r0 + (uint32_t)lane;
            bool lng = false;
            int64_t b = 0, e = 0;
            uint32_t m = 0u, y = 0u;
            if (q < nsq) {
                const uint2 en = E[q];
                m = en.x;
                y = en.y;
                const int64_t i = c0 + (y & 1023u);
                const uint32_t d = y >> 16;
                b = V.row_ptr[i];
                e = d == 255u ? V.row_ptr[i + 1] : b + (int64_t)d;
                lng = e - b > SS_WIDE_SHORT;
            }
            bool inf = false;
            int32_t first = -1;
            if (q < nsq && !lng) inf = scan_lane(V, C, b, e, m, first, w);
            scan_rows(V, C, __ballot(lng), b, e, m, lane, inf, first, w);
            if (q < nsq && inf) {
                const uint32_t slot = y & 1023u;
                const int64_t i = c0 + slot;
                if (((y >> 10) & 1u) == 0u) {
                    res[slot] = (uint8_t)V.s_target;
                    dir_change(dir, V, V.s_code, V.s_target, e - b);
                    busy = true;
                }
                if (C->history) ss_hist_append(C->history + i * SS_HIST_W, first);
            }
        }
    }
    // next states: the queued agents' from LDS, the others unchanged
    uint4 nsv = sv;
    if (nq > 0u) {
        wave_sync();
        if (enq) {
            const uint4 r = *(const uint4 *)(res + 16 * lane);
            const uint32_t m0 = byte_mask4(enq & 15u), m1 = byte_mask4((enq >> 4) & 15u);
            const uint32_t m2 = byte_mask4((enq >> 8) & 15u), m3 = byte_mask4(enq >> 12);
            nsv.x = (r.x & m0) | (sv.x & ~m0);
            nsv.y = (r.y & m1) | (sv.y & ~m1);
            nsv.z = (r.z & m2) | (sv.z & ~m2);
            nsv.w = (r.w & m3) | (sv.w & ~m3);
        }
    }
    // fused halo refresh: the lane's ghost slots take the next step's halo
    // (what the refresh kernel would write before that step); a lane with
    // ghosts starts at or after hbase (a multiple of 64, at most n_local).
    // Compiled only into the kernels that halo runs launch.
    if (HALO && nl < nv && C->hnext != nullptr) {
        const SS_G int32_t *rr = (const SS_G int32_t *)C->hrec + (i0 - C->hbase);
        const SS_G uint8_t *hv = (const SS_G uint8_t *)C->hnext;
        const bool pk = C->hpacked != 0u;
        int32_t rc[16];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint4 x = ld4(rr + 4 * q);
            rc[4 * q] = (int32_t)x.x; rc[4 * q + 1] = (int32_t)x.y; rc[4 * q + 2] = (int32_t)x.z; rc[4 * q + 3] = (int32_t)x.w;
        }
        uint32_t gv[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            gv[j] = 0u;
            if (j >= nl && j < nv && rc[j] >= 0) gv[j] = hv[pk ? (rc[j] >> 2) : rc[j]];
        }
        uint32_t wv[4] = {nsv.x, nsv.y, nsv.z, nsv.w};
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            if (j >= nl && j < nv && rc[j] >= 0) {
                const uint32_t c = pk ? (gv[j] >> (2 * (rc[j] & 3))) & 3u : gv[j] & 0xFFu;
                wv[j >> 2] = (wv[j >> 2] & ~(0xFFu << (8 * (j & 3)))) | c << (8 * (j & 3));
            }
        }
        nsv = make_uint4(wv[0], wv[1], wv[2], wv[3]);
    }
    if (nv == 16) {
        st4(V.nxt + i0, nsv);
    } else if (nv > 0) {
#pragma unroll 1
        for (int j = 0; j < nv; ++j) V.nxt[i0 + j] = (int8_t)((u4_word(nsv, j >> 2) >> (8 * (j & 3))) & 0xFFu);
    }
    // infectious bits of the lane's agents; lanes 4g .. 4g+3 fill bitmap word g
    // of a chunk of whole words, otherwise each lane sets and clears its own 16
    // bits of a word it shares (chunk is wave-uniform)
    const uint32_t X = 0x01010101u * (uint32_t)(uint8_t)V.inf_code;
    const uint32_t ib = (eq_bytes4(nsv.x, X) | eq_bytes4(nsv.y, X) << 4 | eq_bytes4(nsv.z, X) << 8 |
                         eq_bytes4(nsv.w, X) << 12) & vmask;
    if (chunk & 63) {
        if (nv > 0) {
            const int sh = (int)(i0 & 63);
            put_bits(V.nb, i0 >> 6, (uint64_t)ib << sh, (uint64_t)(vmask & ~ib) << sh);
        }
    } else {
        uint32_t lo = (lane & 2) ? 0u : ib << (16 * (lane & 1));
        uint32_t hi = (lane & 2) ? ib << (16 * (lane & 1)) : 0u;
        lo |= (uint32_t)__shfl_xor((int32_t)lo, 1);
        hi |= (uint32_t)__shfl_xor((int32_t)hi, 1);
        lo |= (uint32_t)__shfl_xor((int32_t)lo, 2);
        hi |= (uint32_t)__shfl_xor((int32_t)hi, 2);
        if ((lane & 3) == 0 && nv > 0) V.nb[i0 >> 6] = (uint64_t)hi << 32 | lo;
    }
    if (V.push_next) {
        uint32_t pend = ib & lmask;                  // infectious local agents: their rows are pushed
        // next tick's tile tag: a state changed (the other buffer is stale), an
        // agent will draw or is infectious (its row is pushed every tick);
        // candidates get theirs from the rows pushed below
        if (__ballot(busy || pend != 0u) && lane == 0) {
            const Pusher P(C);
            for (int64_t q = P.slot_of(c0); q <= P.slot_of(end - 1); ++q) P.tnext[q] = P.ptag;
        }
        while (__ballot(pend != 0u)) {
            const bool has = pend != 0u;
            int64_t b = 0, e = 0;
            if (has) {
                const int64_t i = i0 + (__ffs(pend) - 1);
                pend &= pend - 1u;
                b = V.row_ptr[i];
                e = V.row_ptr[i + 1];
            }
            const uint32_t n = push_rows(V, C, __ballot(has), b, e, lane);
            if (lane == 0) w.pushed += n;
        }
    }
}


// HEAVY tile: SS_HPW agents per wavefront.  Lanes 0..SS_HPW-1 load, draw and
// decide for their agent in parallel; the rows that need a scan are then
// scanned one after another by the whole wavefront.
#define SS_HPW 4
__device__ __forceinline__ void heavy_tile(const TickView V, const LCold *C, int64_t begin, int64_t end,
                                           int64_t tix, Work &w, Dir &dir) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t i = begin + (int64_t)wv * SS_HPW + lane;
    const bool mine = lane < SS_HPW && i < end;
    int8_t s = 0, ns = 0;
    int64_t b = 0, e = 0;
    uint32_t m = 0u;
    bool need = false;
    if (mine) {
        s = V.cur[i];
        ns = s;
        if (i < V.n_local) {
            b = V.row_ptr[i];
            e = V.row_ptr[i + 1];
            uint32_t kc = 1u;                        // pull: every susceptible agent is a candidate
            if (V.cnt) {
                kc = V.cnt[i];
                if (kc != 0u) V.cnt[i] = 0;
            }
            if (s == V.s_code) {
                ++w.s_agents;
                if (e > b && kc != 0u) {
                    ++w.draws;
                    m = draw(V, i);
                    if (V.cnt) {
                        if (m < thr(V, C, (int64_t)kc)) {
                            ns = V.s_target;
                            need = C->history != nullptr;     // scan only for the first infectious neighbour
                            m = 0u;
                        }
                    } else {
                        need = m < thr(V, C, e - b);
                        w.scans += need;
                    }
                }
            } else {
                ns = transition(V, C, i, s, w);
            }
        }
    }
    const uint64_t pend = __ballot(need);
    if (pend) {                                      // wave-uniform
        bool inf = false;
        int32_t first = -1;
        scan_rows(V, C, pend, b, e, m, lane, inf, first, w);
        if (need && inf) {
            ns = V.s_target;
            if (C->history) ss_hist_append(C->history + i * SS_HIST_W, first);
        }
    }
    if (mine) V.nxt[i] = ns;
    if (mine && i < V.n_local) dir_change(dir, V, s, ns, e - b);
    if (V.push_next) {
        if (__ballot(mine && i < V.n_local && (ns != s || ns == V.inf_code || trans_thr(C, ns) != 0u)) && lane == 0)
            ((SS_G uint8_t *)C->tnext)[tix] = C->ptag;
        // rows of the wavefront's infectious agents, one after another, lanes striding
        const uint32_t n = push_rows(V, C, __ballot(mine && i < V.n_local && ns == V.inf_code), b, e, lane);
        if (lane == 0) w.pushed += n;
    }
    const int64_t w0 = begin + (int64_t)wv * SS_HPW;
    if (w0 < end) {
        // the SS_HPW agents of a wavefront sit in one bitmap word
        const uint64_t on = __ballot(mine && ns == V.inf_code), off = __ballot(mine && ns != V.inf_code);
        if (lane == 0) put_bits(V.nb, w0 >> 6, on << (w0 & 63), off << (w0 & 63));
    }
}

// HUB tile: one agent, the whole block.  All 256 lanes count infectious
// neighbours over rounds of 1024 entries; the round's total decides.
__device__ __forceinline__ void hub_tile(const TickView V, const LCold *C, int64_t i, int64_t tix, Work &w,
                                         Dir &dir, int *s_cnt, long long *s_pos) {
    // (s_cnt also carries the final state to the whole block for the push below)
    const int tid = threadIdx.x, lane = tid & 63;
    const int8_t s = V.cur[i];
    int8_t ns = s;
    int32_t first = -1;
    int64_t b = 0, e = 0;
    uint32_t kc = 1u;                                // pull: every susceptible agent is a candidate
    if (V.cnt && i < V.n_local) {                    // block-uniform
        // one thread reads and clears the count; the block agrees on it
        if (tid == 0) {
            const uint32_t c = V.cnt[i];
            if (c != 0u) V.cnt[i] = 0;
            *s_cnt = (int)c;
        }
        __syncthreads();
        kc = (uint32_t)*s_cnt;
        __syncthreads();
    }
    if (i < V.n_local) {
        b = V.row_ptr[i];
        e = V.row_ptr[i + 1];
        if (s == V.s_code) {
            if (tid == 0) ++w.s_agents;
            if (e > b && kc != 0u) {
                if (tid == 0) ++w.draws;
                const uint32_t m0 = draw(V, i);
                // push tick: the count decides; the scan below (m = 0) only
                // looks for the first infectious neighbour of the history
                const bool by_count = V.cnt != nullptr;
                const bool hit_now = by_count && m0 < thr(V, C, (int64_t)kc);
                if (hit_now) ns = V.s_target;
                const uint32_t m = by_count ? 0u : m0;
                if (by_count ? (hit_now && C->history) : m < thr(V, C, e - b)) {
                    if (tid == 0) ++w.scans;
                    int32_t k = 0;
                    for (int64_t base = b; base < e; base += 4 * SS_BLOCK) {
                        if (tid == 0) { *s_cnt = 0; *s_pos = INT64_MAX; }
                        __syncthreads();
                        int32_t c[4];
                        bool h[4];
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int64_t j = base + q * SS_BLOCK + tid;
                            c[q] = j < e ? V.col_idx[j] : 0;
                        }
                        int cnt = 0;
                        long long pos = INT64_MAX;
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int64_t j = base + q * SS_BLOCK + tid;
                            h[q] = j < e && ((V.cb[c[q] >> 6] >> (c[q] & 63)) & 1ull);
                            if (h[q]) {
                                ++cnt;
                                if (j < pos) pos = j;
                            }
                        }
                        const uint64_t any = __ballot(cnt > 0);
                        if (any) {
#pragma unroll
                            for (int off = 32; off > 0; off >>= 1) {
                                cnt += __shfl_down(cnt, off);
                                const long long o = __shfl_down(pos, off);
                                pos = o < pos ? o : pos;
                            }
                            if (lane == 0) {
                                atomicAdd(s_cnt, cnt);
                                atomicMin((unsigned long long *)s_pos, (unsigned long long)pos);
                            }
                        }
                        if (tid == 0) w.edges += (uint32_t)(e - base < 4 * SS_BLOCK ? e - base : 4 * SS_BLOCK);
                        __syncthreads();
                        const int32_t got = *s_cnt;
                        const long long p = *s_pos;
                        __syncthreads();
                        if (got > 0) {
                            if (first < 0) first = V.col_idx[p];
                            k += got;
                            if (m < thr(V, C, k)) {
                                ns = V.s_target;
                                break;
                            }
                        }
                    }
                }
            }
        } else if (tid == 0) {
            ns = transition(V, C, i, s, w);
        }
    }
    if (tid == 0) {
        V.nxt[i] = ns;
        if (ns == V.s_target && s == V.s_code && i < V.n_local && C->history)
            ss_hist_append(C->history + i * SS_HIST_W, first);
        if (i < V.n_local) {
            dir_change(dir, V, s, ns, e - b);
            if (V.push_next && (ns != s || ns == V.inf_code || trans_thr(C, ns) != 0u))
                ((SS_G uint8_t *)C->tnext)[tix] = C->ptag;
        }
        const uint64_t bit = 1ull << (i & 63);
        put_bits(V.nb, i >> 6, ns == V.inf_code ? bit : 0ull, ns == V.inf_code ? 0ull : bit);
    }
    if (V.push_next) {                               // block-uniform
        __syncthreads();
        if (tid == 0) *s_cnt = ns == V.inf_code && i < V.n_local ? 1 : 0;
        __syncthreads();
        if (*s_cnt) {
            Pusher(C).range(V.col_idx, b, e, tid, SS_BLOCK);
            if (tid == 0) w.pushed += (uint32_t)(e - b);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void flush_work(unsigned long long *counters, const Work &w) {
    uint32_t a = w.s_agents, d = w.draws, sc = w.scans, ed = w.edges, pu = w.pushed, sw = w.swept;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off);
        d += __shfl_down(d, off);
        sc += __shfl_down(sc, off);
        ed += __shfl_down(ed, off);
        pu += __shfl_down(pu, off);
        sw += __shfl_down(sw, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&counters[0], (unsigned long long)a);
        atomicAdd(&counters[1], (unsigned long long)d);
        atomicAdd(&counters[2], (unsigned long long)sc);
        atomicAdd(&counters[3], (unsigned long long)ed);
        atomicAdd(&counters[4], (unsigned long long)pu);
        atomicAdd(&counters[6], (unsigned long long)sw);
    }
}

// Block-wide sum of the direction statistics into one slot of the next
// tick's accumulator (called by every thread of the block).
__device__ __forceinline__ void flush_dir(Dir &d, unsigned long long *acc_next, unsigned long long *s_red) {
    // signed changes, added in two's complement
    unsigned long long a = (unsigned long long)d.ideg, c = (unsigned long long)(int64_t)d.sus;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off);
        c += __shfl_down(c, off);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        s_red[2 * wv] = a;
        s_red[2 * wv + 1] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0 && acc_next) {
        unsigned long long ta = 0, tc = 0;
        for (int q = 0; q < SS_BLOCK / 64; ++q) {
            ta += s_red[2 * q];
            tc += s_red[2 * q + 1];
        }
        const int slot = (int)(blockIdx.x % SS_ACC_SLOTS);
        SS_G unsigned long long *g = (SS_G unsigned long long *)acc_next;
        if (ta) __hip_atomic_fetch_add(g + 2 * slot, ta, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (tc) __hip_atomic_fetch_add(g + 2 * slot + 1, tc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    d = Dir{};
}

// LDS queue of the light tiles: the wide tile's for 16 agents per lane.
template <int A> struct LightQueue { typedef ScanQueue type; };
template <> struct LightQueue<SS_WIDE_A> { typedef WideQueue type; };

// Launch step k of a tick loop: every live replica r runs tick reps[r].tick0 + k.
// Blocks stride over the tile list (block b takes tiles b, b + G, ...; the
// host orders the list so that every block gets a share of the expensive
// tiles); no queue, no atomics.
// Four waves per SIMD are asked for: the few values that then live in scratch
// are read once per tile, while a fifth of the latency hiding would be lost
// at three waves.
// CNT: the work counters are kept (a separate instantiation, so that the
// counting registers cost nothing when counting is off).
template <int A, bool CNT, bool HALO = false>
__global__ __launch_bounds__(SS_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void tick_kernel(const ss_rep_dev *__restrict__ reps,
                                                        const ss_tile *__restrict__ tiles, int n_tiles, int k,
                                                        unsigned long long *counters) {
    __shared__ Cold cold;
    LCold *const C = (LCold *)&cold;
    __shared__ int s_cnt;
    __shared__ long long s_pos;
    __shared__ unsigned long long s_red[2 * SS_BLOCK / 64];
    __shared__ typename LightQueue<A>::type Q;
    Work w;
    Dir dir;
    int cached = -1;
    bool live = false;
    bool first = false;                              // the replica's first tile of this block
    bool first_run = true;                           // ...and whether its tags let it run
    const SS_G uint8_t *tcur = nullptr;              // tile tags of this tick, null when every tile runs
    unsigned long long *acc_next = nullptr;          // statistics of the next tick
    TickView V;
    ss_tile nt{};
    if ((int)blockIdx.x < n_tiles) nt = tiles[blockIdx.x];
    for (int tix = blockIdx.x; tix < n_tiles; tix += gridDim.x) {
        // the next tile's descriptor is loaded while this one runs
        const int32_t rep = nt.rep, kind = nt.kind, tag_lo = nt.tag_lo, tag_hi = nt.tag_hi;
        const int64_t begin = nt.begin, end = nt.end;
        if (tix + (int)gridDim.x < n_tiles) nt = tiles[tix + gridDim.x];
        first = false;
        if (rep != cached) {                         // block-uniform
            if (cached >= 0) flush_dir(dir, acc_next, s_red);   // barriers inside: C is free after
            const ss_rep_dev *R = reps + rep;
            const int64_t t = R->tick0 + k;
            live = t < R->tick_end;
            if (live) {
                const int32_t last = R->inf_T_last;
                // what this tick needs from the replica beyond its descriptor,
                // loaded before the barrier so that the loads overlap: the
                // threshold table, the push decision of this tick (tags exist
                // when the previous tick decided so), the rule for the next
                // tick, and this tile's tags
                const uint32_t tval =
                    threadIdx.x < SS_TLDS ? R->inf_T[(int32_t)threadIdx.x < last ? (int32_t)threadIdx.x : last] : 0u;
                const uint32_t dec = R->cnt[0] ? R->dec[t & 1] : 0u;
                const bool produce = push_rule(R, t);
                tcur = dec == 1u ? (const SS_G uint8_t *)R->tact[t & 1] : nullptr;
                acc_next = R->acc ? acc_slot(R, t + 1) : nullptr;
                V.tag = ss_front_tag(t);
                first = true;
                first_run = true;
                if (tcur) {
                    first_run = false;
                    for (int32_t q = tag_lo; q <= tag_hi; ++q) first_run |= tcur[q] == V.tag;
                }
                __syncthreads();                     // the previous replica's fields are no longer read
                const int64_t t1 = t + 1;
                if (threadIdx.x < SS_TLDS) C->T[threadIdx.x] = tval;
                if (threadIdx.x == 0) {
                    C->inf_T = R->inf_T;
                    C->history = R->history;
                    C->pcnt = R->cnt[t1 & 1];
                    C->tnext = R->tact[t1 & 1];
                    C->n_hub = R->n_hub;
                    C->h64 = R->h64;
                    C->hub0 = R->hub0;
                    C->heavy0 = R->heavy0;
                    C->light0 = R->light0;
                    C->ptag = ss_front_tag(t1);
                    C->pcur = R->push_sus_only ? (const int8_t *)((t & 1) != 0 ? R->state[1] : R->state[0]) : nullptr;
                    C->s_code = R->s_code;
                    if constexpr (HALO) {
                        C->hnext = k + 1 < R->halo_n ? R->halo_next[k + 1] : nullptr;
                        C->hrec = R->halo_rec;
                        C->hbase = R->halo_base;
                        C->hpacked = R->halo_packed;
                    }
                }
                if (threadIdx.x < 4) {
                    C->tT[threadIdx.x] = R->trans_T[threadIdx.x];
                    C->tN[threadIdx.x] = R->trans_next[threadIdx.x];
                }
                __syncthreads();
                const bool odd = (t & 1) != 0;
                V.row_ptr = (const SS_G int64_t *)R->row_ptr;
                V.col_idx = (const SS_G int32_t *)R->col_idx;
                V.rng_lo = (const SS_G uint32_t *)R->rng_lo;
                V.rng_hi = (const SS_G uint32_t *)R->rng_hi;
                V.cur = (const SS_G int8_t *)(odd ? R->state[1] : R->state[0]);
                V.nxt = (SS_G int8_t *)(odd ? R->state[0] : R->state[1]);
                V.cb = (const SS_G uint64_t *)(odd ? R->ibits[1] : R->ibits[0]);
                V.nb = (SS_G uint64_t *)(odd ? R->ibits[0] : R->ibits[1]);
                V.deg8 = (const SS_G uint8_t *)R->deg8;
                // this tick writes the next tick's tags when the rule says so
                V.cnt = dec == 0u ? nullptr : (SS_G uint16_t *)R->cnt[t & 1];
                V.push_next = produce ? 1 : 0;
                if (tix == R->lead_tile && threadIdx.x == 0) {
                    // one block per replica: the decision for the next tick, the
                    // next tick's statistics started from this tick's totals
                    // (every tile adds its changes), the slot of t + 2 cleared
                    if (R->cnt[0]) R->dec[(t + 1) & 1] = produce ? 1u : 0u;
                    unsigned long long d, c;
                    dir_totals(R, t, d, c);
                    unsigned long long *nx = acc_slot(R, t + 1);
                    atomicAdd(&nx[0], d);
                    atomicAdd(&nx[1], c);
                    unsigned long long *z = acc_slot(R, t + 2);
                    for (int q = 0; q < SS_ACC_SLOTS * 2; ++q) z[q] = 0ull;
                    if (CNT && dec != 0u) atomicAdd(&counters[5], 1ull);
                }
                V.last = last;
                V.s_code = R->s_code;
                V.inf_code = R->inf_code;
                V.s_target = R->s_target;
                V.seed_lo = R->seed_lo;
                V.seed_hi = R->seed_hi;
                V.t_lo = (uint32_t)t;
                V.t_hi = (uint32_t)((uint64_t)t >> 32);
                V.n_local = R->n_local;
            }
            cached = rep;
        }
        if (!live) continue;                         // replica finished: block-uniform
        // push tick with tile tags: a tile without a tag has no candidate, no
        // agent that draws and the same states in both buffers (block-uniform)
        if (first) {
            if (!first_run) continue;
        } else if (tcur) {
            bool run = false;
            for (int32_t q = tag_lo; q <= tag_hi; ++q) run |= tcur[q] == V.tag;
            if (!run) continue;
        }
        if (threadIdx.x == 0) w.swept += (uint32_t)(end - begin);
        if (kind == SS_TILE_LIGHT) {
            if constexpr (A == SS_WIDE_A)
                light_wide<HALO>(V, C, begin, end, w, dir, Q);
            else
                light_tile<A>(V, C, begin, end, tag_lo, tag_hi, w, dir, Q);
        }
        else if (kind == SS_TILE_HEAVY)
            heavy_tile(V, C, begin, end, tag_lo, w, dir);
        else
            hub_tile(V, C, begin, tag_lo, w, dir, &s_cnt, &s_pos);
    }
    if (cached >= 0) flush_dir(dir, acc_next, s_red);
    if constexpr (CNT) flush_work(counters, w);
}

// Infectious-neighbour counts of the current tick of one replica (cleared
// before), pushed along the rows of its infectious agents, when *dec is set.
// Rows [0, heavy) (the rows of degree >= 64 of a degree-sorted replica) are
// walked by a wavefront each, the others by a thread.
__global__ void count_init_kernel(const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                                  const uint64_t *__restrict__ bits, int64_t n_local, int64_t heavy,
                                  uint16_t *cnt, const uint32_t *dec) {
    if (*dec == 0u) return;
    const int lane = threadIdx.x & 63;
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = gt >> 6; i < heavy; i += nt >> 6) {
        if (!((bits[i >> 6] >> (i & 63)) & 1ull)) continue;
        const int64_t b = row_ptr[i], e = row_ptr[i + 1];
        for (int64_t j = b + lane; j < e; j += 64) ss_count_add(cnt, col[j]);
    }
    for (int64_t i = heavy + gt; i < n_local; i += nt) {
        if (!((bits[i >> 6] >> (i & 63)) & 1ull)) continue;
        const int64_t b = row_ptr[i], e = row_ptr[i + 1];
        for (int64_t j = b; j < e; ++j) ss_count_add(cnt, col[j]);
    }
}

// The push decision of the current tick t from the statistics of tick t (one
// thread): push when 16 * sum_deg(I) <= push_num * #S.
__global__ void dec_init_kernel(const unsigned long long *acc_t, uint32_t push_num, uint32_t *dec) {
    unsigned long long d = 0, c = 0;
    for (int q = 0; q < SS_ACC_SLOTS; ++q) {
        d += acc_t[2 * q];
        c += acc_t[2 * q + 1];
    }
    // the counts of this tick come from here, not from a previous tick's tile tags
    *dec = push_num != 0u && d * 16ull <= (unsigned long long)push_num * c ? 2u : 0u;
}

__global__ void dec_demote_kernel(uint32_t *dec) {
    if (threadIdx.x < 2 && dec[threadIdx.x] == 1u) dec[threadIdx.x] = 2u;
}

// Direction statistics of one state buffer (local agents): acc[0] += sum of
// degrees of agents in state inf, acc[1] += number of agents in state s_code.
__global__ void acc_init_kernel(const int8_t *__restrict__ st, const int64_t *__restrict__ row_ptr, int64_t n_local,
                                int8_t s_code, int8_t inf, unsigned long long *acc) {
    unsigned long long d = 0, c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_local;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int8_t s = st[i];
        if (s == inf) d += (unsigned long long)(row_ptr[i + 1] - row_ptr[i]);
        c += s == s_code;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        d += __shfl_down(d, off);
        c += __shfl_down(c, off);
    }
    // one pair of atomics per block, spread over the slots (same-address
    // atomics serialise)
    __shared__ unsigned long long red[2 * 256 / 64];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[2 * wv] = d;
        red[2 * wv + 1] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
            d += red[2 * q];
            c += red[2 * q + 1];
        }
        const int slot = (int)(blockIdx.x % SS_ACC_SLOTS);
        if (d) atomicAdd(&acc[2 * slot], d);
        if (c) atomicAdd(&acc[2 * slot + 1], c);
    }
}

inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ---- launchers (host) ----
hipError_t ss_launch_tick(const ss_rep_dev *reps, const ss_tile *tiles, int n_tiles, int k, int apl, int grid,
                          unsigned long long *counters, bool halo, hipStream_t st) {
    if (n_tiles <= 0) return hipSuccess;
    if (halo) {                                  // fused halo refresh: 16-agent wide tiles only
        if (apl != 16) return hipErrorInvalidValue;
        if (counters)
            hipLaunchKernelGGL((tick_kernel<16, true, true>), dim3(grid), dim3(SS_BLOCK), 0, st, reps, tiles, n_tiles,
                               k, counters);
        else
            hipLaunchKernelGGL((tick_kernel<16, false, true>), dim3(grid), dim3(SS_BLOCK), 0, st, reps, tiles,
                               n_tiles, k, counters);
        return hipGetLastError();
    }
#define SS_TICK_CASE(a)                                                                                     \
    case a:                                                                                                 \
        if (counters)                                                                                       \
            hipLaunchKernelGGL((tick_kernel<a, true>), dim3(grid), dim3(SS_BLOCK), 0, st, reps, tiles, n_tiles, k, \
                               counters);                                                                   \
        else                                                                                                \
            hipLaunchKernelGGL((tick_kernel<a, false>), dim3(grid), dim3(SS_BLOCK), 0, st, reps, tiles, n_tiles, \
                               k, counters);                                                                \
        break;
    switch (apl) {
        SS_TICK_CASE(1)
        SS_TICK_CASE(2)
        SS_TICK_CASE(4)
        SS_TICK_CASE(8)
        SS_TICK_CASE(16)
    default:
        return hipErrorInvalidValue;
    }
#undef SS_TICK_CASE
    return hipGetLastError();
}

hipError_t ss_launch_count_init(const int64_t *row_ptr, const int32_t *col_idx, const uint64_t *bits,
                                int64_t n_local, int64_t heavy, uint16_t *cnt, uint32_t *dec,
                                const unsigned long long *acc_t, uint32_t push_num, int grid, hipStream_t s) {
    hipLaunchKernelGGL(dec_init_kernel, dim3(1), dim3(1), 0, s, acc_t, push_num, dec);
    if (n_local <= 0) return hipGetLastError();
    hipLaunchKernelGGL(count_init_kernel, dim3(grid), dim3(256), 0, s, row_ptr, col_idx, bits, n_local, heavy, cnt,
                       (const uint32_t *)dec);
    return hipGetLastError();
}

hipError_t ss_launch_dec_demote(uint32_t *dec, hipStream_t s) {
    hipLaunchKernelGGL(dec_demote_kernel, dim3(1), dim3(64), 0, s, dec);
    return hipGetLastError();
}

hipError_t ss_launch_acc_init(const int8_t *st, const int64_t *row_ptr, int64_t n_local, int8_t s_code, int8_t inf_code,
                              unsigned long long *acc, int slot, int grid, hipStream_t s) {
    if (n_local <= 0) return hipSuccess;
    int g = (int)std::min<int64_t>(grid, (int64_t)grid_for(n_local, 256));
    hipLaunchKernelGGL(acc_init_kernel, dim3(g), dim3(256), 0, s, st, row_ptr, n_local, s_code, inf_code,
                       acc + slot * (SS_ACC_SLOTS * 2));
    return hipGetLastError();
}

hipError_t ss_occupancy_tick(int apl, int *blocks_per_cu) {
    switch (apl) {
    case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, tick_kernel<1, false>, SS_BLOCK, 0);
    case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, tick_kernel<2, false>, SS_BLOCK, 0);
    case 4: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, tick_kernel<4, false>, SS_BLOCK, 0);
    case 8: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, tick_kernel<8, false>, SS_BLOCK, 0);
    case 16: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, tick_kernel<16, false>, SS_BLOCK, 0);
    default: return hipErrorInvalidValue;
    }
}
