// Device-side data layout and inline device functions of the per-tick agent
// step (gfx950).  Shared by the fused tick kernels, the halo kernels and the
// hipRTC-compiled step functions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SS_HIST_W 32
#define SS_U23_MASK 0x7FFFFFu

// A replica is one independent simulation: its CSR, its int8 state ping-pong
// buffers, its infectious-neighbour bitmaps and its compartment model.  The
// plain path runs one replica; the ensemble and the serving loop run many in
// the same launch.  Slots [0, n_local) are stepped; [n_local, n_total) are
// ghost slots whose state is carried over and refreshed from the halo.
struct ss_rep_dev {
    const int64_t *row_ptr;     // [n_total + 1]; ghost rows are empty
    const int32_t *col_idx;     // [nnz] slot indices
    const uint32_t *rng_lo;     // [n_local] low word of the agent's RNG key id
    const uint32_t *rng_hi;     // [n_local] high word, or null when all ids < 2^32
    int8_t *state[2];           // ping-pong; tick t reads state[t & 1]
    uint64_t *ibits[2];         // bit j of word w: state[.][64w + j] == inf_code
    int32_t *history;           // [n_local][SS_HIST_W] or null
    const uint32_t *inf_T;      // inf_T[k] = ceil(p_k * 2^23), k = #infectious neighbours
    int32_t inf_T_last;         // last index of inf_T (k is clamped to it)
    uint32_t push_ok;           // CSR is symmetric and there are no ghosts: candidates may be pushed
    uint32_t push_sus_only;     // no state returns to s_code: counts go only to agents susceptible now
    uint32_t trans_T[4];        // spontaneous transition thresholds per state (0 = none)
    int8_t trans_next[4];       // ...and target states
    int8_t s_code, inf_code, s_target, n_states;
    uint32_t seed_lo, seed_hi;
    int64_t tick0;              // tick of launch step k = 0
    int64_t tick_end;           // ticks >= tick_end are not run (replica finished)
    int64_t n_local, n_total;
    const uint8_t *deg8;        // [n_local] min(degree, 255); 255 means "read row_ptr"
    // Push direction.  cnt[t & 1][i] = number of CSR entries of agent i that
    // are infectious at tick t, pushed by tick t - 1 along the rows of the
    // agents it leaves infectious (the CSR is symmetric) when it decides to,
    // and cleared by tick t as it reads them.  dec[t & 1] records the
    // decision: 0 = pull, 1 = push with tile tags, 2 = push, every tile runs.
    // 16-bit counts (rows shorter than 65535 entries, checked at creation),
    // added to as halves of aligned 32-bit words (ss_count_add).  Null when
    // !push_ok.
    uint16_t *cnt[2];
    uint32_t *dec;
    // Direction statistics, [tick % 3][slot][2]: [0] = sum of degrees of the
    // infectious agents, [1] = number of susceptible agents at that tick,
    // spread over SS_ACC_SLOTS slots to keep atomics apart.  Tick t adds its
    // changes to the totals of tick t.
    unsigned long long *acc;
    uint32_t push_num;          // push when 16 * sum_deg(I) <= push_num * #S (0: never)
    int32_t lead_tile;          // the tile whose block writes the replica's per-tick flags
    // Tag slot of an agent (for the tile tags of the push direction): agents
    // [0, n_hub) have a HUB tile each, whose index is its slot, from hub0 on;
    // [n_hub, h64) HEAVY tiles of SS_HEAVY_TILE agents, slots from heavy0 on;
    // the rest blocks of 2^SS_LTAG_SHIFT agents, slots from light0 on (a light
    // tile runs when one of the blocks it overlaps is tagged).
    // tact[t & 1][slot] == tag(t): the slot's agents may change at tick t.
    uint8_t *tact[2];
    int64_t n_hub, h64;
    int32_t hub0, heavy0, light0, light_cap;
    // Fused halo refresh (a run fed by a host halo): halo_next[k] is the
    // device copy of the halo of launch step k of the current call, or null
    // where step k starts with its own refresh; step k - 1 then writes that
    // halo into the ghost slots of the next buffer (and their bitmap bits) as
    // the refresh kernel would before step k.  halo_rec[g] is the record of
    // slot halo_base + g (-1: not bound), padded to whole 16-slot groups.
    const uint8_t *const *halo_next;
    const int32_t *halo_rec;
    int64_t halo_base;
    int32_t halo_n;             // entries of halo_next (0: no fused refresh)
    uint32_t halo_packed;       // 2-bit codes, four records a byte
};

#define SS_LTAG_SHIFT 10

#define SS_ACC_SLOTS 16
#define SS_ACC_WORDS (3 * SS_ACC_SLOTS * 2)

// Tile tag of tick t: never 0 (the cleared value).  A tag equal to that of a
// tick 510 steps back (same parity) only makes an idle tile run, which
// changes no result.
__host__ __device__ __forceinline__ uint8_t ss_front_tag(int64_t t) {
    return (uint8_t)(1 + (uint64_t)t % 255u);
}

// A tile is an agent range of one replica.  LIGHT tiles (64-aligned) map
// agents to lanes; HEAVY tiles give each agent's scan to a wavefront; a HUB
// tile is one agent whose scan takes the whole block.
enum { SS_TILE_LIGHT = 0, SS_TILE_HEAVY = 1, SS_TILE_HUB = 2 };
struct ss_tile {
    int32_t rep;
    int32_t kind;
    int64_t begin;
    int64_t end;      // exclusive; end - begin <= tile capacity
    int32_t tag_lo, tag_hi;   // tag slots the tile's agents map to (inclusive)
};

// One more infectious neighbour for agent c: the 16-bit count is one half of
// an aligned 32-bit word (no carry: a count never reaches 65536).
__device__ __forceinline__ void ss_count_add(uint16_t *cnt, int64_t c) {
    __hip_atomic_fetch_add((__attribute__((address_space(1))) uint32_t *)(cnt + (c & ~(int64_t)1)),
                           1u << (16 * (int)(c & 1)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// a ^ b ^ c: one V_BITOP3_B32 (truth table 0x96) on the device.
__host__ __device__ __forceinline__ uint32_t ss_xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

// Philox-4x32-10 (Salmon et al., SC'11), word 0 only.
__host__ __device__ __forceinline__ uint32_t ss_philox4x32_w0(uint32_t c0, uint32_t c1, uint32_t c2,
                                                              uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = ss_xor3((uint32_t)(p1 >> 32), c1, k0);
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = ss_xor3((uint32_t)(p0 >> 32), c3, k1);
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    return c0;
}

// The draw of agent `id` at tick t: key = seed, counter = (t, id) as 64-bit
// halves; 23 low bits of word 0 (the reference's float32-mantissa conversion,
// math_utils.py:189, kept as the integer numerator m of u = m / 2^23).
__device__ __forceinline__ uint32_t ss_draw(const ss_rep_dev &R, int64_t t, int64_t i) {
    uint32_t hi = R.rng_hi ? R.rng_hi[i] : 0u;
    return ss_philox4x32_w0((uint32_t)t, (uint32_t)((uint64_t)t >> 32), R.rng_lo[i], hi, R.seed_lo,
                            R.seed_hi) & SS_U23_MASK;
}

__device__ __forceinline__ bool ss_bit(const uint64_t *bits, int64_t slot) {
    return (bits[slot >> 6] >> (slot & 63)) & 1ull;
}

// Append `who` to a -1 padded history row (first free entry; when the row is
// full the oldest entry is dropped).
// Rare path (once per infection): kept as rolled loops so it does not raise
// the register budget of the kernels that inline it.
__device__ __forceinline__ void ss_hist_append(int32_t *row, int32_t who) {
#pragma unroll 1
    for (int j = 0; j < SS_HIST_W; ++j)
        if (row[j] == -1) { row[j] = who; return; }
#pragma unroll 1
    for (int j = 0; j < SS_HIST_W - 1; ++j) row[j] = row[j + 1];
    row[SS_HIST_W - 1] = who;
}
