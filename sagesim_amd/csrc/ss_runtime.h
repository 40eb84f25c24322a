// Host-side runtime structures of libsagestep (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <map>
#include <string>
#include <vector>

#include "../../include/sagestep.h"
#include "ss_device.h"
#include "ss_map.h"

#define SS_BLOCK 256
// device slots of a bound halo (and pinned staging buffers): the host runs at
// most this many ticks ahead of the refreshes that read them
#define SS_HALO_RING 4
// page-locked halo input: ticks per group and device bytes per group buffer
#define SS_HALO_GROUP 32
#define SS_HALO_GROUP_BYTES (64ll << 20)
#define SS_HALO_SPANS 8                         // copies per group: spans of 1, 1, 2, 4, ... ticks
// work counters: S agents visited, draws, scans, CSR entries scanned,
// counts pushed, replica-ticks taken in the push direction, agents of the
// tiles that ran
#define SS_WORK_N 7

// kernel launchers (tick_kernels.hip)
hipError_t ss_launch_tick(const ss_rep_dev *reps, const ss_tile *tiles, int n_tiles, int k, int apl, int grid,
                          unsigned long long *counters, bool halo, hipStream_t st);
// agents of one HEAVY tile (4 wavefronts x 4 agents, tick_kernels.hip SS_HPW)
#define SS_HEAVY_TILE 16
#define SS_HEAVY_SHIFT 4
// degree thresholds that have a precomputed prefix length per replica
#define SS_N_THRESHOLDS 16
static const int32_t SS_THRESHOLDS[SS_N_THRESHOLDS] = {8, 16, 24, 32, 48, 64, 96, 128, 192, 256, 512,
                                                        1024, 2048, 4096, 16384, 1 << 30};
hipError_t ss_occupancy_tick(int apl, int *blocks_per_cu);
hipError_t ss_launch_bits(const int8_t *st, uint64_t *bits, int64_t w0, int64_t w1, int64_t n_total, int8_t inf,
                          hipStream_t s);
hipError_t ss_launch_map_lookup(const unsigned long long *keys, const int32_t *vals, int64_t cap, const int64_t *ids,
                                int32_t *out, int64_t n, int32_t *misses, hipStream_t s);
// ghost refresh from a bound halo: slot base + g <- vals[rec[g]] (rec -1: kept),
// both buffers, and the bitmap words of [base, n_total) of the buffer `cur`;
// packed: vals holds 2-bit codes, four records a byte
hipError_t ss_launch_halo_gather(int8_t *s0, int8_t *s1, const int8_t *cur, uint64_t *bits, const int32_t *rec,
                                 const int8_t *vals, int64_t base, int64_t n_total, int8_t inf, bool packed,
                                 hipStream_t s);
hipError_t ss_launch_halo_unpack_ids(const unsigned long long *keys, const int32_t *mvals, int64_t cap,
                                     const int64_t *ids, const int8_t *vals, int64_t n, int8_t *s0, int8_t *s1,
                                     int32_t *misses, hipStream_t s);
hipError_t ss_launch_gather(const int8_t *st, const int32_t *slots, int8_t *out, int64_t n, hipStream_t s);
// out[h / 4] bits 2 (h % 4).. = st[slots[h]] & 3
hipError_t ss_launch_gather2(const int8_t *st, const int32_t *slots, uint8_t *out, int64_t n, hipStream_t s);
hipError_t ss_launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t s);
hipError_t ss_launch_copy(const void *src, void *dst, int64_t bytes, int grid, hipStream_t s);
// n scattered 4-byte reads of src[0 .. n_words); out holds grid * 256 words
hipError_t ss_launch_gather_probe(const uint32_t *src, int64_t n_words, int64_t n, uint32_t *out, int grid,
                                  hipStream_t s);
hipError_t ss_launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, int grid, hipStream_t s);
hipError_t ss_launch_check_cols(const int32_t *col, int64_t nnz, int64_t n_total, unsigned long long *first_bad,
                                int grid, hipStream_t s);
hipError_t ss_launch_stats(const int8_t *st, const int64_t *row_ptr, int64_t n_local, int8_t s_code,
                           unsigned long long *out, int grid, hipStream_t s);
// push decision of the current tick from its statistics acc_t, then (when
// pushing) its infectious-neighbour counts (cleared before) from the bitmap
hipError_t ss_launch_count_init(const int64_t *row_ptr, const int32_t *col_idx, const uint64_t *bits,
                                int64_t n_local, int64_t heavy, uint16_t *cnt, uint32_t *dec,
                                const unsigned long long *acc_t, uint32_t push_num, int grid, hipStream_t s);
// a push decision with tile tags (1) becomes "push, every tile runs" (2)
hipError_t ss_launch_dec_demote(uint32_t *dec, hipStream_t s);
// direction statistics of the state buffer st into statistics slot `slot` (zeroed first by the caller)
hipError_t ss_launch_acc_init(const int8_t *st, const int64_t *row_ptr, int64_t n_local, int8_t s_code, int8_t inf_code,
                              unsigned long long *acc, int slot, int grid, hipStream_t s);
// sums of two 64-bit hashes of every entry (i, j) and of its mirror (j, i): out[0..1] and out[2..3]
hipError_t ss_launch_symcheck(const int64_t *row_ptr, const int32_t *col_idx, int64_t n_rows,
                              unsigned long long *out, int grid, hipStream_t s);

// One replica's device allocations and host bookkeeping.
struct ss_replica {
    bool live = false;
    int64_t n_local = 0, n_ghost = 0, n_total = 0, nnz = 0;
    int64_t n_words = 0;                       // bitmap words per buffer
    // device memory (owned)
    int64_t *row_ptr = nullptr;
    int32_t *col_idx = nullptr;
    uint32_t *rng_lo = nullptr, *rng_hi = nullptr;
    uint8_t *deg8 = nullptr;
    int8_t *state[2] = {nullptr, nullptr};
    int8_t *init_state = nullptr;
    uint64_t *ibits[2] = {nullptr, nullptr};
    int32_t *history = nullptr, *init_history = nullptr;
    uint32_t *inf_T = nullptr;
    uint16_t *cnt[2] = {nullptr, nullptr};     // push-direction infectious-neighbour counts [n_local] by tick parity
    uint32_t *dec = nullptr;                   // push decisions [2] by tick parity
    unsigned long long *acc = nullptr;         // direction statistics [SS_ACC_WORDS]
    bool symmetric = false;                    // every entry (i, j) has its mirror (j, i)
    // id -> slot map for ghosts
    unsigned long long *map_keys = nullptr;
    int32_t *map_vals = nullptr;
    int64_t map_cap = 0, map_size = 0;
    // bound halo: slot of every record of the per-tick halo buffer
    int32_t *halo_slots = nullptr;
    int64_t n_halo = 0;
    int32_t *halo_rec = nullptr;                // record of slot halo_base + g, -1 when not bound
    int64_t halo_base = 0;                      // first slot refreshed (a multiple of 64)
    int8_t *halo_dev[SS_HALO_RING] = {};       // device staging of halo values, a ring by tick
    int8_t *halo_grp[2] = {};                   // device staging of a group of ticks (page-locked input)
    int64_t halo_grp_bytes = 0;                 // bytes of each halo_grp buffer
    const uint8_t **halo_ptrs = nullptr;        // per launch step of a call: its staged halo (fused refresh)
    int64_t halo_ptrs_cap = 0;
    // bookkeeping
    int64_t tick = 0, tick_init = 0;
    int64_t tick_limit = INT64_MAX;            // replica finishes at this tick
    int64_t max_deg = 0;
    std::vector<int64_t> heavy_prefix_deg;      // thresholds
    std::vector<int64_t> heavy_prefix_len;      // rows at the front with deg >= threshold
    std::vector<int64_t> edge64;                // CSR entries before agents 0, 64, 128, ... and n_total
    ss_rep_dev dev{};                           // device descriptor mirror (pointers filled)
    ss_model model{};                           // the compartment model last applied
    uint64_t csr_hash[2] = {0, 0};              // multiset hash of the CSR entries (checkpoint identity)
    uint64_t key_hash = 0;                      // hash of the RNG key ids
    size_t bytes = 0;                           // device bytes held
};

struct ss_ctx {
    int device = 0;
    int num_cus = 0;
    int64_t total_mem = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;
    std::vector<ss_replica> reps;
    ss_rep_dev *d_reps = nullptr;              // device descriptor table
    ss_rep_dev *h_reps = nullptr;              // pinned host mirror
    int reps_cap = 0;
    ss_tile *d_tiles = nullptr;
    uint8_t *d_tact[2] = {nullptr, nullptr};   // tile tags of the push direction by tick parity
    int64_t tiles_cap = 0;
    int64_t slots_cap = 0;                      // tag slots allocated in d_tact
    int n_tiles = 0;
    bool tiles_dirty = true;
    // tiling (autotuned by the host library)
    int hub_deg = 1024;
    int heavy_deg = 256;
    int apl = 16;
    int grid_blocks = 0;                        // 0: persistent grid from occupancy
    int use_graph = 1;
    int grid = 0;
    std::map<int64_t, hipGraphExec_t> graphs;   // by tick count
    int32_t *d_counters = nullptr;              // misc device counters (misses, overflow)
    unsigned long long *d_work = nullptr;       // [SS_WORK_N] tick-kernel work counters
    unsigned long long *d_stats = nullptr;      // [8] census output
    bool counting = false;
    uint32_t push_num = 4;                      // direction rule, 16 * ratio (0: pull only); ratio 0.25
    hipEvent_t t_start = nullptr, t_stop = nullptr;
    int8_t *h_halo_pin[SS_HALO_RING] = {};      // pinned staging for per-tick halo values
    int64_t h_halo_cap = 0;
    hipEvent_t halo_ev[SS_HALO_RING] = {};      // ring slot b copied to the device
    hipEvent_t halo_used[SS_HALO_RING] = {};    // the refresh that read ring slot b is done
    int64_t halo_pos = 0;                       // ring position of the next halo tick
    hipEvent_t grp_done[2] = {};                // the last refresh that read group buffer b is done
    hipEvent_t span_ev[2][SS_HALO_SPANS] = {};  // copy span s of group buffer b is on the device
    int64_t grp_pos = 0;                        // groups issued (buffer = grp_pos & 1)
    bool halo_fused = false;                    // this call's ticks fuse the next tick's halo refresh
    std::string err;
    size_t bytes = 0;
};
