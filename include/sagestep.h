/*
 * libsagestep: C ABI of the MI355X (gfx950) per-tick agent step.
 *
 * The reference has no native library: its hot path is Python + CuPy JIT
 * (sagesim/model.py, sagesim/gpu_kernels.py).  Each entry point below names
 * the reference code whose job it takes over.  Paths are relative to the
 * reference repository root.
 *
 * Conventions
 *   - Every function returns int: SS_OK (0) or a negative SS_ERR_* code; the
 *     message is available from ss_last_error(ctx) (ctx may be NULL for
 *     errors raised before a context exists).
 *   - Host pointers are borrowed for the duration of the call.  The context
 *     owns all device memory.
 *   - A context is single-threaded: the caller serialises calls on it.
 *     Several contexts may coexist (each has its own HIP stream).
 *   - Work is ordered on the context's stream.  ss_simulate returns once the
 *     ticks are queued; ss_sync (and every download) waits for them.
 *   - Nothing throws across the ABI.
 */
#ifndef SAGESTEP_H
#define SAGESTEP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SS_OK 0
#define SS_ERR_INVALID (-1)   /* bad argument or shape */
#define SS_ERR_HIP (-2)       /* HIP runtime error */
#define SS_ERR_OOM (-3)       /* device allocation failed */
#define SS_ERR_RANGE (-4)     /* index or id outside its range */
#define SS_ERR_STATE (-5)     /* call not valid in the current state */
#define SS_ERR_RTC (-6)       /* hipRTC compile or load error */
#define SS_ERR_FULL (-7)      /* hash map or budget exhausted */

#define SS_HISTORY_WIDTH 32

typedef struct ss_ctx ss_ctx;

/* Compartment model of one replica.  An agent in s_code with k neighbours
 * in inf_code moves to s_target when u < 1 - (1-beta)^k; an agent in state s
 * moves to trans_next[s] when u < trans_p[s].  u = m / 2^23 where m is the 23
 * low bits of word 0 of Philox-4x32-10(key = seed, counter = (tick, agent
 * id)); the reference's conversion (math_utils.py:189).  The power is formed
 * by repeated multiplication in double: q_0 = 1, q_k = q_{k-1} * (1 - beta).
 * SIR: {S=0, I=1, R=2}; SEIR: {S=0, E=1, I=2, R=3}.  States are int8. */
typedef struct ss_model {
    int32_t n_states;
    int32_t s_code;
    int32_t inf_code;
    int32_t s_target;
    double beta;
    double trans_p[4];
    int32_t trans_next[4];
} ss_model;

/* One replica (an independent simulation).  Slots [0, n_local) are stepped;
 * slots [n_local, n_local + n_ghost) are ghosts (model.py:1069-1073 puts
 * ghost slots after the local ones) whose rows are empty.  row_ptr has
 * n_local + 1 entries; col_idx values are slots in [0, n_local + n_ghost).
 * rng_key_id[i] is the id the RNG of local slot i is keyed on (the original
 * agent id when the caller relabels rows, model.py:1089-1096 keys the RNG on
 * logical ids for the same reason). */
typedef struct ss_replica_spec {
    int64_t n_local;
    int64_t n_ghost;
    int64_t nnz;
    const int64_t *row_ptr;
    const int32_t *col_idx;
    const int64_t *rng_key_id;
    const int8_t *state;          /* n_local + n_ghost initial states */
    const int32_t *history;       /* n_local * SS_HISTORY_WIDTH, -1 padded, or NULL */
    ss_model model;
    uint64_t seed;
    int64_t tick0;                /* tick of the first simulated tick */
    int64_t tick_limit;           /* the replica stops at this tick; <= 0: never */
} ss_replica_spec;

/* ---- context: replaces GPUBufferManager construction (gpu_kernels.py:263-304)
 *      and the device setup of Model.setup (model.py:682-886) ---- */
int ss_create(int device, ss_ctx **out);
int ss_destroy(ss_ctx *ctx);
const char *ss_last_error(const ss_ctx *ctx);
int ss_sync(ss_ctx *ctx);
/* compute units, device memory, and the bytes the context holds */
int ss_device_info(ss_ctx *ctx, int32_t *num_cus, int64_t *total_mem, int64_t *ctx_bytes);

/* ---- replicas: replace _build_gpu_buffers (model.py:1020-1274),
 *      GPUBufferManager.allocate_property_tensors / allocate_write_buffers /
 *      allocate_csr (gpu_kernels.py:306-359) ---- */
int ss_replica_create(ss_ctx *ctx, const ss_replica_spec *spec, int32_t *rep);
int ss_replica_destroy(ss_ctx *ctx, int32_t rep);
/* device bytes a replica of this shape takes (serving admission) */
int64_t ss_replica_footprint(int64_t n_local, int64_t n_ghost, int64_t nnz, int32_t with_history, int64_t max_deg);
/* back to the state, history and tick given at creation (Model.reset, model.py:944-958) */
int ss_replica_reset(ss_ctx *ctx, int32_t rep);
int ss_replica_tick(ss_ctx *ctx, int32_t rep, int64_t *tick);
int ss_replica_set_tick(ss_ctx *ctx, int32_t rep, int64_t tick, int64_t tick_limit);
int ss_replica_bytes(ss_ctx *ctx, int32_t rep, int64_t *bytes);
int ss_set_model(ss_ctx *ctx, int32_t rep, const ss_model *model);
/* Model.set_seed (model.py:652) */
int ss_set_seed(ss_ctx *ctx, int32_t rep, uint64_t seed);

/* ---- state readback / upload: replace get_agent_property_value's slice
 *      .get() (model.py:422-450) and set_agent_property_value ---- */
int ss_state_download(ss_ctx *ctx, int32_t rep, int8_t *out, int64_t n);
int ss_state_upload(ss_ctx *ctx, int32_t rep, const int8_t *in, int64_t n);
int ss_history_download(ss_ctx *ctx, int32_t rep, int32_t *out, int64_t n_rows);

/* ---- checkpoint: replaces the pickle of Model.save / Model.load
 *      (model.py:1289-1310) with the replica's device state, bit for bit ----
 * The blob holds the tick, tick limit, seed, compartment model, the live
 * state buffer and the history rows, with the shape, CSR hash and RNG-key hash
 * of the replica they belong to.  buf NULL: *size = bytes needed.  Restore
 * into a replica of the same network (any context) continues the run exactly
 * as the saved one would; *tick gets the restored tick. */
int ss_replica_save(ss_ctx *ctx, int32_t rep, void *buf, int64_t cap, int64_t *size);
int ss_replica_restore(ss_ctx *ctx, int32_t rep, const void *buf, int64_t size, int64_t *tick);

/* ---- run: replaces Model.simulate -> worker_coroutine -> stepfunc launch
 *      (model.py:961-998, :1314-1640, kernel template :2421-2454) ----
 * Advances every live replica by `ticks` ticks (each stops at its own
 * tick_limit).  Tick loops are captured once per tick count in a hipGraph. */
int ss_simulate(ss_ctx *ctx, int64_t ticks);
/* Launch shape.  Rows are expected in descending-degree order.  The rows of
 * degree >= hub_degree at the front get a whole block each; the following
 * rows of degree >= heavy_degree get a wavefront each for their neighbour
 * scan; the rest run agents_per_lane (1, 2, 4, 8 or 16) agents per lane.  Degrees
 * are one of 8,16,24,32,48,64,96,128,192,256,512,1024,2048,4096,16384,2^30.
 * grid_blocks = 0 sizes the persistent grid from occupancy; use_graph = 0
 * launches tick by tick instead of replaying a captured hipGraph. */
int ss_set_tiling(ss_ctx *ctx, int32_t hub_degree, int32_t heavy_degree, int32_t agents_per_lane,
                  int32_t grid_blocks, int32_t use_graph);
/* tiles and grid of the current launch shape (after ss_simulate built them) */
int ss_launch_shape(ss_ctx *ctx, int32_t *n_tiles, int32_t *grid);

/* ---- ghosts of one partition: replace GPUHashMap (gpu_kernels.py:91-260,
 *      which the reference builds but never queries) and the pack/unpack of
 *      CommunicationManager (gpu_kernels.py:852-909, :984-1066) ---- */
/* The map's arrays are laid out as the reference's GPUHashMap leaves its CPU
 * mirror after the same calls: keys int64 (-1 empty, -2 deleted), values
 * int32 (-1 in empty and deleted slots), home slot floor-mod(id * 2654435761
 * wrapped to int64, capacity), linear probing.
 * ss_map_build: GPUHashMap(capacity).build_from_arrays(ids, slots)
 *   (gpu_kernels.py:122-169, capacity at model.py:1113); capacity =
 *   max(64, ceil(n / load_factor)); the reference's placement rounds, so a
 *   repeated id is stored once per occurrence and lookups find the first
 *   occurrence in input order; the size counts every entry.
 * ss_map_insert: GPUHashMap.insert (gpu_kernels.py:171-191) for every entry
 *   in input order: the probe stops at an empty or deleted slot or at the
 *   id's own key; the value of an existing key is overwritten.  An insert
 *   that finds no slot in a whole sweep of a full table returns SS_ERR_FULL
 *   (the reference would probe forever); grow the map with ss_map_resize, as
 *   the reference's callers are to do when needs_resize() (size > 0.7 *
 *   capacity, gpu_kernels.py:217-219).
 * ss_map_remove: GPUHashMap.remove (gpu_kernels.py:193-206) in input order;
 *   *n_removed counts the entries marked deleted. */
int ss_map_build(ss_ctx *ctx, int32_t rep, const int64_t *ids, const int32_t *slots, int64_t n,
                 double load_factor);
int ss_map_insert(ss_ctx *ctx, int32_t rep, const int64_t *ids, const int32_t *slots, int64_t n);
int ss_map_remove(ss_ctx *ctx, int32_t rep, const int64_t *ids, int64_t n, int64_t *n_removed);
/* GPUHashMap.resize (gpu_kernels.py:221-240): a table of new_capacity slots
 * with every live entry re-inserted in the old table's slot order. */
int ss_map_resize(ss_ctx *ctx, int32_t rep, int64_t new_capacity);
/* GPUHashMap.capacity and .size (gpu_kernels.py:103-106) */
int ss_map_info(ss_ctx *ctx, int32_t rep, int64_t *capacity, int64_t *size);
/* copies of GPUHashMap.keys / .values (capacity entries each) */
int ss_map_download(ss_ctx *ctx, int32_t rep, int64_t *keys, int32_t *vals, int64_t capacity);
/* out[i] = slot of ids[i] or -1 (GPUHashMap.lookup_cpu, gpu_kernels.py:208-215) */
int ss_map_lookup(ss_ctx *ctx, int32_t rep, const int64_t *ids, int32_t *out, int64_t n, int64_t *n_missing);
/* discover_ghost_topology (gpu_kernels.py:35-88) and the receive side of
 * build_communication_maps (gpu_kernels.py:521-850) on the GPU.  nbr_ids are
 * the local agents' neighbour ids (the flattened CSR values; negative entries
 * are padding), owner[id] the rank that owns agent id (-1: none; ids at or
 * beyond n_owner are ignored, as the reference's dense rank table ignores
 * them).  Ghosts are the distinct ids whose owner is neither -1 nor my_rank:
 * ghost_ids gets them in ascending order (np.unique), which is also their
 * ghost-slot order (model.py:1069-1073); for every rank r,
 * recv_index[recv_offsets[r] .. recv_offsets[r + 1]) lists the positions in
 * ghost_ids of the ghosts r owns, ascending (the per-rank np.unique of
 * gpu_kernels.py:610-617).  recv_offsets has n_ranks + 1 entries.  With
 * ghost_ids NULL only *n_ghost is set; an owner rank >= n_ranks is
 * SS_ERR_RANGE; n_ranks is at most 2^20 and n_owner at most 2^31 - 1.  The send side is ss_map_lookup of the ids a peer requests in
 * a map built over the local (and ghost) agents (model.py:1112-1115). */
int ss_ghost_discover(ss_ctx *ctx, const int64_t *nbr_ids, int64_t nnz, const int32_t *owner, int64_t n_owner,
                      int32_t my_rank, int32_t n_ranks, int64_t *ghost_ids, int32_t *recv_offsets,
                      int32_t *recv_index, int64_t cap, int64_t *n_ghost);
/* Fix the record order of the halo buffer: record h refreshes the ghost with
 * id ids[h] (build_communication_maps, gpu_kernels.py:521-850, resolves the
 * recv index lists once). */
int ss_halo_bind(ss_ctx *ctx, int32_t rep, const int64_t *ids, int64_t n, int64_t *n_missing);
/* Refresh ghosts from one halo buffer in bound order (vals[n_bound]);
 * both state buffers are written (gpu_kernels.py:1030-1036). */
int ss_halo_unpack(ss_ctx *ctx, int32_t rep, const int8_t *vals, int64_t n);
/* Refresh ghosts from (id, state) records in any order. */
int ss_halo_unpack_ids(ss_ctx *ctx, int32_t rep, const int64_t *ids, const int8_t *vals, int64_t n,
                       int64_t *n_missing);
/* out[h] = current state of slot slots[h] (the send side, gpu_kernels.py:852-909) */
int ss_halo_pack(ss_ctx *ctx, int32_t rep, const int32_t *slots, int64_t n, int8_t *out);
/* ss_halo_pack with no device allocation.  No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, per-tick
 * exchange loops that must not call hipMalloc).  Its staging is a fixed
 * device array per GPU (SS_NOALLOC_MAX_RECORDS records; n beyond that is
 * refused with SS_ERR_INVALID), held until the process ends.  Calls that
 * overlap on one GPU (several threads or contexts on the same device) race
 * on that array. */
#define SS_NOALLOC_MAX_RECORDS (1 << 20)
int ss_halo_pack_noalloc(ss_ctx *ctx, int32_t rep, const int32_t *slots, int64_t n, int8_t *out);
/* Run `ticks` ticks; before tick j the ghosts are refreshed from
 * vals[j * n_bound ...] (bound order), copied on a second stream while
 * earlier ticks compute: straight from vals when it lies in memory from
 * ss_host_alloc (in groups of up to 32 ticks, each group's copies issued
 * ahead of its ticks; device staging of up to 2 x 64 MB per replica), else
 * staged tick by tick through the library's pinned buffers.  vals is
 * borrowed for the call only. */
int ss_simulate_halo(ss_ctx *ctx, int32_t rep, int64_t ticks, const int8_t *vals);
/* 2-bit halo (no counterpart in the reference, whose halo carries one f32
 * per property and agent, gpu_kernels.py:852-909): the same three calls with
 * state codes packed four records a byte, record h in bits 2*(h%4) and
 * 2*(h%4)+1 of byte h/4, (n + 3) / 4 bytes per halo of n records: a
 * quarter of the bytes over the host link.  Every model has at most 4
 * states (ss_set_model), so its codes fit; the pack side takes states
 * modulo 4. */
int ss_halo_unpack_2bit(ss_ctx *ctx, int32_t rep, const uint8_t *packed, int64_t n);
int ss_halo_pack_2bit(ss_ctx *ctx, int32_t rep, const int32_t *slots, int64_t n, uint8_t *out);
int ss_simulate_halo_2bit(ss_ctx *ctx, int32_t rep, int64_t ticks, const uint8_t *packed);
/* Page-locked host memory for halo buffers (no counterpart in the reference:
 * for callers that feed ghost states every tick, whose per-tick copy then
 * needs no staging).  Freed with ss_host_free. */
int ss_host_alloc(ss_ctx *ctx, int64_t bytes, void **out);
int ss_host_free(ss_ctx *ctx, void *p);

/* ---- step direction: no counterpart in the reference, whose generated
 *      kernel always reads every neighbour (model.py:2302-2408) ----
 * A replica whose CSR is symmetric (checked on the device at creation) and
 * that has no ghost slots may run a tick in the push direction: the tick
 * before adds, along the row of every agent it leaves infectious, one to a
 * per-agent count of infectious neighbours and tags the tiles it touched;
 * the push tick then decides every susceptible agent from its count (no
 * neighbour scan) and skips untagged tiles.  Tick t writes the counts for
 * t + 1 when sum(degree of infectious agents at t) <= ratio * (number of
 * susceptible agents at t); ratio 0 disables it (default 0.25).  Results do
 * not depend on the direction. */
int ss_set_push(ss_ctx *ctx, double ratio);
/* out = 1 when the replica may take the push direction */
int ss_replica_push_ok(ss_ctx *ctx, int32_t rep, int32_t *out);

/* ---- measurement: replaces the verbose_timing dict of worker_coroutine
 *      (model.py:1351-1640) with device-side events and counters ---- */
/* Count the work of every tick launch: out[0] susceptible agents visited,
 * out[1] RNG draws, out[2] neighbour scans started, out[3] CSR entries read,
 * out[4] infectious-neighbour counts pushed, out[5] replica-ticks that took
 * the push direction, out[6] agents of the tiles that ran. */
int ss_set_counting(ss_ctx *ctx, int32_t on);
/* read and clear the work counters (int64 out[7]) */
int ss_read_counters(ss_ctx *ctx, int64_t *out);
/* census of the live state: out[0..3] agents per state, out[4] sum of the
 * degrees of susceptible agents (int64 out[5]) */
int ss_census(ss_ctx *ctx, int32_t rep, int64_t *out);
/* Plain device copy of `bytes` (multiple of 16) repeated `repeats` times:
 * average milliseconds per copy.  The known-size reference for hardware
 * traffic counters and for the achievable HBM bandwidth. */
int ss_copy_probe(ss_ctx *ctx, int64_t bytes, int32_t repeats, float *ms);
/* n_gathers 4-byte reads at scattered positions of a fresh buffer of
 * span_bytes (pick it far larger than the caches), repeated `repeats` times:
 * average milliseconds per pass.  The known-count reference of a random
 * gather for hardware traffic counters (the tick kernel's neighbour reads are
 * gathers, not streams). */
int ss_gather_probe(ss_ctx *ctx, int64_t span_bytes, int64_t n_gathers, int32_t repeats, float *ms);
/* HIP events on the context's stream around queued work */
int ss_timer_start(ss_ctx *ctx);
int ss_timer_stop(ss_ctx *ctx, float *ms);

/* ---- general step functions (hipRTC): replace generate_gpu_func and the
 *      CuPy JIT of the generated module (model.py:2010-2457, :800-865), the
 *      property tensors of _build_gpu_buffers (model.py:1020-1274) and the
 *      write-back of the fused kernel (model.py:2368-2408) ----
 * A program holds n agents (sorted by breed), their CSR, float32 agent ids,
 * logical ids and breed indices (model.py:1085-1096), float64 globals, up to
 * 64 device arrays (properties, breed-local arrays and their index maps;
 * element size 1/2/4/8 bytes, rows x width elements, double-buffered or not),
 * up to 16 kernel-extra buffers, and the kernels compiled from `source`,
 * which includes "sagesim_rt.h" (the device runtime this library embeds). */
typedef struct ss_prog ss_prog;
int ss_prog_create(ss_ctx *ctx, int64_t n_agents, ss_prog **out);
int ss_prog_destroy(ss_prog *prog);
const char *ss_prog_last_error(const ss_prog *prog);
int ss_prog_set_csr(ss_prog *prog, const int64_t *row_ptr, const int32_t *col_idx, int64_t nnz);
/* float32 agent ids, logical ids and breeds as the reference passes them,
 * plus the logical ids as exact int64 */
int ss_prog_set_ids(ss_prog *prog, const float *agent_ids, const float *logical_ids, const int64_t *logical_ids64,
                    const float *breeds);
/* a property: n rows (GPUBufferManager.allocate_property_tensors and
 * allocate_write_buffers, gpu_kernels.py:306-332) */
int ss_prog_add_prop(ss_prog *prog, int32_t elem_bytes, int64_t width, int32_t double_buffered, const void *init,
                     int32_t *prop_id);
/* any device array of `rows` rows: a breed-local array (model.py:1215-1270,
 * rows = agents of the breed) or an index map (rows = n); init NULL zero-fills */
int ss_prog_add_array(ss_prog *prog, int32_t elem_bytes, int64_t rows, int64_t width, int32_t double_buffered,
                      const void *init, int32_t *array_id);
int ss_prog_prop_upload(ss_prog *prog, int32_t prop_id, const void *data);
int ss_prog_prop_download(ss_prog *prog, int32_t prop_id, void *out);
int ss_prog_set_globals(ss_prog *prog, const double *values, int64_t count);
int ss_prog_set_seed(ss_prog *prog, int64_t seed);
/* kernel-extra buffers (the extra kernel arguments of _prepare_kernel_extras,
 * model.py:892-921, :1501, :1526): slot `slot` holds `bytes` bytes, visible to
 * the wrapper kernels as SSProgArgs::extra[slot]; read back after a run */
int ss_prog_set_extra(ss_prog *prog, int32_t slot, int64_t bytes, const void *data);
int ss_prog_get_extra(ss_prog *prog, int32_t slot, int64_t bytes, void *out);
/* Compile `source` for gfx950 and load the kernels named in `kernel_names`
 * ('\n'-separated); the hipRTC log is copied into `log` (may be NULL).  Code
 * objects are cached by a hash of the source, the runtime header and the
 * options, in the process and in $SAGESIM_AMD_RTC_CACHE (default
 * ~/.cache/sagesim_amd/rtc; empty: no disk cache), as the reference caches
 * its generated module (model.py:843-865). */
int ss_prog_compile(ss_prog *prog, const char *source, const char *kernel_names, char *log, int64_t log_cap);
/* *hit = 0: the last compile ran hipRTC; 1: process cache; 2: disk cache */
int ss_prog_cache_hit(ss_prog *prog, int32_t *hit);
/* The launches of one tick, in order: kernel index, first agent, agent count.
 * Every kernel takes (const SSProgArgs*, long long tick, long long start,
 * long long count, const int *order): thread k runs agent order[k], or
 * start + k when order is NULL. */
int ss_prog_set_schedule(ss_prog *prog, int32_t n, const int32_t *kernel, const int64_t *start,
                         const int64_t *count);
/* Agent order of every launch of the schedule, concatenated (each launch's
 * segment a permutation of its range), e.g. longest rows first so that the
 * threads of the longest rows start first; results do not depend on it. */
int ss_prog_set_order(ss_prog *prog, const int32_t *order, int64_t n);
/* Ticks tick0 .. tick0 + ticks - 1: the schedule, then write-back; waits. */
int ss_prog_simulate(ss_prog *prog, int64_t tick0, int64_t ticks);
/* Checkpoint of the program's state (replaces the pickle of Model.save /
 * Model.load, model.py:1289-1310): globals, seed, every array, the caller's
 * tick, and the shape of the network it belongs to.  buf NULL: *size = bytes
 * needed.  Restore checks the shape and CSR and returns the tick. */
int ss_prog_save(ss_prog *prog, int64_t tick, void *buf, int64_t cap, int64_t *size);
int ss_prog_restore(ss_prog *prog, const void *buf, int64_t size, int64_t *tick);

#ifdef __cplusplus
}
#endif
#endif /* SAGESTEP_H */
