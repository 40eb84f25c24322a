// Fused per-tick step kernels for the compartment (SIR/SEIR) workloads on
// gfx950, plus the bitmap, hash-map and halo kernels around them.
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
//     Both shortcuts give the same state as the full count.
#include <algorithm>

#include "ss_device.h"
#include "ss_runtime.h"

namespace {

// Everything one tile needs from its replica, as wave-uniform scalars (no
// arrays: a per-lane index into a local array would send it to scratch).
// Pointers carry the global address space so the loads are global_load, not flat.
#define SS_G __attribute__((address_space(1)))
struct TickView {
    const SS_G int64_t *row_ptr;
    const SS_G int32_t *col_idx;
    const SS_G uint32_t *rng_lo;
    const SS_G uint32_t *rng_hi;
    const SS_G int8_t *cur;
    SS_G int8_t *nxt;
    const SS_G uint64_t *cb;
    SS_G uint64_t *nb;
    int32_t *history;
    const SS_G uint32_t *inf_T;
    int32_t last;
    uint32_t tT0, tT1, tT2, tT3;
    int8_t tN0, tN1, tN2, tN3;
    int8_t s_code, inf_code, s_target;
    uint32_t seed_lo, seed_hi;
    uint32_t t_lo, t_hi;
    int64_t n_local;
    unsigned long long *counters;   // null, or work counters of this launch
};

// Work done by one lane, flushed to the launch counters when they are on:
// [0] susceptible agents visited, [1] draws, [2] neighbour scans started,
// [3] CSR entries read by scans.
struct Work {
    uint32_t s_agents = 0, draws = 0, scans = 0, edges = 0;
};

__device__ __forceinline__ uint32_t draw(const TickView V, int64_t i) {
    const uint32_t hi = V.rng_hi ? V.rng_hi[i] : 0u;
    return ss_philox4x32_w0(V.t_lo, V.t_hi, V.rng_lo[i], hi, V.seed_lo, V.seed_hi) & SS_U23_MASK;
}

// spontaneous transition of a non-susceptible agent (select chain, no indexing)
__device__ __forceinline__ int8_t transition(const TickView V, int64_t i, int8_t s, Work &w) {
    const uint32_t thr = s == 0 ? V.tT0 : s == 1 ? V.tT1 : s == 2 ? V.tT2 : s == 3 ? V.tT3 : 0u;
    if (thr == 0u) return s;
    ++w.draws;
    if (draw(V, i) >= thr) return s;
    return s == 0 ? V.tN0 : s == 1 ? V.tN1 : s == 2 ? V.tN2 : V.tN3;
}

__device__ __forceinline__ bool infects(const TickView V, uint32_t m, int32_t k) {
    return m < V.inf_T[k < V.last ? k : V.last];
}

// One agent, one lane.  Returns the new state.
__device__ __forceinline__ int8_t step_lane(const TickView V, int64_t i, int8_t s, Work &w) {
    if (s != V.s_code) return transition(V, i, s, w);
    ++w.s_agents;
    const int64_t b = V.row_ptr[i], e = V.row_ptr[i + 1];
    const int64_t deg = e - b;
    if (deg <= 0) return s;
    ++w.draws;
    const uint32_t m = draw(V, i);
    if (!infects(V, m, (int32_t)(deg < V.last ? deg : V.last))) return s;
    ++w.scans;
    const SS_G int32_t *__restrict__ col = V.col_idx;
    const SS_G uint64_t *__restrict__ cb = V.cb;
    int32_t k = 0, first = -1;
    int64_t j = b;
    // four independent gathers in flight before the first test
    for (; j + 4 <= e; j += 4) {
        w.edges += 4;
        const int32_t c0 = col[j], c1 = col[j + 1], c2 = col[j + 2], c3 = col[j + 3];
        const uint64_t w0 = cb[c0 >> 6], w1 = cb[c1 >> 6], w2 = cb[c2 >> 6], w3 = cb[c3 >> 6];
        const bool h0 = (w0 >> (c0 & 63)) & 1ull, h1 = (w1 >> (c1 & 63)) & 1ull;
        const bool h2 = (w2 >> (c2 & 63)) & 1ull, h3 = (w3 >> (c3 & 63)) & 1ull;
        if (!(h0 | h1 | h2 | h3)) continue;
        if (h0) { if (first < 0) first = c0; if (infects(V, m, ++k)) goto infected; }
        if (h1) { if (first < 0) first = c1; if (infects(V, m, ++k)) goto infected; }
        if (h2) { if (first < 0) first = c2; if (infects(V, m, ++k)) goto infected; }
        if (h3) { if (first < 0) first = c3; if (infects(V, m, ++k)) goto infected; }
    }
    for (; j < e; ++j) {
        ++w.edges;
        const int32_t c = col[j];
        if ((cb[c >> 6] >> (c & 63)) & 1ull) {
            if (first < 0) first = c;
            if (infects(V, m, ++k)) goto infected;
        }
    }
    return s;
infected:
    if (V.history) ss_hist_append(V.history + i * SS_HIST_W, first);
    return V.s_target;
}

// One agent, one wavefront (all 64 lanes call with the same i and s).
__device__ __forceinline__ int8_t step_wave(const TickView V, int64_t i, int8_t s, int lane, Work &w) {
    // counts are kept by lane 0 only (all lanes see the same agent)
    if (s != V.s_code) {
        Work dummy;
        return transition(V, i, s, lane == 0 ? w : dummy);
    }
    if (lane == 0) ++w.s_agents;
    const int64_t b = V.row_ptr[i], e = V.row_ptr[i + 1];
    const int64_t deg = e - b;
    if (deg <= 0) return s;
    if (lane == 0) ++w.draws;
    const uint32_t m = draw(V, i);
    if (!infects(V, m, (int32_t)(deg < V.last ? deg : V.last))) return s;
    if (lane == 0) ++w.scans;
    const SS_G int32_t *__restrict__ col = V.col_idx;
    int32_t k = 0, first = -1;
    for (int64_t base = b; base < e; base += 64) {
        const int64_t j = base + lane;
        if (lane == 0) w.edges += (uint32_t)(e - base < 64 ? e - base : 64);
        int32_t c = 0;
        bool hit = false;
        if (j < e) {
            c = col[j];
            hit = (V.cb[c >> 6] >> (c & 63)) & 1ull;
        }
        const uint64_t bal = __ballot(hit);
        if (bal) {
            if (first < 0) first = __shfl(c, __ffsll((unsigned long long)bal) - 1);
            k += __popcll(bal);
            if (infects(V, m, k)) {
                if (lane == 0 && V.history) ss_hist_append(V.history + i * SS_HIST_W, first);
                return V.s_target;
            }
        }
    }
    return s;
}

__device__ __forceinline__ void flush_work(const TickView V, const Work &w) {
    if (!V.counters) return;
    uint32_t a = w.s_agents, d = w.draws, sc = w.scans, ed = w.edges;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off);
        d += __shfl_down(d, off);
        sc += __shfl_down(sc, off);
        ed += __shfl_down(ed, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&V.counters[0], (unsigned long long)a);
        atomicAdd(&V.counters[1], (unsigned long long)d);
        atomicAdd(&V.counters[2], (unsigned long long)sc);
        atomicAdd(&V.counters[3], (unsigned long long)ed);
    }
}

template <int APL>
__device__ __forceinline__ void light_tile(const TickView V, int64_t begin, int64_t end) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    Work w;
#pragma unroll
    for (int a = 0; a < APL; ++a) {
        const int64_t g0 = begin + (int64_t)(wv * APL + a) * 64;
        if (g0 >= end) break;                        // wave-uniform
        const int64_t i = g0 + lane;
        bool inf = false;
        if (i < end) {
            const int8_t s = V.cur[i];
            const int8_t ns = i < V.n_local ? step_lane(V, i, s, w) : s;
            V.nxt[i] = ns;
            inf = ns == V.inf_code;
        }
        const uint64_t word = __ballot(inf);
        if (lane == 0) V.nb[g0 >> 6] = word;
    }
    flush_work(V, w);
}

__device__ __forceinline__ void heavy_tile(const TickView V, int64_t begin, int64_t end,
                                           unsigned long long *sword) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    Work w;
    if (threadIdx.x == 0) *sword = 0ull;
    __syncthreads();
    for (int q = wv; q < 64; q += nw) {
        const int64_t i = begin + q;
        if (i >= end) break;                         // wave-uniform
        const int8_t s = V.cur[i];
        const int8_t ns = i < V.n_local ? step_wave(V, i, s, lane, w) : s;
        if (lane == 0) {
            V.nxt[i] = ns;
            if (ns == V.inf_code) atomicOr(sword, 1ull << q);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) V.nb[begin >> 6] = *sword;
    flush_work(V, w);
    __syncthreads();
}

// Launch step k of a tick loop: every live replica r runs tick reps[r].tick0 + k.
template <int APL>
__global__ __launch_bounds__(SS_BLOCK) void tick_kernel(const ss_rep_dev *__restrict__ reps,
                                                        const ss_tile *__restrict__ tiles, int n_tiles,
                                                        int k) {
    __shared__ unsigned long long sword;
    for (int tix = blockIdx.x; tix < n_tiles; tix += gridDim.x) {
        const ss_tile *Tp = tiles + tix;
        const int32_t rep = Tp->rep, kind = Tp->kind;
        const int64_t begin = Tp->begin, end = Tp->end;
        const ss_rep_dev *R = reps + rep;
        const int64_t t = R->tick0 + k;
        if (t >= R->tick_end) continue;              // block-uniform
        const bool odd = (t & 1) != 0;
        TickView V;
        V.row_ptr = (const SS_G int64_t *)R->row_ptr;
        V.col_idx = (const SS_G int32_t *)R->col_idx;
        V.rng_lo = (const SS_G uint32_t *)R->rng_lo;
        V.rng_hi = (const SS_G uint32_t *)R->rng_hi;
        V.cur = (const SS_G int8_t *)(odd ? R->state[1] : R->state[0]);
        V.nxt = (SS_G int8_t *)(odd ? R->state[0] : R->state[1]);
        V.cb = (const SS_G uint64_t *)(odd ? R->ibits[1] : R->ibits[0]);
        V.nb = (SS_G uint64_t *)(odd ? R->ibits[0] : R->ibits[1]);
        V.history = R->history;
        V.inf_T = (const SS_G uint32_t *)R->inf_T;
        V.last = R->inf_T_last;
        V.tT0 = R->trans_T[0]; V.tT1 = R->trans_T[1]; V.tT2 = R->trans_T[2]; V.tT3 = R->trans_T[3];
        V.tN0 = R->trans_next[0]; V.tN1 = R->trans_next[1]; V.tN2 = R->trans_next[2]; V.tN3 = R->trans_next[3];
        V.s_code = R->s_code;
        V.inf_code = R->inf_code;
        V.s_target = R->s_target;
        V.seed_lo = R->seed_lo;
        V.seed_hi = R->seed_hi;
        V.t_lo = (uint32_t)t;
        V.t_hi = (uint32_t)((uint64_t)t >> 32);
        V.n_local = R->n_local;
        V.counters = R->counters;
        if (kind == SS_TILE_HEAVY)
            heavy_tile(V, begin, end, &sword);
        else
            light_tile<APL>(V, begin, end);
    }
}

// Rebuild bitmap words [w0, w1) of one state buffer.
__global__ void bits_kernel(const int8_t *__restrict__ st, uint64_t *__restrict__ bits, int64_t w0,
                            int64_t w1, int64_t n_total, int8_t inf) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = w0 * 64 + g;
    const bool on = i < w1 * 64 && i < n_total && st[i] == inf;
    const uint64_t word = __ballot(on);
    if ((threadIdx.x & 63) == 0 && (i >> 6) < w1) bits[i >> 6] = word;
}

// ---- open-addressing id -> slot map (gpu_kernels.py:91-260 layout) ----
// slot of id: floor-mod of the wrapped int64 product id * 2654435761 by the
// capacity, as the reference's numpy expression computes it
// (gpu_kernels.py:112-114, :146); linear probing; key -1 empty, -2 deleted.
__device__ __forceinline__ int64_t map_home(int64_t id, int64_t cap) {
    const int64_t h = (int64_t)((uint64_t)id * 2654435761ull);
    int64_t s = h % cap;
    return s < 0 ? s + cap : s;
}

__global__ void map_insert_kernel(unsigned long long *keys, int32_t *vals, int64_t cap,
                                  const int64_t *__restrict__ ids, const int32_t *__restrict__ slots,
                                  int64_t n, int32_t *overflow) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int64_t id = ids[g];
    int64_t s = map_home(id, cap);
    for (int64_t probe = 0; probe < cap; ++probe) {
        unsigned long long prev = atomicCAS(&keys[s], (unsigned long long)(-1ll), (unsigned long long)id);
        if (prev == (unsigned long long)(-1ll) || prev == (unsigned long long)id) {
            // equal keys: the lowest slot value wins, independent of thread order
            atomicMin(&vals[s], slots[g]);
            return;
        }
        if (prev == (unsigned long long)(-2ll)) {
            prev = atomicCAS(&keys[s], (unsigned long long)(-2ll), (unsigned long long)id);
            if (prev == (unsigned long long)(-2ll) || prev == (unsigned long long)id) {
                atomicMin(&vals[s], slots[g]);
                return;
            }
        }
        s = s + 1 == cap ? 0 : s + 1;
    }
    atomicAdd(overflow, 1);
}

__device__ __forceinline__ int32_t map_find(const unsigned long long *__restrict__ keys,
                                            const int32_t *__restrict__ vals, int64_t cap, int64_t id) {
    int64_t s = map_home(id, cap);
    for (int64_t probe = 0; probe < cap; ++probe) {
        const unsigned long long k = keys[s];
        if (k == (unsigned long long)id) return vals[s];
        if (k == (unsigned long long)(-1ll)) return -1;
        s = s + 1 == cap ? 0 : s + 1;
    }
    return -1;
}

__global__ void map_lookup_kernel(const unsigned long long *__restrict__ keys, const int32_t *__restrict__ vals,
                                  int64_t cap, const int64_t *__restrict__ ids, int32_t *__restrict__ out,
                                  int64_t n, int32_t *misses) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int32_t v = map_find(keys, vals, cap, ids[g]);
    out[g] = v;
    if (v < 0 && misses) atomicAdd(misses, 1);
}

__global__ void map_remove_kernel(unsigned long long *keys, int32_t *vals, int64_t cap,
                                  const int64_t *__restrict__ ids, int64_t n, int32_t *removed) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int64_t id = ids[g];
    int64_t s = map_home(id, cap);
    for (int64_t probe = 0; probe < cap; ++probe) {
        const unsigned long long k = keys[s];
        if (k == (unsigned long long)(-1ll)) return;
        if (k == (unsigned long long)id) {
            if (atomicCAS(&keys[s], (unsigned long long)id, (unsigned long long)(-2ll)) == (unsigned long long)id) {
                vals[s] = INT32_MAX;   // free slots hold INT32_MAX so a later claim can atomicMin
                atomicAdd(removed, 1);
            }
            return;
        }
        s = s + 1 == cap ? 0 : s + 1;
    }
}

__global__ void fill_i64_kernel(unsigned long long *p, int64_t n, unsigned long long v) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) p[g] = v;
}
__global__ void fill_i32_kernel(int32_t *p, int64_t n, int32_t v) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) p[g] = v;
}

// Ghost refresh (gpu_kernels.py:1030-1036 writes ghosts into the read and the
// write buffer): record h of the halo goes to slot slots[h].
__global__ void halo_scatter_kernel(int8_t *s0, int8_t *s1, const int32_t *__restrict__ slots,
                                    const int8_t *__restrict__ vals, int64_t n) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int32_t s = slots[g];
    if (s < 0) return;
    const int8_t v = vals[g];
    s0[s] = v;
    s1[s] = v;
}

// Ghost refresh keyed by id: resolve through the map, then scatter.
__global__ void halo_unpack_ids_kernel(const unsigned long long *__restrict__ keys,
                                       const int32_t *__restrict__ mvals, int64_t cap,
                                       const int64_t *__restrict__ ids, const int8_t *__restrict__ vals,
                                       int64_t n, int8_t *s0, int8_t *s1, int32_t *misses) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int32_t s = map_find(keys, mvals, cap, ids[g]);
    if (s < 0) { atomicAdd(misses, 1); return; }
    s0[s] = vals[g];
    s1[s] = vals[g];
}

// Pack (gpu_kernels.py:852-909 gathers rows of the send list): out[h] = st[slots[h]].
__global__ void gather_kernel(const int8_t *__restrict__ st, const int32_t *__restrict__ slots,
                              int8_t *__restrict__ out, int64_t n) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) out[g] = st[slots[g]];
}

// Census of one state buffer: out[s] += #agents in state s (s < 4) over the
// local slots, out[4] += sum of degrees of agents in state s_code.
__global__ void stats_kernel(const int8_t *__restrict__ st, const int64_t *__restrict__ row_ptr, int64_t n_local,
                             int8_t s_code, unsigned long long *out) {
    unsigned long long c0 = 0, c1 = 0, c2 = 0, c3 = 0, dsum = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_local;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int8_t s = st[i];
        c0 += s == 0; c1 += s == 1; c2 += s == 2; c3 += s == 3;
        if (s == s_code) dsum += (unsigned long long)(row_ptr[i + 1] - row_ptr[i]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        c0 += __shfl_down(c0, off); c1 += __shfl_down(c1, off); c2 += __shfl_down(c2, off);
        c3 += __shfl_down(c3, off); dsum += __shfl_down(dsum, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&out[0], c0); atomicAdd(&out[1], c1); atomicAdd(&out[2], c2);
        atomicAdd(&out[3], c3); atomicAdd(&out[4], dsum);
    }
}

inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ---- launchers (host) ----
hipError_t ss_launch_tick(const ss_rep_dev *reps, const ss_tile *tiles, int n_tiles, int k, int apl,
                          int grid, hipStream_t st) {
    if (n_tiles <= 0) return hipSuccess;
    switch (apl) {
    case 1: hipLaunchKernelGGL(tick_kernel<1>, dim3(grid), dim3(SS_BLOCK), 0, st, reps, tiles, n_tiles, k); break;
    case 2: hipLaunchKernelGGL(tick_kernel<2>, dim3(grid), dim3(SS_BLOCK), 0, st, reps, tiles, n_tiles, k); break;
    case 4: hipLaunchKernelGGL(tick_kernel<4>, dim3(grid), dim3(SS_BLOCK), 0, st, reps, tiles, n_tiles, k); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t ss_occupancy_tick(int apl, int *blocks_per_cu) {
    switch (apl) {
    case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, tick_kernel<1>, SS_BLOCK, 0);
    case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, tick_kernel<2>, SS_BLOCK, 0);
    case 4: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, tick_kernel<4>, SS_BLOCK, 0);
    default: return hipErrorInvalidValue;
    }
}

hipError_t ss_launch_bits(const int8_t *st, uint64_t *bits, int64_t w0, int64_t w1, int64_t n_total, int8_t inf,
                          hipStream_t s) {
    if (w1 <= w0) return hipSuccess;
    hipLaunchKernelGGL(bits_kernel, dim3(grid_for((w1 - w0) * 64, 256)), dim3(256), 0, s, st, bits, w0, w1, n_total,
                       inf);
    return hipGetLastError();
}

hipError_t ss_launch_map_clear(unsigned long long *keys, int32_t *vals, int64_t cap, hipStream_t s) {
    hipLaunchKernelGGL(fill_i64_kernel, dim3(grid_for(cap, 256)), dim3(256), 0, s, keys, cap,
                       (unsigned long long)(-1ll));
    hipLaunchKernelGGL(fill_i32_kernel, dim3(grid_for(cap, 256)), dim3(256), 0, s, vals, cap, (int32_t)INT32_MAX);
    return hipGetLastError();
}

hipError_t ss_launch_map_insert(unsigned long long *keys, int32_t *vals, int64_t cap, const int64_t *ids,
                                const int32_t *slots, int64_t n, int32_t *overflow, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(map_insert_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, keys, vals, cap, ids, slots, n,
                       overflow);
    return hipGetLastError();
}

hipError_t ss_launch_map_lookup(const unsigned long long *keys, const int32_t *vals, int64_t cap, const int64_t *ids,
                                int32_t *out, int64_t n, int32_t *misses, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(map_lookup_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, keys, vals, cap, ids, out, n,
                       misses);
    return hipGetLastError();
}

hipError_t ss_launch_map_remove(unsigned long long *keys, int32_t *vals, int64_t cap, const int64_t *ids, int64_t n,
                                int32_t *removed, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(map_remove_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, keys, vals, cap, ids, n, removed);
    return hipGetLastError();
}

hipError_t ss_launch_halo_scatter(int8_t *s0, int8_t *s1, const int32_t *slots, const int8_t *vals, int64_t n,
                                  hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(halo_scatter_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, s0, s1, slots, vals, n);
    return hipGetLastError();
}

hipError_t ss_launch_halo_unpack_ids(const unsigned long long *keys, const int32_t *mvals, int64_t cap,
                                     const int64_t *ids, const int8_t *vals, int64_t n, int8_t *s0, int8_t *s1,
                                     int32_t *misses, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(halo_unpack_ids_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, keys, mvals, cap, ids, vals,
                       n, s0, s1, misses);
    return hipGetLastError();
}

hipError_t ss_launch_gather(const int8_t *st, const int32_t *slots, int8_t *out, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, st, slots, out, n);
    return hipGetLastError();
}

hipError_t ss_launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_i32_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t ss_launch_stats(const int8_t *st, const int64_t *row_ptr, int64_t n_local, int8_t s_code,
                           unsigned long long *out, int grid, hipStream_t s) {
    if (n_local <= 0) return hipSuccess;
    int g = (int)std::min<int64_t>(grid, (int64_t)grid_for(n_local, 256));
    hipLaunchKernelGGL(stats_kernel, dim3(g), dim3(256), 0, s, st, row_ptr, n_local, s_code, out);
    return hipGetLastError();
}
