// Open-addressing id -> slot map of one partition's ghosts, laid out as the
// reference's GPUHashMap (gpu_kernels.py:91-260): keys int64 (-1 empty,
// -2 deleted), values int32 (-1 in empty and deleted slots), home slot the
// floor-mod of the wrapped int64 product id * 2654435761 by the capacity
// (gpu_kernels.py:112-114, :146), linear probing.  Every operation leaves the
// arrays exactly as the reference's CPU mirror would be after the same calls:
//   * build (gpu_kernels.py:122-169) runs the reference's rounds: in every
//     round the remaining entry of lowest input index among those probing an
//     empty slot takes it, all others advance one slot;
//   * insert, remove (gpu_kernels.py:171-206) and the re-insertion of a
//     resize (gpu_kernels.py:221-247) are sequential in the reference, and one
//     entry's probe depends on the writes of the entries before it, so they
//     run in input order on one wavefront, 64 slots of the probe per load.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ss_map.h"

namespace {

constexpr unsigned long long kEmpty = (unsigned long long)(-1ll);
constexpr unsigned long long kDeleted = (unsigned long long)(-2ll);

__global__ void map_fill_kernel(unsigned long long *keys, int32_t *vals, int64_t cap) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < cap) {
        keys[g] = kEmpty;
        vals[g] = -1;
    }
}

// ---- build: the reference's placement rounds ----
__global__ void map_round_init_kernel(const int64_t *__restrict__ ids, int64_t n, int64_t cap,
                                      int64_t *__restrict__ probe, int32_t *__restrict__ claim) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) probe[g] = ss_map_home(ids[g], cap);
    if (g < cap) claim[g] = INT32_MAX;
}

// every remaining entry whose probe slot is empty bids its index for it
__global__ void map_round_claim_kernel(const unsigned long long *__restrict__ keys, int64_t n,
                                       const int64_t *__restrict__ probe, int32_t *claim) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int64_t s = probe[g];
    if (s >= 0 && keys[s] == kEmpty) atomicMin(&claim[s], (int32_t)g);
}

// the lowest bidder of a slot takes it (and clears the bid for the next
// round); every other remaining entry moves to the next slot.  Grid-stride
// over the entries with a bounded grid, the count left summed in LDS: one
// global atomic per block, not one per wavefront on a single address
// (those serialise at the L2 and cost ~1 ms per round over 6M entries).
__global__ __launch_bounds__(256) void map_round_place_kernel(unsigned long long *keys, int32_t *vals, int64_t cap,
                                                              const int64_t *__restrict__ ids,
                                                              const int32_t *__restrict__ slot_vals, int64_t n,
                                                              int64_t *probe, int32_t *claim, int32_t *remaining) {
    __shared__ int32_t block_left;
    if (threadIdx.x == 0) block_left = 0;
    __syncthreads();
    int32_t left = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += stride) {
        const int64_t s = probe[g];
        if (s < 0) continue;
        if (claim[s] == (int32_t)g) {
            keys[s] = (unsigned long long)ids[g];
            vals[s] = slot_vals[g];
            claim[s] = INT32_MAX;
            probe[g] = -1;
        } else {
            probe[g] = s + 1 == cap ? 0 : s + 1;
            ++left;
        }
    }
    if (left) atomicAdd(&block_left, left);
    __syncthreads();
    if (threadIdx.x == 0 && block_left) atomicAdd(remaining, block_left);
}

// ---- sequential operations on one wavefront ----
// op SS_MAP_INSERT: stop at an empty or deleted slot or at the id's own key
//    (gpu_kernels.py:175-179); a slot that does not hold the id yet counts as
//    a new entry (:181-182); key and value are written.
// op SS_MAP_REMOVE: stop at an empty slot (not found) or at the id's key,
//    which becomes deleted with value -1 (gpu_kernels.py:193-206).
// op SS_MAP_RAW: stop at an empty or deleted slot (gpu_kernels.py:242-247).
// out[0] += entries added (insert, raw) or removed (remove); out[1] += ops
// that found no stop slot within one sweep of the table (a full table; the
// reference would probe forever).
__global__ __launch_bounds__(64) void map_seq_kernel(unsigned long long *keys, int32_t *vals, int64_t cap, int op,
                                                     const int64_t *__restrict__ ids,
                                                     const int32_t *__restrict__ slot_vals, int64_t n, int32_t *out) {
    const int lane = threadIdx.x;
    int32_t changed = 0, failed = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t id = ids[i];
        const unsigned long long key = (unsigned long long)id;
        int64_t base = ss_map_home(id, cap);
        int64_t stop = -1;
        for (int64_t swept = 0; swept < cap; swept += 64) {
            int64_t s = base + lane;
            if (s >= cap) s -= cap;            // base < cap and lane < 64 <= cap
            const unsigned long long k = keys[s];
            bool hit;
            if (op == SS_MAP_INSERT) hit = k == kEmpty || k == kDeleted || k == key;
            else if (op == SS_MAP_REMOVE) hit = k == kEmpty || k == key;
            else hit = k == kEmpty || k == kDeleted;
            // slots past one whole sweep of the table do not count
            hit = hit && swept + lane < cap;
            const unsigned long long b = __ballot(hit);
            if (b) {
                const int f = __ffsll((long long)b) - 1;
                stop = __shfl(s, f);
                break;
            }
            base += 64;
            if (base >= cap) base -= cap;
        }
        if (stop < 0) {
            ++failed;
        } else if (lane == 0) {
            const unsigned long long k = keys[stop];
            if (op == SS_MAP_REMOVE) {
                if (k == key) {
                    keys[stop] = kDeleted;
                    vals[stop] = -1;
                    ++changed;
                }
            } else {
                if (op == SS_MAP_RAW || k != key) ++changed;
                keys[stop] = key;
                vals[stop] = slot_vals[i];
            }
        }
        // the next entry's probe must see this entry's write: wait for the
        // store and drop this CU's cached lines of the table
        __threadfence();
    }
    if (lane == 0) {
        out[0] += changed;
        out[1] += failed;
    }
}

inline unsigned grid_of(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

hipError_t ss_launch_map_clear(unsigned long long *keys, int32_t *vals, int64_t cap, hipStream_t s) {
    if (cap <= 0) return hipSuccess;
    hipLaunchKernelGGL(map_fill_kernel, dim3(grid_of(cap, 256)), dim3(256), 0, s, keys, vals, cap);
    return hipGetLastError();
}

hipError_t ss_launch_map_build_init(const int64_t *ids, int64_t n, int64_t cap, int64_t *probe, int32_t *claim,
                                    hipStream_t s) {
    const int64_t m = n > cap ? n : cap;
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(map_round_init_kernel, dim3(grid_of(m, 256)), dim3(256), 0, s, ids, n, cap, probe, claim);
    return hipGetLastError();
}

hipError_t ss_launch_map_build_round(unsigned long long *keys, int32_t *vals, int64_t cap, const int64_t *ids,
                                     const int32_t *slot_vals, int64_t n, int64_t *probe, int32_t *claim,
                                     int32_t *remaining, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(map_round_claim_kernel, dim3(grid_of(n, 256)), dim3(256), 0, s, keys, n, probe, claim);
    const unsigned place_grid = grid_of(n, 256) < 2048u ? grid_of(n, 256) : 2048u;
    hipLaunchKernelGGL(map_round_place_kernel, dim3(place_grid), dim3(256), 0, s, keys, vals, cap, ids,
                       slot_vals, n, probe, claim, remaining);
    return hipGetLastError();
}

hipError_t ss_launch_map_seq(unsigned long long *keys, int32_t *vals, int64_t cap, int op, const int64_t *ids,
                             const int32_t *slot_vals, int64_t n, int32_t *out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(map_seq_kernel, dim3(1), dim3(64), 0, s, keys, vals, cap, op, ids, slot_vals, n, out);
    return hipGetLastError();
}
