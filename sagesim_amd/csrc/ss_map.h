// Id -> slot map of one partition's ghosts: home slot and launchers
// (hashmap.hip).  Layout and probing of the reference's GPUHashMap
// (gpu_kernels.py:91-260).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SS_MAP_INSERT 0
#define SS_MAP_REMOVE 1
#define SS_MAP_RAW 2

// floor-mod of the wrapped int64 product id * 2654435761 by the capacity, as
// the reference's numpy expression computes it (gpu_kernels.py:112-114, :146)
__host__ __device__ __forceinline__ int64_t ss_map_home(int64_t id, int64_t cap) {
    const int64_t h = (int64_t)((uint64_t)id * 2654435761ull);
    int64_t s = h % cap;
    return s < 0 ? s + cap : s;
}

// keys = -1, values = -1 over the whole table
hipError_t ss_launch_map_clear(unsigned long long *keys, int32_t *vals, int64_t cap, hipStream_t s);
// build: probe[i] = home of ids[i], claim[0 .. cap) = INT32_MAX
hipError_t ss_launch_map_build_init(const int64_t *ids, int64_t n, int64_t cap, int64_t *probe, int32_t *claim,
                                    hipStream_t s);
// one placement round of the build; *remaining += entries still unplaced after it
hipError_t ss_launch_map_build_round(unsigned long long *keys, int32_t *vals, int64_t cap, const int64_t *ids,
                                     const int32_t *slot_vals, int64_t n, int64_t *probe, int32_t *claim,
                                     int32_t *remaining, hipStream_t s);
// insert / remove / raw re-insert of n entries in input order; out[0] +=
// entries added or removed, out[1] += entries that found no slot
hipError_t ss_launch_map_seq(unsigned long long *keys, int32_t *vals, int64_t cap, int op, const int64_t *ids,
                             const int32_t *slot_vals, int64_t n, int32_t *out, hipStream_t s);
