// Kernels around the tick: CSR symmetry check, bitmap rebuilds, id -> slot
// lookups in the ghost map, the halo refresh and packing, census, and the
// two probes the measurement uses (a streaming copy and a random gather).
// The tick kernel itself and its direction-state kernels are in
// tick_kernels.hip.
#include <algorithm>

#include "ss_device.h"
#include "ss_runtime.h"

namespace {

// 64-bit finaliser of splitmix64 (Steele, Lea, Flood 2014).
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Symmetry of a CSR as a multiset test: the sums of two independent hashes
// over all entries (i, j) equal those over the mirrored entries (j, i) when
// every entry has its mirror.  One thread per row.
__global__ void symcheck_kernel(const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col, int64_t n,
                                unsigned long long *out) {
    unsigned long long a0 = 0, a1 = 0, b0 = 0, b1 = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = row_ptr[i], e = row_ptr[i + 1];
#pragma unroll 4
        for (int64_t j = b; j < e; ++j) {
            const uint64_t c = (uint32_t)col[j];
            const uint64_t fwd = ((uint64_t)i << 32) | c, bwd = (c << 32) | (uint64_t)i;
            a0 += mix64(fwd ^ 0x9E3779B97F4A7C15ull);
            a1 += mix64(fwd * 0xD6E8FEB86659FD93ull + 0x2545F4914F6CDD1Dull);
            b0 += mix64(bwd ^ 0x9E3779B97F4A7C15ull);
            b1 += mix64(bwd * 0xD6E8FEB86659FD93ull + 0x2545F4914F6CDD1Dull);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a0 += __shfl_down(a0, off); a1 += __shfl_down(a1, off);
        b0 += __shfl_down(b0, off); b1 += __shfl_down(b1, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&out[0], a0); atomicAdd(&out[1], a1);
        atomicAdd(&out[2], b0); atomicAdd(&out[3], b1);
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

// ---- id -> slot lookups (gpu_kernels.py:208-215): probe from the home slot
// until the id or an empty slot ----
__device__ __forceinline__ int32_t map_find(const unsigned long long *__restrict__ keys,
                                            const int32_t *__restrict__ vals, int64_t cap, int64_t id) {
    int64_t s = ss_map_home(id, cap);
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

__global__ void fill_i32_kernel(int32_t *p, int64_t n, int32_t v) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) p[g] = v;
}

// Ghost refresh (gpu_kernels.py:1030-1036 writes ghosts into the read and the
// write buffer): record h of the halo goes to slot slots[h].
// Ghost refresh in slot order: slot base + g takes record rec[g] of the halo
// (-1: not bound, state kept) into both state buffers, and the bitmap words
// of the slots [base, n_total) are rebuilt from the states of the buffer the
// next tick reads (`cur`, `bits`).  base is a multiple of 64.  One gather per
// slot and whole-word writes, where a scatter in record order would write
// single bytes all over the slots.
// PACKED: the halo holds 2-bit state codes, record h in bits 2 (h % 4) and
// 2 (h % 4) + 1 of byte h / 4.
template <bool PACKED>
__global__ void halo_gather_kernel(int8_t *s0, int8_t *s1, const int8_t *cur, uint64_t *__restrict__ bits,
                                   const int32_t *__restrict__ rec, const int8_t *__restrict__ vals, int64_t base,
                                   int64_t n_total, int8_t inf) {
    // four slots per thread (base is a multiple of 64, the record list is
    // padded to whole groups of four with -1): one 16-byte record load, four
    // value reads, one 4-byte store per state buffer; a wavefront covers 256
    // slots, four bitmap words
    const int lane = threadIdx.x & 63;
    const int64_t i = base + 4 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    uint32_t nib = 0u;                                   // infectious bits of the four slots
    if (i < n_total) {
        const int4 h = *(const int4 *)(rec + (i - base));
        const int32_t hh[4] = {h.x, h.y, h.z, h.w};
        uint32_t v4 = 0u;
        bool any_rec = false, all_rec = true;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            any_rec |= hh[q] >= 0;
            all_rec &= hh[q] >= 0;
        }
        if (!all_rec) {
            // slots without a record keep their state; a whole group in range
            // reads it as one word
            if (i + 3 < n_total) {
                v4 = *(const uint32_t *)(cur + i);
            } else {
                for (int q = 0; q < 4 && i + q < n_total; ++q) v4 |= (uint32_t)(uint8_t)cur[i + q] << (8 * q);
            }
        }
        if (any_rec) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (hh[q] >= 0) {
                    uint32_t v;
                    if constexpr (PACKED)
                        v = ((uint32_t)(uint8_t)vals[hh[q] >> 2] >> ((hh[q] & 3) * 2)) & 3u;
                    else
                        v = (uint32_t)(uint8_t)vals[hh[q]];
                    v4 = (v4 & ~(0xFFu << (8 * q))) | v << (8 * q);
                }
            }
            if (all_rec && i + 3 < n_total) {
                *(uint32_t *)(s0 + i) = v4;
                *(uint32_t *)(s1 + i) = v4;
            } else {
                // only the slots with a record are written
                for (int q = 0; q < 4 && i + q < n_total; ++q) {
                    if (hh[q] < 0) continue;
                    const int8_t v = (int8_t)(v4 >> (8 * q));
                    s0[i + q] = v;
                    s1[i + q] = v;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (i + q < n_total && (int8_t)(v4 >> (8 * q)) == inf) nib |= 1u << q;
    }
    // lanes 16w .. 16w + 15 hold the 64 slots of word w of the wavefront's four
    uint64_t x = (uint64_t)nib << (4 * (lane & 15));
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) x |= (uint64_t)__shfl_xor((long long)x, off);
    if ((lane & 15) == 0 && i < n_total) bits[i >> 6] = x;
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

// The same as 2-bit codes, four records a byte (record h in bits 2 (h % 4)
// and 2 (h % 4) + 1 of byte h / 4); states are taken modulo 4.
__global__ void gather2_kernel(const int8_t *__restrict__ st, const int32_t *__restrict__ slots,
                               uint8_t *__restrict__ out, int64_t n) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (4 * j >= n) return;
    uint32_t b = 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (4 * j + q < n) b |= ((uint32_t)(uint8_t)st[slots[4 * j + q]] & 3u) << (2 * q);
    out[j] = (uint8_t)b;
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

// Plain streaming copy (16 B per lane per step): the known-size reference for
// calibrating hardware traffic counters and the achievable HBM bandwidth.
__global__ void copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// n 4-byte reads at scattered positions of src (n_words words, far larger
// than the caches): the known-count random-gather reference for the traffic
// counters.  One word per thread is written so the reads stay.
__global__ void gather_probe_kernel(const uint32_t *__restrict__ src, int64_t n_words, int64_t n,
                                    uint32_t *__restrict__ out) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t acc = 0u;
    for (int64_t i = g; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t h = (uint64_t)i * 0x9E3779B97F4A7C15ull;
        acc += src[(int64_t)((h >> 17) % (uint64_t)n_words)];
    }
    out[g] = acc;
}

__global__ void fill_u32_kernel(uint32_t *p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v ^ (uint32_t)i;
}

// Range check of an uploaded CSR index array: first_bad = smallest e with
// col[e] outside [0, n_total) (stays ~0 when every entry is in range).
__global__ void check_cols_kernel(const int32_t *__restrict__ col, int64_t nnz, int64_t n_total,
                                  unsigned long long *first_bad) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = col[e];
        if (c < 0 || (int64_t)c >= n_total) atomicMin(first_bad, (unsigned long long)e);
    }
}

inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ---- launchers (host) ----
hipError_t ss_launch_symcheck(const int64_t *row_ptr, const int32_t *col_idx, int64_t n_rows,
                              unsigned long long *out, int grid, hipStream_t s) {
    if (n_rows <= 0) return hipSuccess;
    int g = (int)std::min<int64_t>(grid, (int64_t)grid_for(n_rows, 256));
    hipLaunchKernelGGL(symcheck_kernel, dim3(g), dim3(256), 0, s, row_ptr, col_idx, n_rows, out);
    return hipGetLastError();
}

hipError_t ss_launch_bits(const int8_t *st, uint64_t *bits, int64_t w0, int64_t w1, int64_t n_total, int8_t inf,
                          hipStream_t s) {
    if (w1 <= w0) return hipSuccess;
    hipLaunchKernelGGL(bits_kernel, dim3(grid_for((w1 - w0) * 64, 256)), dim3(256), 0, s, st, bits, w0, w1, n_total,
                       inf);
    return hipGetLastError();
}

hipError_t ss_launch_map_lookup(const unsigned long long *keys, const int32_t *vals, int64_t cap, const int64_t *ids,
                                int32_t *out, int64_t n, int32_t *misses, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(map_lookup_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, keys, vals, cap, ids, out, n,
                       misses);
    return hipGetLastError();
}

hipError_t ss_launch_halo_gather(int8_t *s0, int8_t *s1, const int8_t *cur, uint64_t *bits, const int32_t *rec,
                                 const int8_t *vals, int64_t base, int64_t n_total, int8_t inf, bool packed,
                                 hipStream_t s) {
    if (n_total <= base) return hipSuccess;
    if (packed)
        hipLaunchKernelGGL(halo_gather_kernel<true>, dim3(grid_for(n_total - base, 4 * 256)), dim3(256), 0, s, s0,
                           s1, cur, bits, rec, vals, base, n_total, inf);
    else
        hipLaunchKernelGGL(halo_gather_kernel<false>, dim3(grid_for(n_total - base, 4 * 256)), dim3(256), 0, s, s0,
                           s1, cur, bits, rec, vals, base, n_total, inf);
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

hipError_t ss_launch_gather2(const int8_t *st, const int32_t *slots, uint8_t *out, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather2_kernel, dim3(grid_for((n + 3) / 4, 256)), dim3(256), 0, s, st, slots, out, n);
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

hipError_t ss_launch_copy(const void *src, void *dst, int64_t bytes, int grid, hipStream_t s) {
    hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, s, (const uint4 *)src, (uint4 *)dst, bytes / 16);
    return hipGetLastError();
}

hipError_t ss_launch_gather_probe(const uint32_t *src, int64_t n_words, int64_t n, uint32_t *out, int grid,
                                  hipStream_t s) {
    hipLaunchKernelGGL(gather_probe_kernel, dim3(grid), dim3(256), 0, s, src, n_words, n, out);
    return hipGetLastError();
}

hipError_t ss_launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, int grid, hipStream_t s) {
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t ss_launch_check_cols(const int32_t *col, int64_t nnz, int64_t n_total, unsigned long long *first_bad,
                                int grid, hipStream_t s) {
    if (nnz <= 0) return hipSuccess;
    hipLaunchKernelGGL(check_cols_kernel, dim3(grid), dim3(256), 0, s, col, nnz, n_total, first_bad);
    return hipGetLastError();
}
