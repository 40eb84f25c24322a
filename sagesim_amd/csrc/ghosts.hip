// Ghost topology of one partition on the GPU: replaces
// discover_ghost_topology (gpu_kernels.py:35-88), which flattens the local
// neighbour lists, looks every neighbour id up in a dense rank table and takes
// np.unique of the ids owned by another rank, and the receive side of
// build_communication_maps (gpu_kernels.py:521-850), which groups those ids by
// owner rank (np.unique per rank, :610-617).
//
// On the device: one pass over the neighbour entries sets a bit per remote id
// in a bitmap over the id range (so the set comes out sorted and unique), a
// scan of the words' populations places every id, and a stable radix sort by
// owner rank of the ghost positions gives the per-rank lists in id order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdio>
#include <vector>

#include "../../include/sagestep.h"
#include "ss_runtime.h"

namespace {

__global__ void ghost_mark_kernel(const int64_t *__restrict__ nbr, int64_t nnz, const int32_t *__restrict__ owner,
                                  int64_t n_owner, int32_t my_rank, uint32_t *__restrict__ bits) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += stride) {
        const int64_t id = nbr[e];
        if (id < 0 || id >= n_owner) continue;           // padding (-1) and ids outside the rank table
        const int32_t r = owner[id];
        if (r < 0 || r == my_rank) continue;
        const uint32_t bit = 1u << (id & 31);
        // most repeated ids are already marked: read before the atomic
        if ((bits[id >> 5] & bit) == 0u) atomicOr(&bits[id >> 5], bit);
    }
}

__global__ void ghost_popc_kernel(const uint32_t *__restrict__ bits, int64_t n_words, int32_t *__restrict__ cnt) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < n_words) cnt[w] = __popc(bits[w]);
}

// ids of the set bits of word w at positions pos[w] ..; their owner ranks;
// ranks outside [0, n_ranks) are counted in *bad (an error: the host stops)
__global__ void ghost_emit_kernel(const uint32_t *__restrict__ bits, int64_t n_words, const int32_t *__restrict__ pos,
                                  const int32_t *__restrict__ owner, int32_t n_ranks, int64_t *__restrict__ ids,
                                  int32_t *__restrict__ rank, int32_t *__restrict__ index, int32_t *bad) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_words) return;
    uint32_t b = bits[w];
    int32_t p = pos[w];
    while (b) {
        const int j = __ffs(b) - 1;
        b &= b - 1u;
        const int64_t id = w * 32 + j;
        const int32_t r = owner[id];
        ids[p] = id;
        index[p] = p;
        if (r >= n_ranks) {
            atomicAdd(bad, 1);
            rank[p] = 0;
        } else {
            rank[p] = r;
        }
        ++p;
    }
}

// Receive offsets from the ranks sorted ascending: offsets[q] is the first
// position whose rank is >= q (offsets[n_ranks] = total).  Each position
// where the rank steps up writes the offsets of the ranks it steps over, so
// no counter is shared (per-rank atomic counts serialise on a few addresses).
__global__ void ghost_offsets_kernel(const int32_t *__restrict__ sorted_rank, int64_t total, int32_t n_ranks,
                                     int32_t *__restrict__ offsets) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p > total) return;
    const int32_t lo = p == 0 ? -1 : sorted_rank[p - 1];
    const int32_t hi = p == total ? n_ranks : sorted_rank[p];
    for (int32_t q = lo + 1; q <= hi; ++q) offsets[q] = (int32_t)p;
}

int bits_for(int32_t n) {
    int b = 1;
    while ((1ll << b) < (int64_t)n) ++b;
    return b;
}

int fail(ss_ctx *ctx, int code, const char *msg) {
    ctx->err = msg;
    return code;
}

// device scratch of one discovery, freed by the destructor
struct Scratch {
    std::vector<void *> p;
    ~Scratch() {
        for (void *q : p)
            if (q) (void)hipFree(q);
    }
    template <class T>
    hipError_t get(T **out, size_t count) {
        *out = nullptr;
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, sizeof(T) * (count ? count : 1));
        if (e == hipSuccess) {
            p.push_back(q);
            *out = static_cast<T *>(q);
        }
        return e;
    }
};

}  // namespace

extern "C" int ss_ghost_discover(ss_ctx *ctx, const int64_t *nbr_ids, int64_t nnz, const int32_t *owner,
                                 int64_t n_owner, int32_t my_rank, int32_t n_ranks, int64_t *ghost_ids,
                                 int32_t *recv_offsets, int32_t *recv_index, int64_t cap, int64_t *n_ghost) {
    if (!ctx) return SS_ERR_INVALID;
    if (!n_ghost || nnz < 0 || n_owner < 0 || (nnz && !nbr_ids) || (n_owner && !owner))
        return fail(ctx, SS_ERR_INVALID, "ss_ghost_discover: NULL argument or negative size");
    // ghost positions and counts are int32: at most INT32_MAX possible ghost ids
    if (n_ranks < 1 || n_ranks > (1 << 20) || n_owner > (int64_t)INT32_MAX)
        return fail(ctx, SS_ERR_INVALID, "ss_ghost_discover: n_ranks in [1, 2^20], n_owner at most 2^31 - 1");
    if (ghost_ids && (!recv_offsets || !recv_index))
        return fail(ctx, SS_ERR_INVALID, "ss_ghost_discover: recv_offsets and recv_index are needed with ghost_ids");
    (void)hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    const int64_t n_words = (n_owner + 31) / 32;
    Scratch s;
    int64_t *d_nbr = nullptr;
    int32_t *d_owner = nullptr, *d_cnt = nullptr, *d_pos = nullptr, *d_misc = nullptr;
    uint32_t *d_bits = nullptr;
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = s.get(&d_nbr, (size_t)nnz);
    if (e == hipSuccess) e = s.get(&d_owner, (size_t)n_owner);
    if (e == hipSuccess) e = s.get(&d_bits, (size_t)n_words);
    if (e == hipSuccess) e = s.get(&d_cnt, (size_t)n_words + 1);
    if (e == hipSuccess) e = s.get(&d_pos, (size_t)n_words + 1);
    if (e == hipSuccess) e = s.get(&d_misc, 1);
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(d_nbr, nbr_ids, sizeof(int64_t) * nnz, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && n_owner)
        e = hipMemcpyAsync(d_owner, owner, sizeof(int32_t) * n_owner, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(d_bits, 0, sizeof(uint32_t) * (size_t)std::max<int64_t>(n_words, 1), st);
    if (e == hipSuccess) e = hipMemsetAsync(d_cnt, 0, sizeof(int32_t) * (size_t)(n_words + 1), st);
    if (e == hipSuccess) e = hipMemsetAsync(d_misc, 0, sizeof(int32_t), st);
    if (e == hipSuccess && nnz) {
        const int grid = (int)std::min<int64_t>((nnz + 255) / 256, (int64_t)ctx->num_cus * 16);
        hipLaunchKernelGGL(ghost_mark_kernel, dim3(grid), dim3(256), 0, st, d_nbr, nnz, d_owner, n_owner, my_rank,
                           d_bits);
        e = hipGetLastError();
    }
    if (e == hipSuccess && n_words) {
        hipLaunchKernelGGL(ghost_popc_kernel, dim3((unsigned)((n_words + 255) / 256)), dim3(256), 0, st, d_bits,
                           n_words, d_cnt);
        e = hipGetLastError();
    }
    // positions: exclusive sum over n_words + 1 counts (the last is the total)
    size_t tmp_bytes = 0;
    char *d_tmp = nullptr;
    if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, d_cnt, d_pos, (int)(n_words + 1), st);
    if (e == hipSuccess) e = s.get(&d_tmp, tmp_bytes);
    if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp_bytes, d_cnt, d_pos, (int)(n_words + 1), st);
    int32_t total = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&total, d_pos + n_words, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        char b[256];
        std::snprintf(b, sizeof b, "ss_ghost_discover: %s", hipGetErrorString(e));
        return fail(ctx, e == hipErrorOutOfMemory ? SS_ERR_OOM : SS_ERR_HIP, b);
    }
    *n_ghost = total;
    if (!ghost_ids) return SS_OK;
    if (cap < total) return fail(ctx, SS_ERR_INVALID, "ss_ghost_discover: cap is below the number of ghosts");
    int64_t *d_ids = nullptr;
    int32_t *d_rank = nullptr, *d_rank2 = nullptr, *d_idx = nullptr, *d_idx2 = nullptr, *d_off = nullptr;
    e = s.get(&d_ids, (size_t)total);
    if (e == hipSuccess) e = s.get(&d_off, (size_t)n_ranks + 1);
    if (e == hipSuccess) e = s.get(&d_rank, (size_t)total);
    if (e == hipSuccess) e = s.get(&d_rank2, (size_t)total);
    if (e == hipSuccess) e = s.get(&d_idx, (size_t)total);
    if (e == hipSuccess) e = s.get(&d_idx2, (size_t)total);
    if (e == hipSuccess && n_words) {
        hipLaunchKernelGGL(ghost_emit_kernel, dim3((unsigned)((n_words + 255) / 256)), dim3(256), 0, st, d_bits, n_words,
                           d_pos, d_owner, n_ranks, d_ids, d_rank, d_idx, d_misc);
        e = hipGetLastError();
    }
    // ghost positions grouped by owner rank, id order kept (the sort is stable)
    if (e == hipSuccess && total) {
        size_t sb = 0;
        e = hipcub::DeviceRadixSort::SortPairs(nullptr, sb, d_rank, d_rank2, d_idx, d_idx2, (int)total, 0,
                                               bits_for(n_ranks), st);
        char *d_sort = nullptr;
        if (e == hipSuccess) e = s.get(&d_sort, sb);
        if (e == hipSuccess)
            e = hipcub::DeviceRadixSort::SortPairs(d_sort, sb, d_rank, d_rank2, d_idx, d_idx2, (int)total, 0,
                                                   bits_for(n_ranks), st);
    }
    if (e == hipSuccess) {
        // ranks sorted into d_rank2 (total == 0: every offset is 0)
        hipLaunchKernelGGL(ghost_offsets_kernel, dim3((unsigned)((total + 1 + 255) / 256)), dim3(256), 0, st,
                           total ? d_rank2 : d_rank, (int64_t)total, n_ranks, d_off);
        e = hipGetLastError();
    }
    int32_t bad = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&bad, d_misc, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(recv_offsets, d_off, sizeof(int32_t) * ((size_t)n_ranks + 1), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && total) e = hipMemcpyAsync(ghost_ids, d_ids, sizeof(int64_t) * total, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && total)
        e = hipMemcpyAsync(recv_index, d_idx2, sizeof(int32_t) * total, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        char b[256];
        std::snprintf(b, sizeof b, "ss_ghost_discover: %s", hipGetErrorString(e));
        return fail(ctx, e == hipErrorOutOfMemory ? SS_ERR_OOM : SS_ERR_HIP, b);
    }
    if (bad) return fail(ctx, SS_ERR_RANGE, "ss_ghost_discover: an owner rank is >= n_ranks");
    return SS_OK;
}
