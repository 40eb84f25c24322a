// ss_halo_pack_noalloc: the send-side halo pack with no device allocation,
// for callers that cannot allocate (code under HIP graph capture, per-tick
// exchange loops that must not call hipMalloc).
//
// Its staging lives in `scratch`, a file-scope device array of fixed size
// that each GPU holds until the process ends.  Host code never passes the
// array's name to a kernel or a copy (in host code it is only a stand-in for
// the symbol): the device address of each GPU's copy is taken once per device
// with hipGetSymbolAddress, with that device current, on the entry point's
// first call there, and used from then on.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (two threads, or two contexts on the same device) race on it.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/sagestep.h"
#include "ss_runtime.h"

namespace {

constexpr size_t kSlotBytes = sizeof(int32_t) * (size_t)SS_NOALLOC_MAX_RECORDS;
constexpr size_t kScratchBytes = kSlotBytes + (size_t)SS_NOALLOC_MAX_RECORDS;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

// device address of this GPU's scratch (the device must be current)
hipError_t scratch_ptr(int device, void **out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_ptr[device] = p;
    }
    *out = g_ptr[device];
    return hipSuccess;
}

int fail(ss_ctx *ctx, int code, const char *msg, long long a = 0, long long b = 0) {
    char buf[256];
    std::snprintf(buf, sizeof buf, msg, a, b);
    ctx->err = buf;
    return code;
}

}  // namespace

extern "C" int ss_halo_pack_noalloc(ss_ctx *ctx, int32_t rep, const int32_t *slots, int64_t n, int8_t *out) {
    if (!ctx) return SS_ERR_INVALID;
    if (rep < 0 || rep >= (int32_t)ctx->reps.size() || !ctx->reps[rep].live || n < 0 || (n && (!slots || !out)))
        return fail(ctx, SS_ERR_INVALID, "bad replica or NULL");
    if (n > SS_NOALLOC_MAX_RECORDS)
        return fail(ctx, SS_ERR_INVALID, "n = %lld records exceeds the %lld this entry point holds", (long long)n,
                    (long long)SS_NOALLOC_MAX_RECORDS);
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return fail(ctx, SS_ERR_INVALID, "device out of range");
    const ss_replica &r = ctx->reps[rep];
    for (int64_t i = 0; i < n; ++i)
        if (slots[i] < 0 || slots[i] >= r.n_total)
            return fail(ctx, SS_ERR_RANGE, "slot %lld outside the replica", (long long)slots[i]);
    if (n == 0) return SS_OK;
    hipError_t e = hipSetDevice(ctx->device);
    void *base = nullptr;
    if (e == hipSuccess) e = scratch_ptr(ctx->device, &base);
    if (e != hipSuccess) return fail(ctx, SS_ERR_HIP, "scratch address: error %lld", (long long)e);
    int32_t *d_slots = static_cast<int32_t *>(base);
    int8_t *d_out = static_cast<int8_t *>(base) + kSlotBytes;
    e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_slots, slots, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = ss_launch_gather(r.state[r.tick & 1], d_slots, d_out, n, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(ctx, SS_ERR_HIP, "halo pack: error %lld", (long long)e);
    return SS_OK;
}
