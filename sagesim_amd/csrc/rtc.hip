// General step functions: user HIP device code compiled with hipRTC.
//
// Replaces generate_gpu_func + the CuPy JIT of the generated module
// (model.py:2010-2457, :800-865) for models whose step functions are not the
// built-in compartment step: the host side writes one wrapper kernel per
// priority group around the user's __device__ functions, this file compiles
// it for gfx950, loads it, and runs the per-tick launch schedule followed by
// the write-back of double-buffered properties (model.py:2368-2408).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <hip/hiprtc.h>

#include "../../include/sagestep.h"
#include "ss_runtime.h"

namespace {

// the device runtime header, embedded at build time from step_rt.h
const char *kStepRtHeader =
#include "step_rt_embed.inc"
    ;

#define SS_MAXP_HOST 48

struct ProgArgsHost {           // layout of SSProgArgs in step_rt.h
    long long n_agents;
    const long long *row_ptr;
    const int *col_idx;
    const float *agent_ids;
    const float *logical_ids;
    const float *breeds;
    const double *globals;
    void *prop_r[SS_MAXP_HOST];
    void *prop_w[SS_MAXP_HOST];
    int seed;
    const long long *logical_ids64;
};

struct ProgProp {
    int32_t elem_bytes = 0;
    int64_t width = 1;
    bool double_buffered = true;
    void *r = nullptr, *w = nullptr;
    size_t bytes() const { return (size_t)elem_bytes * (size_t)width; }
};

struct Launch {
    int32_t kernel;
    int64_t start, count;
};

}  // namespace

struct ss_prog {
    ss_ctx *ctx = nullptr;
    int64_t n = 0;
    int64_t *row_ptr = nullptr;
    int32_t *col_idx = nullptr;
    int64_t nnz = 0;
    float *agent_ids = nullptr, *logical_ids = nullptr, *breeds = nullptr;
    int64_t *logical_ids64 = nullptr;
    double *globals = nullptr;
    int64_t n_globals = 0;
    std::vector<ProgProp> props;
    ProgArgsHost host{};
    ProgArgsHost *d_args = nullptr;
    hipModule_t module = nullptr;
    std::vector<hipFunction_t> kernels;
    std::vector<Launch> schedule;
    int32_t seed = 0;
    std::string err;
};

static thread_local std::string g_prog_err;

#define P_FAIL(p, code, ...)                                      \
    do {                                                          \
        char _b[1024];                                            \
        std::snprintf(_b, sizeof _b, __VA_ARGS__);                \
        if (p) (p)->err = _b; else g_prog_err = _b;               \
        return code;                                              \
    } while (0)

#define P_HIP(p, expr)                                                                   \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            P_FAIL(p, _e == hipErrorOutOfMemory ? SS_ERR_OOM : SS_ERR_HIP, "%s: %s", #expr, \
                   hipGetErrorString(_e));                                               \
    } while (0)

static void pfree(void *q) {
    if (q) (void)hipFree(q);
}

static int push_args(ss_prog *p) {
    p->host.n_agents = p->n;
    p->host.row_ptr = (const long long *)p->row_ptr;
    p->host.col_idx = p->col_idx;
    p->host.agent_ids = p->agent_ids;
    p->host.logical_ids = p->logical_ids;
    p->host.breeds = p->breeds;
    p->host.globals = p->globals;
    for (size_t i = 0; i < p->props.size(); ++i) {
        p->host.prop_r[i] = p->props[i].r;
        p->host.prop_w[i] = p->props[i].w;
    }
    p->host.seed = p->seed;
    p->host.logical_ids64 = (const long long *)p->logical_ids64;
    P_HIP(p, hipMemcpy(p->d_args, &p->host, sizeof p->host, hipMemcpyHostToDevice));
    return SS_OK;
}

extern "C" {

const char *ss_prog_last_error(const ss_prog *p) { return p ? p->err.c_str() : g_prog_err.c_str(); }

int ss_prog_create(ss_ctx *ctx, int64_t n_agents, ss_prog **out) {
    if (!ctx || !out || n_agents < 0 || n_agents > INT32_MAX) P_FAIL((ss_prog *)nullptr, SS_ERR_INVALID, "bad arguments");
    auto *p = new ss_prog();
    p->ctx = ctx;
    p->n = n_agents;
    const size_t nn = (size_t)std::max<int64_t>(n_agents, 1);
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipMalloc((void **)&p->d_args, sizeof(ProgArgsHost));
    if (e == hipSuccess) e = hipMalloc((void **)&p->agent_ids, nn * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&p->logical_ids, nn * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&p->breeds, nn * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&p->logical_ids64, nn * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc((void **)&p->row_ptr, (nn + 1) * sizeof(int64_t));
    if (e == hipSuccess) e = hipMemset(p->row_ptr, 0, (nn + 1) * sizeof(int64_t));
    if (e != hipSuccess) {
        g_prog_err = std::string("ss_prog_create: ") + hipGetErrorString(e);
        pfree(p->d_args); pfree(p->agent_ids); pfree(p->logical_ids); pfree(p->breeds); pfree(p->row_ptr);
        pfree(p->logical_ids64);
        delete p;
        return e == hipErrorOutOfMemory ? SS_ERR_OOM : SS_ERR_HIP;
    }
    *out = p;
    return SS_OK;
}

int ss_prog_destroy(ss_prog *p) {
    if (!p) return SS_OK;
    (void)hipStreamSynchronize(p->ctx->stream);
    if (p->module) (void)hipModuleUnload(p->module);
    pfree(p->d_args); pfree(p->agent_ids); pfree(p->logical_ids); pfree(p->breeds);
    pfree(p->row_ptr); pfree(p->col_idx); pfree(p->globals); pfree(p->logical_ids64);
    for (auto &q : p->props) {
        if (q.w != q.r) pfree(q.w);
        pfree(q.r);
    }
    delete p;
    return SS_OK;
}

int ss_prog_set_csr(ss_prog *p, const int64_t *row_ptr, const int32_t *col_idx, int64_t nnz) {
    if (!p || !row_ptr || (nnz && !col_idx)) P_FAIL(p, SS_ERR_INVALID, "NULL argument");
    if (row_ptr[0] != 0 || row_ptr[p->n] != nnz) P_FAIL(p, SS_ERR_INVALID, "row_ptr must start at 0 and end at nnz");
    for (int64_t i = 0; i < p->n; ++i)
        if (row_ptr[i + 1] < row_ptr[i]) P_FAIL(p, SS_ERR_INVALID, "row_ptr decreases at row %lld", (long long)i);
    for (int64_t e = 0; e < nnz; ++e)
        if (col_idx[e] < 0 || col_idx[e] >= p->n)
            P_FAIL(p, SS_ERR_RANGE, "col_idx[%lld] = %d outside [0, %lld)", (long long)e, col_idx[e], (long long)p->n);
    P_HIP(p, hipStreamSynchronize(p->ctx->stream));
    pfree(p->col_idx);
    p->col_idx = nullptr;
    P_HIP(p, hipMalloc((void **)&p->col_idx, sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)));
    if (nnz) P_HIP(p, hipMemcpy(p->col_idx, col_idx, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
    P_HIP(p, hipMemcpy(p->row_ptr, row_ptr, sizeof(int64_t) * (p->n + 1), hipMemcpyHostToDevice));
    p->nnz = nnz;
    return push_args(p);
}

int ss_prog_set_ids(ss_prog *p, const float *agent_ids, const float *logical_ids, const int64_t *logical_ids64,
                    const float *breeds) {
    if (!p || (p->n && (!agent_ids || !logical_ids || !logical_ids64 || !breeds)))
        P_FAIL(p, SS_ERR_INVALID, "NULL argument");
    if (p->n) {
        P_HIP(p, hipMemcpy(p->agent_ids, agent_ids, sizeof(float) * p->n, hipMemcpyHostToDevice));
        P_HIP(p, hipMemcpy(p->logical_ids, logical_ids, sizeof(float) * p->n, hipMemcpyHostToDevice));
        P_HIP(p, hipMemcpy(p->logical_ids64, logical_ids64, sizeof(int64_t) * p->n, hipMemcpyHostToDevice));
        P_HIP(p, hipMemcpy(p->breeds, breeds, sizeof(float) * p->n, hipMemcpyHostToDevice));
    }
    return push_args(p);
}

int ss_prog_add_prop(ss_prog *p, int32_t elem_bytes, int64_t width, int32_t double_buffered, const void *init,
                     int32_t *id) {
    if (!p || !id) P_FAIL(p, SS_ERR_INVALID, "NULL argument");
    if (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8)
        P_FAIL(p, SS_ERR_INVALID, "elem_bytes must be 1, 2, 4 or 8");
    if (width < 1) P_FAIL(p, SS_ERR_INVALID, "width must be >= 1");
    if (p->props.size() >= SS_MAXP_HOST) P_FAIL(p, SS_ERR_FULL, "at most %d properties", SS_MAXP_HOST);
    ProgProp q;
    q.elem_bytes = elem_bytes;
    q.width = width;
    q.double_buffered = double_buffered != 0;
    const size_t b = q.bytes() * (size_t)std::max<int64_t>(p->n, 1);
    P_HIP(p, hipMalloc(&q.r, b));
    if (q.double_buffered) {
        hipError_t e = hipMalloc(&q.w, b);
        if (e != hipSuccess) { pfree(q.r); P_HIP(p, e); }
    } else {
        q.w = q.r;
    }
    p->props.push_back(q);
    *id = (int32_t)p->props.size() - 1;
    if (init) {
        const size_t nb = q.bytes() * (size_t)p->n;
        if (nb) {
            P_HIP(p, hipMemcpy(q.r, init, nb, hipMemcpyHostToDevice));
            if (q.w != q.r) P_HIP(p, hipMemcpy(q.w, init, nb, hipMemcpyHostToDevice));
        }
    }
    return push_args(p);
}

int ss_prog_prop_upload(ss_prog *p, int32_t id, const void *data) {
    if (!p || id < 0 || id >= (int32_t)p->props.size() || !data) P_FAIL(p, SS_ERR_INVALID, "bad property or NULL");
    const ProgProp &q = p->props[id];
    const size_t nb = q.bytes() * (size_t)p->n;
    P_HIP(p, hipStreamSynchronize(p->ctx->stream));
    if (nb) {
        P_HIP(p, hipMemcpy(q.r, data, nb, hipMemcpyHostToDevice));
        if (q.w != q.r) P_HIP(p, hipMemcpy(q.w, data, nb, hipMemcpyHostToDevice));
    }
    return SS_OK;
}

int ss_prog_prop_download(ss_prog *p, int32_t id, void *out) {
    if (!p || id < 0 || id >= (int32_t)p->props.size() || !out) P_FAIL(p, SS_ERR_INVALID, "bad property or NULL");
    const ProgProp &q = p->props[id];
    P_HIP(p, hipStreamSynchronize(p->ctx->stream));
    const size_t nb = q.bytes() * (size_t)p->n;
    if (nb) P_HIP(p, hipMemcpy(out, q.r, nb, hipMemcpyDeviceToHost));
    return SS_OK;
}

int ss_prog_set_globals(ss_prog *p, const double *values, int64_t count) {
    if (!p || (count && !values) || count < 0) P_FAIL(p, SS_ERR_INVALID, "bad globals");
    P_HIP(p, hipStreamSynchronize(p->ctx->stream));
    if (count != p->n_globals) {
        pfree(p->globals);
        p->globals = nullptr;
        P_HIP(p, hipMalloc((void **)&p->globals, sizeof(double) * (size_t)std::max<int64_t>(count, 1)));
        p->n_globals = count;
    }
    if (count) P_HIP(p, hipMemcpy(p->globals, values, sizeof(double) * count, hipMemcpyHostToDevice));
    return push_args(p);
}

static int push_seed(ss_prog *p) {
    if (!p->module) return SS_OK;
    hipDeviceptr_t dp = nullptr;
    size_t sz = 0;
    P_HIP(p, hipModuleGetGlobal(&dp, &sz, p->module, "ss_rt_seed"));
    if (sz != sizeof(int32_t)) P_FAIL(p, SS_ERR_RTC, "ss_rt_seed has size %zu", sz);
    P_HIP(p, hipMemcpy((void *)dp, &p->seed, sizeof(int32_t), hipMemcpyHostToDevice));
    return SS_OK;
}

// The reference keeps the seed as int32 (model.py:1104) and reduces it mod
// 2^32 in the RNG (math_utils.py:108-111); the low 32 bits are what matter.
int ss_prog_set_seed(ss_prog *p, int64_t seed) {
    if (!p) P_FAIL(p, SS_ERR_INVALID, "NULL program");
    P_HIP(p, hipStreamSynchronize(p->ctx->stream));
    p->seed = (int32_t)(uint32_t)(uint64_t)seed;
    int rc = push_args(p);
    if (rc) return rc;
    return push_seed(p);
}

int ss_prog_compile(ss_prog *p, const char *source, const char *kernel_names, char *log, int64_t log_cap) {
    if (!p || !source || !kernel_names) P_FAIL(p, SS_ERR_INVALID, "NULL argument");
    if (log && log_cap > 0) log[0] = 0;
    hiprtcProgram prog;
    const char *hdrs[] = {kStepRtHeader};
    const char *names[] = {"sagesim_rt.h"};
    if (hiprtcCreateProgram(&prog, source, "step_functions.hip", 1, hdrs, names) != HIPRTC_SUCCESS)
        P_FAIL(p, SS_ERR_RTC, "hiprtcCreateProgram failed");
    // packed-FP32 instructions stay out of the generated code object
    const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++20", "-ffp-contract=off",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string lg(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &lg[0]);
    if (log && log_cap > 0) {
        std::strncpy(log, lg.c_str(), (size_t)log_cap - 1);
        log[log_cap - 1] = 0;
    }
    if (rc != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        P_FAIL(p, SS_ERR_RTC, "hipRTC compile failed: %s", lg.substr(0, 900).c_str());
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    std::vector<char> code(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    P_HIP(p, hipStreamSynchronize(p->ctx->stream));
    if (p->module) {
        (void)hipModuleUnload(p->module);
        p->module = nullptr;
    }
    p->kernels.clear();
    P_HIP(p, hipModuleLoadData(&p->module, code.data()));
    std::string names_s(kernel_names);
    size_t pos = 0;
    while (pos <= names_s.size()) {
        size_t nl = names_s.find('\n', pos);
        std::string nm = names_s.substr(pos, nl == std::string::npos ? std::string::npos : nl - pos);
        if (!nm.empty()) {
            hipFunction_t f = nullptr;
            hipError_t e = hipModuleGetFunction(&f, p->module, nm.c_str());
            if (e != hipSuccess) P_FAIL(p, SS_ERR_RTC, "kernel %s not in the compiled module", nm.c_str());
            p->kernels.push_back(f);
        }
        if (nl == std::string::npos) break;
        pos = nl + 1;
    }
    return push_seed(p);
}

int ss_prog_set_schedule(ss_prog *p, int32_t n, const int32_t *kernel, const int64_t *start, const int64_t *count) {
    if (!p || n < 0 || (n && (!kernel || !start || !count))) P_FAIL(p, SS_ERR_INVALID, "bad schedule");
    std::vector<Launch> s;
    for (int32_t i = 0; i < n; ++i) {
        if (kernel[i] < 0 || kernel[i] >= (int32_t)p->kernels.size())
            P_FAIL(p, SS_ERR_RANGE, "launch %d names kernel %d of %zu", i, kernel[i], p->kernels.size());
        if (start[i] < 0 || count[i] < 0 || start[i] + count[i] > p->n)
            P_FAIL(p, SS_ERR_RANGE, "launch %d covers agents outside [0, %lld)", i, (long long)p->n);
        s.push_back({kernel[i], start[i], count[i]});
    }
    p->schedule = s;
    return SS_OK;
}

// Every tick: the launches of the schedule in order (a kernel boundary
// between priority groups is the grid barrier of model.py:2356-2366), then the
// copy of every double-buffered property's write buffer into its read buffer
// (model.py:2378-2389).
int ss_prog_simulate(ss_prog *p, int64_t tick0, int64_t ticks) {
    if (!p || ticks < 0) P_FAIL(p, SS_ERR_INVALID, "bad arguments");
    if (!p->module) P_FAIL(p, SS_ERR_STATE, "no compiled step functions");
    hipStream_t st = p->ctx->stream;
    for (int64_t k = 0; k < ticks; ++k) {
        long long tick = (long long)(tick0 + k);
        for (const Launch &l : p->schedule) {
            if (l.count == 0) continue;
            const ProgArgsHost *a = p->d_args;
            long long start = l.start, count = l.count;
            void *args[] = {(void *)&a, (void *)&tick, (void *)&start, (void *)&count};
            const unsigned grid = (unsigned)((count + 255) / 256);
            P_HIP(p, hipModuleLaunchKernel(p->kernels[l.kernel], grid, 1, 1, 256, 1, 1, 0, st, args, nullptr));
        }
        for (const ProgProp &q : p->props)
            if (q.double_buffered && p->n)
                P_HIP(p, hipMemcpyAsync(q.r, q.w, q.bytes() * (size_t)p->n, hipMemcpyDeviceToDevice, st));
    }
    P_HIP(p, hipStreamSynchronize(st));
    return SS_OK;
}

}  // extern "C"
