// General step functions: user HIP device code compiled with hipRTC.
//
// Replaces generate_gpu_func + the CuPy JIT of the generated module
// (model.py:2010-2457, :800-865) for models whose step functions are not the
// built-in compartment step: the host side writes one wrapper kernel per
// priority group around the user's __device__ functions, this file compiles
// it for gfx950 (code objects cached by a hash of the source, as the
// reference caches its generated module, model.py:843-865), loads it, and
// runs the per-tick launch schedule followed by the write-back of
// double-buffered arrays (model.py:2368-2408).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

#include <hip/hiprtc.h>

#include "../../include/sagestep.h"
#include "ss_runtime.h"

namespace {

// the device runtime header, embedded at build time from step_rt.h
const char *kStepRtHeader =
#include "step_rt_embed.inc"
    ;

#define SS_MAXP_HOST 64
#define SS_MAXX_HOST 16

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
    void *extra[SS_MAXX_HOST];
};

// a device array of the program: rows x width elements, double-buffered
// (read and write buffer) or one buffer for both
struct ProgProp {
    int32_t elem_bytes = 0;
    int64_t rows = 0;
    int64_t width = 1;
    bool double_buffered = true;
    void *r = nullptr, *w = nullptr;
    size_t bytes() const { return (size_t)elem_bytes * (size_t)width * (size_t)rows; }
};

struct Launch {
    int32_t kernel;
    int64_t start, count;
    int64_t order_off;          // first entry of this launch in the order array, -1: agents in index order
};

struct Extra {
    void *d = nullptr;
    int64_t bytes = 0;
};

uint64_t fnv1a(uint64_t h, const void *p, size_t n) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

// Loops of unknown trip count (a step function's walk over its neighbour row)
// are unrolled at run time, so one thread has several neighbour loads in
// flight instead of one round trip per neighbour; SAGESIM_AMD_RTC_UNROLL
// sets the count (1 disables it) and SAGESIM_AMD_RTC_UNROLL_THRESHOLD the
// size the unrolled body may reach (LLVM lowers the count of a loop whose
// body would grow past it: a heavy loop body unrolls less).
std::vector<std::string> rtc_options() {
    std::vector<std::string> o = {"--offload-arch=gfx950", "-O3", "-std=c++20", "-ffp-contract=off",
                                  "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    int unroll = 48;
    std::string threshold = "1200";
    if (const char *u = std::getenv("SAGESIM_AMD_RTC_UNROLL")) unroll = std::atoi(u);
    if (const char *t = std::getenv("SAGESIM_AMD_RTC_UNROLL_THRESHOLD")) threshold = t;
    if (unroll > 1) {
        o.insert(o.end(), {"-mllvm", "-unroll-runtime", "-mllvm", "-unroll-count=" + std::to_string(unroll)});
        if (!threshold.empty()) o.insert(o.end(), {"-mllvm", "-unroll-threshold=" + threshold});
    }
    return o;
}

// code objects by hash of (options, runtime header, source), for the process
std::mutex g_cache_mu;
std::map<uint64_t, std::vector<char>> g_code_cache;

std::string disk_cache_dir() {
    if (const char *d = std::getenv("SAGESIM_AMD_RTC_CACHE")) return d;   // "" disables the disk cache
    const char *home = std::getenv("HOME");
    return home ? std::string(home) + "/.cache/sagesim_amd/rtc" : std::string();
}

bool read_file(const std::string &path, std::vector<char> &out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return !out.empty();
}

void write_file_atomic(const std::string &dir, const std::string &name, const std::vector<char> &data) {
    std::string acc;
    for (size_t i = 0; i <= dir.size(); ++i) {          // mkdir -p
        if (i == dir.size() || dir[i] == '/') {
            if (!acc.empty()) ::mkdir(acc.c_str(), 0755);
        }
        if (i < dir.size()) acc += dir[i];
    }
    const std::string tmp = dir + "/" + name + ".tmp" + std::to_string((long long)::getpid());
    {
        std::ofstream f(tmp, std::ios::binary);
        if (!f) return;
        f.write(data.data(), (std::streamsize)data.size());
        if (!f) return;
    }
    std::rename(tmp.c_str(), (dir + "/" + name).c_str());
}

__global__ void writeback_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n16,
                                 const unsigned char *__restrict__ src_b, unsigned char *__restrict__ dst_b,
                                 int64_t tail_from, int64_t n_bytes) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t i = g; i < n16; i += stride) dst[i] = src[i];
    for (int64_t i = tail_from + g; i < n_bytes; i += stride) dst_b[i] = src_b[i];
}

}  // namespace

struct ss_prog {
    ss_ctx *ctx = nullptr;
    int64_t n = 0;
    int64_t *row_ptr = nullptr;
    int32_t *col_idx = nullptr;
    int64_t nnz = 0;
    uint64_t csr_hash = 0;
    float *agent_ids = nullptr, *logical_ids = nullptr, *breeds = nullptr;
    int64_t *logical_ids64 = nullptr;
    double *globals = nullptr;
    int64_t n_globals = 0;
    std::vector<double> h_globals;
    std::vector<ProgProp> props;
    Extra extras[SS_MAXX_HOST];
    ProgArgsHost host{};
    ProgArgsHost *d_args = nullptr;
    hipModule_t module = nullptr;
    std::vector<hipFunction_t> kernels;
    std::vector<Launch> schedule;
    int32_t *order = nullptr;                   // agent order of every launch (concatenated)
    int64_t order_n = 0;
    int32_t seed = 0;
    int32_t cache_hit = 0;                      // the last compile came from the code-object cache
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
    for (int i = 0; i < SS_MAXX_HOST; ++i) p->host.extra[i] = p->extras[i].d;
    P_HIP(p, hipMemcpy(p->d_args, &p->host, sizeof p->host, hipMemcpyHostToDevice));
    return SS_OK;
}

// the module globals neighbor_offsets / neighbor_values (step_rt.h) -> the CSR
static int push_csr_globals(ss_prog *p) {
    if (!p->module) return SS_OK;
    const void *vals[2] = {p->row_ptr, p->col_idx};
    const char *names[2] = {"neighbor_offsets", "neighbor_values"};
    for (int k = 0; k < 2; ++k) {
        hipDeviceptr_t dp = nullptr;
        size_t sz = 0;
        P_HIP(p, hipModuleGetGlobal(&dp, &sz, p->module, names[k]));
        if (sz != sizeof(void *)) P_FAIL(p, SS_ERR_RTC, "%s has size %zu", names[k], sz);
        P_HIP(p, hipMemcpy((void *)dp, &vals[k], sizeof(void *), hipMemcpyHostToDevice));
    }
    return SS_OK;
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

// copy every double-buffered array's write buffer into its read buffer
static int write_back(ss_prog *p, hipStream_t st) {
    for (const ProgProp &q : p->props) {
        if (!q.double_buffered || q.rows == 0) continue;
        const int64_t nb = (int64_t)q.bytes();
        const int64_t n16 = nb / 16;
        int grid = (int)std::min<int64_t>(std::max<int64_t>((n16 + 255) / 256, 1), (int64_t)p->ctx->num_cus * 4);
        hipLaunchKernelGGL(writeback_kernel, dim3(grid), dim3(256), 0, st, (const uint4 *)q.w, (uint4 *)q.r, n16,
                           (const unsigned char *)q.w, (unsigned char *)q.r, n16 * 16, nb);
        P_HIP(p, hipGetLastError());
    }
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
    pfree(p->row_ptr); pfree(p->col_idx); pfree(p->globals); pfree(p->logical_ids64); pfree(p->order);
    for (auto &q : p->props) {
        if (q.w != q.r) pfree(q.w);
        pfree(q.r);
    }
    for (auto &x : p->extras) pfree(x.d);
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
    p->csr_hash = fnv1a(fnv1a(1469598103934665603ull, row_ptr, sizeof(int64_t) * (size_t)(p->n + 1)), col_idx,
                        sizeof(int32_t) * (size_t)nnz);
    int rc = push_csr_globals(p);
    if (rc) return rc;
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

int ss_prog_add_array(ss_prog *p, int32_t elem_bytes, int64_t rows, int64_t width, int32_t double_buffered,
                      const void *init, int32_t *id) {
    if (!p || !id) P_FAIL(p, SS_ERR_INVALID, "NULL argument");
    if (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8)
        P_FAIL(p, SS_ERR_INVALID, "elem_bytes must be 1, 2, 4 or 8");
    if (width < 1 || rows < 0) P_FAIL(p, SS_ERR_INVALID, "width must be >= 1 and rows >= 0");
    if (p->props.size() >= SS_MAXP_HOST) P_FAIL(p, SS_ERR_FULL, "at most %d device arrays", SS_MAXP_HOST);
    ProgProp q;
    q.elem_bytes = elem_bytes;
    q.rows = rows;
    q.width = width;
    q.double_buffered = double_buffered != 0;
    const size_t b = std::max<size_t>(q.bytes(), 16);
    P_HIP(p, hipMalloc(&q.r, b));
    if (q.double_buffered) {
        hipError_t e = hipMalloc(&q.w, b);
        if (e != hipSuccess) { pfree(q.r); P_HIP(p, e); }
    } else {
        q.w = q.r;
    }
    const size_t nb = q.bytes();
    hipError_t e = hipSuccess;
    if (nb) {
        if (init) e = hipMemcpy(q.r, init, nb, hipMemcpyHostToDevice);
        else e = hipMemset(q.r, 0, nb);
        if (e == hipSuccess && q.w != q.r) e = hipMemcpy(q.w, q.r, nb, hipMemcpyDeviceToDevice);
    }
    if (e != hipSuccess) {
        if (q.w != q.r) pfree(q.w);
        pfree(q.r);
        P_HIP(p, e);
    }
    p->props.push_back(q);
    *id = (int32_t)p->props.size() - 1;
    return push_args(p);
}

int ss_prog_add_prop(ss_prog *p, int32_t elem_bytes, int64_t width, int32_t double_buffered, const void *init,
                     int32_t *id) {
    if (!p) P_FAIL(p, SS_ERR_INVALID, "NULL program");
    return ss_prog_add_array(p, elem_bytes, p->n, width, double_buffered, init, id);
}

int ss_prog_prop_upload(ss_prog *p, int32_t id, const void *data) {
    if (!p || id < 0 || id >= (int32_t)p->props.size() || !data) P_FAIL(p, SS_ERR_INVALID, "bad array or NULL");
    const ProgProp &q = p->props[id];
    const size_t nb = q.bytes();
    P_HIP(p, hipStreamSynchronize(p->ctx->stream));
    if (nb) {
        P_HIP(p, hipMemcpy(q.r, data, nb, hipMemcpyHostToDevice));
        if (q.w != q.r) P_HIP(p, hipMemcpy(q.w, data, nb, hipMemcpyHostToDevice));
    }
    return SS_OK;
}

int ss_prog_prop_download(ss_prog *p, int32_t id, void *out) {
    if (!p || id < 0 || id >= (int32_t)p->props.size() || !out) P_FAIL(p, SS_ERR_INVALID, "bad array or NULL");
    const ProgProp &q = p->props[id];
    P_HIP(p, hipStreamSynchronize(p->ctx->stream));
    const size_t nb = q.bytes();
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
    p->h_globals.assign(values, values + count);
    return push_args(p);
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

int ss_prog_set_extra(ss_prog *p, int32_t slot, int64_t bytes, const void *data) {
    if (!p || slot < 0 || slot >= SS_MAXX_HOST || bytes < 0 || (bytes && !data))
        P_FAIL(p, SS_ERR_INVALID, "bad extra slot, size or NULL data");
    P_HIP(p, hipStreamSynchronize(p->ctx->stream));
    Extra &x = p->extras[slot];
    if (x.bytes != bytes || !x.d) {
        pfree(x.d);
        x.d = nullptr;
        x.bytes = 0;
        P_HIP(p, hipMalloc(&x.d, (size_t)std::max<int64_t>(bytes, 16)));
        x.bytes = bytes;
    }
    if (bytes) P_HIP(p, hipMemcpy(x.d, data, (size_t)bytes, hipMemcpyHostToDevice));
    return push_args(p);
}

int ss_prog_get_extra(ss_prog *p, int32_t slot, int64_t bytes, void *out) {
    if (!p || slot < 0 || slot >= SS_MAXX_HOST || (bytes && !out)) P_FAIL(p, SS_ERR_INVALID, "bad extra slot or NULL");
    const Extra &x = p->extras[slot];
    if (!x.d || bytes != x.bytes) P_FAIL(p, SS_ERR_INVALID, "extra slot %d holds %lld bytes", slot, (long long)x.bytes);
    P_HIP(p, hipStreamSynchronize(p->ctx->stream));
    if (bytes) P_HIP(p, hipMemcpy(out, x.d, (size_t)bytes, hipMemcpyDeviceToHost));
    return SS_OK;
}

int ss_prog_compile(ss_prog *p, const char *source, const char *kernel_names, char *log, int64_t log_cap) {
    if (!p || !source || !kernel_names) P_FAIL(p, SS_ERR_INVALID, "NULL argument");
    if (log && log_cap > 0) log[0] = 0;
    const std::vector<std::string> opts_s = rtc_options();
    std::vector<const char *> opts;
    for (const auto &o : opts_s) opts.push_back(o.c_str());
    uint64_t key = 1469598103934665603ull;
    for (const auto &o : opts_s) key = fnv1a(key, o.c_str(), o.size() + 1);
    key = fnv1a(key, kStepRtHeader, std::strlen(kStepRtHeader));
    key = fnv1a(key, source, std::strlen(source));
    char key_hex[32];
    std::snprintf(key_hex, sizeof key_hex, "%016llx.co", (unsigned long long)key);
    std::vector<char> code;
    p->cache_hit = 0;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_code_cache.find(key);
        if (it != g_code_cache.end()) {
            code = it->second;
            p->cache_hit = 1;
        }
    }
    const std::string dir = disk_cache_dir();
    if (code.empty() && !dir.empty() && read_file(dir + "/" + key_hex, code)) p->cache_hit = 2;
    if (code.empty()) {
        hiprtcProgram prog;
        const char *hdrs[] = {kStepRtHeader};
        const char *names[] = {"sagesim_rt.h"};
        if (hiprtcCreateProgram(&prog, source, "step_functions.hip", 1, hdrs, names) != HIPRTC_SUCCESS)
            P_FAIL(p, SS_ERR_RTC, "hiprtcCreateProgram failed");
        // packed-FP32 instructions stay out of the generated code object
        hiprtcResult rc = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
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
        code.resize(cs);
        hiprtcGetCode(prog, code.data());
        hiprtcDestroyProgram(&prog);
        if (!dir.empty()) write_file_atomic(dir, key_hex, code);
    }
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        g_code_cache.emplace(key, code);
    }
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
    int rc = push_csr_globals(p);
    if (rc) return rc;
    return push_seed(p);
}

int ss_prog_cache_hit(ss_prog *p, int32_t *hit) {
    if (!p || !hit) P_FAIL(p, SS_ERR_INVALID, "NULL argument");
    *hit = p->cache_hit;
    return SS_OK;
}

int ss_prog_set_schedule(ss_prog *p, int32_t n, const int32_t *kernel, const int64_t *start, const int64_t *count) {
    if (!p || n < 0 || (n && (!kernel || !start || !count))) P_FAIL(p, SS_ERR_INVALID, "bad schedule");
    std::vector<Launch> s;
    for (int32_t i = 0; i < n; ++i) {
        if (kernel[i] < 0 || kernel[i] >= (int32_t)p->kernels.size())
            P_FAIL(p, SS_ERR_RANGE, "launch %d names kernel %d of %zu", i, kernel[i], p->kernels.size());
        if (start[i] < 0 || count[i] < 0 || start[i] + count[i] > p->n)
            P_FAIL(p, SS_ERR_RANGE, "launch %d covers agents outside [0, %lld)", i, (long long)p->n);
        s.push_back({kernel[i], start[i], count[i], -1});
    }
    p->schedule = s;
    pfree(p->order);
    p->order = nullptr;
    p->order_n = 0;
    return SS_OK;
}

int ss_prog_set_order(ss_prog *p, const int32_t *order, int64_t n) {
    if (!p || n < 0 || (n && !order)) P_FAIL(p, SS_ERR_INVALID, "bad order");
    int64_t need = 0;
    for (const Launch &l : p->schedule) need += l.count;
    if (n != need) P_FAIL(p, SS_ERR_INVALID, "order has %lld entries, the schedule covers %lld", (long long)n,
                          (long long)need);
    // each launch's segment is a permutation of its range
    int64_t off = 0;
    std::vector<char> seen;
    for (Launch &l : p->schedule) {
        seen.assign((size_t)l.count, 0);
        for (int64_t k = 0; k < l.count; ++k) {
            const int64_t a = (int64_t)order[off + k] - l.start;
            if (a < 0 || a >= l.count || seen[(size_t)a])
                P_FAIL(p, SS_ERR_RANGE, "order of a launch is not a permutation of its agents");
            seen[(size_t)a] = 1;
        }
        l.order_off = off;
        off += l.count;
    }
    P_HIP(p, hipStreamSynchronize(p->ctx->stream));
    pfree(p->order);
    p->order = nullptr;
    P_HIP(p, hipMalloc((void **)&p->order, sizeof(int32_t) * (size_t)std::max<int64_t>(n, 1)));
    if (n) P_HIP(p, hipMemcpy(p->order, order, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));
    p->order_n = n;
    return SS_OK;
}

// Every tick: the launches of the schedule in order (a kernel boundary
// between priority groups is the grid barrier of model.py:2356-2366), then the
// copy of every double-buffered array's write buffer into its read buffer
// (model.py:2378-2406).
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
            const int32_t *ord = l.order_off >= 0 && p->order ? p->order + l.order_off : nullptr;
            void *args[] = {(void *)&a, (void *)&tick, (void *)&start, (void *)&count, (void *)&ord};
            const unsigned grid = (unsigned)((count + 255) / 256);
            P_HIP(p, hipModuleLaunchKernel(p->kernels[l.kernel], grid, 1, 1, 256, 1, 1, 0, st, args, nullptr));
        }
        int rc = write_back(p, st);
        if (rc) return rc;
    }
    P_HIP(p, hipStreamSynchronize(st));
    return SS_OK;
}

// ---- checkpoint: replaces the pickle of Model.save / Model.load
// (model.py:1289-1310) with the program's device state, bit for bit ----
// layout: magic, version, n, nnz, csr_hash, tick, seed, n_globals, n_arrays,
// globals, then per array (elem_bytes, rows, width, double_buffered) and its
// read buffer (the write buffer equals it between ticks)
static const uint64_t kProgMagic = 0x53534147455052ull;   // "SSAGEPR"

int ss_prog_save(ss_prog *p, int64_t tick, void *buf, int64_t cap, int64_t *size) {
    if (!p || !size || cap < 0 || (cap && !buf)) P_FAIL(p, SS_ERR_INVALID, "bad arguments");
    int64_t need = 9 * 8 + 8 * p->n_globals;
    for (const ProgProp &q : p->props) need += 4 * 8 + (int64_t)((q.bytes() + 7) / 8 * 8);
    *size = need;
    if (!buf) return SS_OK;
    if (cap < need) P_FAIL(p, SS_ERR_INVALID, "buffer of %lld bytes, checkpoint needs %lld", (long long)cap,
                           (long long)need);
    P_HIP(p, hipStreamSynchronize(p->ctx->stream));
    std::memset(buf, 0, (size_t)need);
    int64_t *h = static_cast<int64_t *>(buf);
    h[0] = (int64_t)kProgMagic; h[1] = 1; h[2] = p->n; h[3] = p->nnz; h[4] = (int64_t)p->csr_hash;
    h[5] = tick; h[6] = p->seed; h[7] = p->n_globals; h[8] = (int64_t)p->props.size();
    char *c = static_cast<char *>(buf) + 9 * 8;
    if (p->n_globals) std::memcpy(c, p->h_globals.data(), 8 * (size_t)p->n_globals);
    c += 8 * p->n_globals;
    for (const ProgProp &q : p->props) {
        int64_t d[4] = {q.elem_bytes, q.rows, q.width, q.double_buffered ? 1 : 0};
        std::memcpy(c, d, sizeof d);
        c += sizeof d;
        if (q.bytes()) P_HIP(p, hipMemcpy(c, q.r, q.bytes(), hipMemcpyDeviceToHost));
        c += (q.bytes() + 7) / 8 * 8;
    }
    return SS_OK;
}

int ss_prog_restore(ss_prog *p, const void *buf, int64_t size, int64_t *tick) {
    if (!p || !buf || size < 9 * 8) P_FAIL(p, SS_ERR_INVALID, "bad arguments");
    const int64_t *h = static_cast<const int64_t *>(buf);
    if (h[0] != (int64_t)kProgMagic || h[1] != 1) P_FAIL(p, SS_ERR_INVALID, "not a program checkpoint");
    if (h[2] != p->n || h[3] != p->nnz || h[4] != (int64_t)p->csr_hash)
        P_FAIL(p, SS_ERR_INVALID, "checkpoint of another network (agents, entries or CSR differ)");
    if (h[7] != p->n_globals || h[8] != (int64_t)p->props.size())
        P_FAIL(p, SS_ERR_INVALID, "checkpoint has %lld globals and %lld arrays, the program %lld and %zu",
               (long long)h[7], (long long)h[8], (long long)p->n_globals, p->props.size());
    int64_t need = 9 * 8 + 8 * p->n_globals;
    for (const ProgProp &q : p->props) need += 4 * 8 + (int64_t)((q.bytes() + 7) / 8 * 8);
    if (size != need) P_FAIL(p, SS_ERR_INVALID, "checkpoint has %lld bytes, expected %lld", (long long)size,
                             (long long)need);
    const char *c = static_cast<const char *>(buf) + 9 * 8 + 8 * p->n_globals;
    for (const ProgProp &q : p->props) {
        int64_t d[4];
        std::memcpy(d, c, sizeof d);
        if (d[0] != q.elem_bytes || d[1] != q.rows || d[2] != q.width || d[3] != (q.double_buffered ? 1 : 0))
            P_FAIL(p, SS_ERR_INVALID, "checkpoint arrays differ in shape or buffering");
        c += sizeof d + (q.bytes() + 7) / 8 * 8;
    }
    int rc = ss_prog_set_globals(p, reinterpret_cast<const double *>(static_cast<const char *>(buf) + 9 * 8), h[7]);
    if (rc) return rc;
    if ((rc = ss_prog_set_seed(p, h[6]))) return rc;
    c = static_cast<const char *>(buf) + 9 * 8 + 8 * p->n_globals;
    for (size_t i = 0; i < p->props.size(); ++i) {
        c += 4 * 8;
        if (p->props[i].bytes() && (rc = ss_prog_prop_upload(p, (int32_t)i, c))) return rc;
        c += (p->props[i].bytes() + 7) / 8 * 8;
    }
    if (tick) *tick = h[5];
    return SS_OK;
}

}  // extern "C"
