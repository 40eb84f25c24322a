// Device runtime of user step functions compiled with hipRTC.
//
// A step function keeps the reference's argument order
// (tick, agent_index, globals..., agent_ids, breeds, locations, properties...)
// and is written as a HIP __device__ function; parameters may be declared
// `auto`.  The types below give the reference's semantics without rewriting
// the source:
//   * reading prop[i] reads the tick-start snapshot (read buffer);
//   * prop[i] = v writes the write buffer, prop[i] op= v writes
//     write[i] = read[i] op v (model.py:2159-2220 turns both into writes of the
//     write buffer); with no_double_buffer both buffers are the same array;
//   * locations[i] is agent i's CSR row (model.py:1643-1793), len(row) its size;
//   * rand_uniform_philox(tick, agent_index, salt) draws the reference's
//     Philox-2x32-10 uniform keyed on the agent's logical id and the model seed
//     (math_utils.py:87-189 with the injection of model.py:1903-1953).
#pragma once

#define SS_MAXP 48

struct SSProgArgs {
    long long n_agents;
    const long long *row_ptr;
    const int *col_idx;
    const float *agent_ids;
    const float *logical_ids;
    const float *breeds;
    const double *globals;
    void *prop_r[SS_MAXP];
    void *prop_w[SS_MAXP];
    int seed;
    const long long *logical_ids64;   // exact ids (the float32 ids round above 2^24)
};

// Seed for draws keyed on a plain id rather than the running agent.
__device__ int ss_rt_seed;

struct AgentIndex {
    long long i;
    float logical_id;     // the reference's float32 logical id (model.py:1090-1096)
    int seed;
    long long id64;       // the same id, exact
    __device__ operator long long() const { return i; }
};

// ---- reference RNG: Philox-2x32-10 round structure, word 0, low 23 bits ----
// The multiplier is the one math_utils.py:111 multiplies by, 3529668003
// (0xD26275A3; the comment at math_utils.py:81 names another hex value, the
// code's decimal is what the reference computes with).  Key bump 0x9E3779B9.
#define SS_RT_PHILOX2_M 3529668003u
__device__ inline unsigned int ss_rt_philox2x32(unsigned int key, unsigned int c0, unsigned int c1) {
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        unsigned long long p = (unsigned long long)c0 * SS_RT_PHILOX2_M;
        unsigned int n0 = (unsigned int)(p >> 32) ^ key ^ c1;
        c1 = (unsigned int)p;
        key += 0x9E3779B9u;
        c0 = n0;
    }
    unsigned long long p = (unsigned long long)c0 * SS_RT_PHILOX2_M;
    return (unsigned int)(p >> 32) ^ key ^ c1;
}

// counter ((tick * 1000003 + salt) mod 2^32, id mod 2^32), key seed mod 2^32
__device__ inline double ss_rt_uniform(int seed, long long tick, long long id, long long salt) {
    const unsigned int c0 = (unsigned int)(unsigned long long)(tick * 1000003LL + salt);
    const unsigned int c1 = (unsigned int)(unsigned long long)id;
    const unsigned int w = ss_rt_philox2x32((unsigned int)seed, c0, c1);
    return (double)(w & 0x7FFFFFu) / 8388608.0;
}

__device__ inline double rand_uniform_philox(long long tick, const AgentIndex &a, long long salt) {
    return ss_rt_uniform(a.seed, tick, (long long)a.logical_id, salt);
}

template <class Id>
__device__ inline double rand_uniform_philox(long long tick, Id id, long long salt) {
    return ss_rt_uniform(ss_rt_seed, tick, (long long)id, salt);
}

// ---- Philox-4x32-10 draw of the compartment workloads ----
// key = (seed, 0), counter = (tick lo, tick hi, id lo, id hi), word 0, 23 low
// bits: the draw the built-in fused kernel makes for the same agent and tick.
__device__ inline unsigned int ss_rt_philox4x32_w0(unsigned int c0, unsigned int c1, unsigned int c2,
                                                   unsigned int c3, unsigned int k0, unsigned int k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        unsigned long long p0 = (unsigned long long)0xD2511F53u * c0;
        unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c2;
        unsigned int n0 = (unsigned int)(p1 >> 32) ^ c1 ^ k0;
        unsigned int n1 = (unsigned int)p1;
        unsigned int n2 = (unsigned int)(p0 >> 32) ^ c3 ^ k1;
        unsigned int n3 = (unsigned int)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    return c0;
}

__device__ inline unsigned int rand_u23_philox4x32(long long tick, const AgentIndex &a) {
    const unsigned long long t = (unsigned long long)tick, id = (unsigned long long)a.id64;
    return ss_rt_philox4x32_w0((unsigned int)t, (unsigned int)(t >> 32), (unsigned int)id, (unsigned int)(id >> 32),
                               (unsigned int)a.seed, 0u) & 0x7FFFFFu;
}

__device__ inline double rand_uniform_philox4x32(long long tick, const AgentIndex &a) {
    return (double)rand_u23_philox4x32(tick, a) / 8388608.0;
}

// ---- properties ----
template <class T>
struct PropRef {
    const T *r;
    T *w;
    __device__ operator T() const { return *r; }
    __device__ PropRef &operator=(const PropRef &o) { *w = (T)(*o.r); return *this; }
    template <class V> __device__ PropRef &operator=(V v) { *w = (T)v; return *this; }
    template <class V> __device__ PropRef &operator+=(V v) { *w = (T)(*r + v); return *this; }
    template <class V> __device__ PropRef &operator-=(V v) { *w = (T)(*r - v); return *this; }
    template <class V> __device__ PropRef &operator*=(V v) { *w = (T)(*r * v); return *this; }
    template <class V> __device__ PropRef &operator/=(V v) { *w = (T)(*r / v); return *this; }
};

template <class T>
struct Prop {
    const T *r;
    T *w;
    __device__ PropRef<T> operator[](long long i) const { return PropRef<T>{r + i, w + i}; }
};

template <class T>
struct PropRowRef {
    const T *r;
    T *w;
    int width;
    __device__ PropRef<T> operator[](long long j) const { return PropRef<T>{r + j, w + j}; }
};

template <class T>
__device__ inline int len(const PropRowRef<T> &row) { return row.width; }

template <class T>
struct PropRows {
    const T *r;
    T *w;
    int width;
    __device__ PropRowRef<T> operator[](long long i) const {
        return PropRowRef<T>{r + i * (long long)width, w + i * (long long)width, width};
    }
};

// ---- network (CSR) ----
struct NeighborRow {
    const int *p;
    int n;
    __device__ int operator[](long long j) const { return p[j]; }
};
__device__ inline int len(const NeighborRow &row) { return row.n; }

struct Locations {
    const long long *row_ptr;
    const int *col_idx;
    __device__ NeighborRow operator[](long long i) const {
        const long long b = row_ptr[i];
        return NeighborRow{col_idx + b, (int)(row_ptr[i + 1] - b)};
    }
};

struct FloatArray {
    const float *p;
    __device__ float operator[](long long i) const { return p[i]; }
};

struct GlobalArray {
    const double *p;
    long long n;
    __device__ double operator[](long long i) const { return p[i]; }
};

// ---- helpers of utils.py:37-81 ----
__device__ inline int get_num_neighbors(long long i, const Locations &loc) { return loc[i].n; }
__device__ inline int get_neighbor(long long i, long long j, const Locations &loc) { return loc[i][j]; }
template <class P>
__device__ inline auto get_this_agent_data_from_tensor(long long i, const P &t) { return t[i]; }
template <class P, class V>
__device__ inline void set_this_agent_data_from_tensor(long long i, const P &t, V v) { t[i] = v; }
