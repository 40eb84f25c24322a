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

#define SS_MAXP 64        // device arrays of a program: properties, breed-local arrays and their index maps
#define SS_MAXX 16        // kernel-extra buffers

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
    void *extra[SS_MAXX];             // kernel-extra buffers (model.py:892-921)
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
// A write to the running agent's own slot of a double-buffered scalar
// property is held in an OwnSlot in registers and stored once after the step
// function returns (the wrapper kernel does that), so a step function that
// writes its own slot inside its neighbour loop keeps no store in that loop
// and the loop's loads can overlap.  Writes to any other slot, and every
// write to a no_double_buffer property or a row, go to memory at once.
// Reads always come from the read buffer, as in the reference.
// Device memory is read and written through global-address-space pointers,
// so loads and stores are global_* (not flat_*, which also count on the LDS
// counter and take the flat aperture check); the proxies keep plain pointers.
#define SS_RT_G __attribute__((address_space(1)))
template <class T>
__device__ inline T ss_ld(const T *p) { return *(const SS_RT_G T *)p; }
template <class T>
__device__ inline void ss_st(T *p, T v) { *(SS_RT_G T *)p = v; }

template <class T>
struct OwnSlot {
    T v;
    bool dirty;
};

template <class T>
struct PropRef {
    const T *r;
    T *w;
    OwnSlot<T> *own;      // non-null: this is the running agent's own slot
    __device__ operator T() const { return ss_ld(r); }
    __device__ void store(T v) const {
        if (own) {
            own->v = v;
            own->dirty = true;
        } else {
            ss_st(w, v);
        }
    }
    __device__ const PropRef &operator=(const PropRef &o) const { store((T)ss_ld(o.r)); return *this; }
    template <class V> __device__ const PropRef &operator=(V v) const { store((T)v); return *this; }
    template <class V> __device__ const PropRef &operator+=(V v) const { store((T)(ss_ld(r) + v)); return *this; }
    template <class V> __device__ const PropRef &operator-=(V v) const { store((T)(ss_ld(r) - v)); return *this; }
    template <class V> __device__ const PropRef &operator*=(V v) const { store((T)(ss_ld(r) * v)); return *this; }
    template <class V> __device__ const PropRef &operator/=(V v) const { store((T)(ss_ld(r) / v)); return *this; }
};

template <class T>
struct Prop {
    const T *r;
    T *w;
    long long self;       // the running agent's slot, -1 when writes are not held
    OwnSlot<T> *own;
    __device__ PropRef<T> operator[](long long i) const {
        const bool mine = i == self;
#ifdef SS_RT_OFF32
        // every property array is below 2 GiB (the wrapper's source says so):
        // a 32-bit byte offset from the array's base, which the loads take as
        // one register beside the base in scalar registers (an index outside
        // the array is outside it either way)
        const unsigned int o = (unsigned int)i * (unsigned int)sizeof(T);
        return PropRef<T>{(const T *)((const char *)r + o), (T *)((char *)w + o), mine ? own : (OwnSlot<T> *)nullptr};
#else
        return PropRef<T>{r + i, w + i, mine ? own : (OwnSlot<T> *)nullptr};
#endif
    }
};

// store a held own-slot write (called by the wrapper after the step returns)
template <class T>
__device__ inline void ss_flush(const Prop<T> &p, const OwnSlot<T> &o) {
    if (o.dirty) ss_st(p.w + p.self, o.v);
}

template <class T>
struct PropRowRef {
    const T *r;
    T *w;
    int width;
    __device__ PropRef<T> operator[](long long j) const { return PropRef<T>{r + j, w + j, nullptr}; }
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

// ---- breed-local arrays (model.py:620-651): one row per agent of a breed,
// `array[idx_map[agent_index]][j]` (shape (a,) per agent) or
// `array[idx_map[agent_index]][j][k]` (shape (a, b)); read from the read
// buffer, written to the write buffer (the same array with no_double_buffer)
template <class T>
struct BlaMatRef {
    const T *r;
    T *w;
    int a, b;
    __device__ PropRowRef<T> operator[](long long j) const {
        return PropRowRef<T>{r + j * (long long)b, w + j * (long long)b, b};
    }
};

template <class T>
__device__ inline int len(const BlaMatRef<T> &m) { return m.a; }

template <class T>
struct BlaRows1 {
    const T *r;
    T *w;
    long long rows;
    int a;
    __device__ PropRowRef<T> operator[](long long row) const {
        return PropRowRef<T>{r + row * (long long)a, w + row * (long long)a, a};
    }
};

template <class T>
struct BlaRows2 {
    const T *r;
    T *w;
    long long rows;
    int a, b;
    __device__ BlaMatRef<T> operator[](long long row) const {
        const long long o = row * (long long)a * (long long)b;
        return BlaMatRef<T>{r + o, w + o, a, b};
    }
};

template <class T>
__device__ inline long long len(const BlaRows1<T> &x) { return x.rows; }
template <class T>
__device__ inline long long len(const BlaRows2<T> &x) { return x.rows; }

// buffer index -> row of a breed-local array (-1 for agents of other breeds)
struct IdxMap {
    const int *p;
    __device__ int operator[](long long i) const { return ss_ld(p + i); }
};

// ---- network (CSR) ----
struct NeighborRow {
    const int *p;
    int n;
    __device__ int operator[](long long j) const { return ss_ld(p + j); }
};
__device__ inline int len(const NeighborRow &row) { return row.n; }

struct Locations {
    const long long *row_ptr;
    const int *col_idx;
    __device__ NeighborRow operator[](long long i) const {
        const long long b = ss_ld(row_ptr + i);
        return NeighborRow{col_idx + b, (int)(ss_ld(row_ptr + i + 1) - b)};
    }
};

struct FloatArray {
    const float *p;
    __device__ float operator[](long long i) const { return ss_ld(p + i); }
};

struct GlobalArray {
    const double *p;
    long long n;
    __device__ double operator[](long long i) const { return p[i]; }
};

// ---- helpers of utils.py:34-57 ----
// The reference's step functions see the CSR as `neighbor_offsets` and
// `neighbor_values` (model.py:1845-1900 renames `locations`), and its helpers
// take those: get_num_neighbors(agent_index, neighbor_offsets),
// get_neighbor(agent_index, i, neighbor_offsets, neighbor_values).  Here the
// two names are module globals the library points at the program's CSR, so
// those calls compile unchanged; the forms taking `locations` work too.
struct NeighborOffsets {
    const long long *p;
    __device__ long long operator[](long long i) const { return p[i]; }
};
struct NeighborValues {
    const int *p;
    __device__ int operator[](long long i) const { return p[i]; }
};
__device__ NeighborOffsets neighbor_offsets;
__device__ NeighborValues neighbor_values;

__device__ inline long long get_num_neighbors(long long i, const NeighborOffsets &o) { return o[i + 1] - o[i]; }
__device__ inline int get_neighbor(long long i, long long j, const NeighborOffsets &o, const NeighborValues &v) {
    return v[o[i] + j];
}
__device__ inline int get_num_neighbors(long long i, const Locations &loc) { return loc[i].n; }
__device__ inline int get_neighbor(long long i, long long j, const Locations &loc) { return loc[i][j]; }
template <class P>
__device__ inline auto get_this_agent_data_from_tensor(long long i, const P &t) { return t[i]; }
template <class P, class V>
__device__ inline void set_this_agent_data_from_tensor(long long i, const P &t, V v) { t[i] = v; }
