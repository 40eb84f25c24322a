/*
 * CPU ORACLE for the per-tick agent step path.  TEST INFRASTRUCTURE ONLY:
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this file's library; the product path (sagesim_amd) never does.
 *
 * It is a plain restatement of the reference semantics, written for being
 * obviously right rather than fast (one pass over every agent, every
 * neighbour, every tick; OpenMP over agents only).  or_sir_run_fast is the
 * exception: the same results with the work a tuned CPU code would skip left
 * out, timed as the CPU baseline.  Citations are relative to the reference
 * repository root.
 *
 * Pinning (what checks each function against something other than itself):
 *   or_philox2x32_ref   - tests/test_oracle_rng.py: pure-Python big-int
 *                          transcription of math_utils.py:106-189 using the
 *                          reference's own // and % arithmetic, special
 *                          counters (0, all-ones, 2^32 wrap of tick*1000003).
 *   or_philox4x32_10    - tests/test_oracle_rng.py: torch's
 *                          ATen/core/PhiloxRNGEngine.h and rocRAND's
 *                          rocrand_philox4x32_10.h compiled on the host, plus
 *                          the pure-Python transcription.
 *   or_sir_run          - known-answer tests (chain wave, beta=0, gamma=1,
 *                          p=1 BFS front), invariants (monotone S->I->R,
 *                          counts conserved) and brute force on tiny graphs
 *                          written independently in tests/test_oracle_sir.py.
 *   or_sir_run_fast     - or_sir_run on the same inputs, bit for bit
 *                          (tests/test_oracle_fast.py), with thresholds at
 *                          the exact u == p boundary.
 *   or_ref_sir_run      - the reference's own known answers:
 *                          tests/test_simulation.py:95-115 (chain) and
 *                          tests/test_double_buffer.py:247-271 (hierarchy).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Reference RNG: Philox-2x32-10, math_utils.py:106-189.               */
/* counter = ((tick*1000003 + salt) mod 2^32, id mod 2^32), key = seed */
/* mod 2^32; ten rounds with M = 3529668003 (0xD26275A3, the decimal    */
/* math_utils.py:111 multiplies by) and key bump 0x9E3779B9; the output */
/* is word 0.                                                          */
/* ------------------------------------------------------------------ */
uint32_t or_philox2x32_ref(uint32_t key, uint32_t c0, uint32_t c1)
{
    for (int r = 0; r < 10; ++r) {
        uint64_t prod = (uint64_t)c0 * 3529668003ull;      /* math_utils.py:117 */
        uint32_t hi = (uint32_t)(prod >> 32);              /* :118 */
        uint32_t lo = (uint32_t)prod;                      /* :119 */
        uint32_t n0 = hi ^ key ^ c1;                       /* :120 */
        if (r == 9) return n0;                             /* round 10 stops at :186 */
        c1 = lo;                                           /* :121 */
        key = key + 2654435769u;                           /* :122 */
        c0 = n0;
    }
    return c0; /* unreachable */
}

/* Python floor-mod by 2^32 of a signed 64-bit value (math_utils.py:109-111). */
static uint32_t pymod32(int64_t v)
{
    int64_t m = v % 4294967296ll;
    if (m < 0) m += 4294967296ll;
    return (uint32_t)m;
}

/* rand_uniform_philox(seed, tick, agent_id, salt) after the seed/id
 * injection of model.py:1934-1951: float(c0 % 2^23) / 2^23 (math_utils.py:189). */
double or_rand_uniform_philox(int64_t seed, int64_t tick, int64_t agent_id, int64_t salt)
{
    uint32_t c0 = pymod32(tick * 1000003ll + salt);
    uint32_t c1 = pymod32(agent_id);
    uint32_t k = pymod32(seed);
    uint32_t w = or_philox2x32_ref(k, c0, c1);
    return (double)(w % 8388608u) / 8388608.0;
}

/* ------------------------------------------------------------------ */
/* Philox-4x32-10 (Salmon et al., SC'11): multipliers 0xD2511F53 and    */
/* 0xCD9E8D57, key bumps 0x9E3779B9 and 0xBB67AE85.                     */
/* ------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

/* Per-agent draw of the SIR workloads: key = seed (lo, hi),
 * counter = (tick lo, tick hi, agent_id lo, agent_id hi); word 0. */
uint32_t or_sir_word(uint64_t seed, int64_t tick, int64_t agent_id)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {(uint32_t)tick, (uint32_t)((uint64_t)tick >> 32),
                       (uint32_t)agent_id, (uint32_t)((uint64_t)agent_id >> 32)};
    uint32_t o[4];
    or_philox4x32_10(ctr, key, o);
    return o[0];
}

/* p_k = 1 - (1-beta)^k with the power formed by repeated multiplication in
 * IEEE double, q_0 = 1, q_k = q_{k-1} * (1 - beta). */
void or_infection_probs(double beta, int64_t kmax, double *p_out)
{
    double q = 1.0, b = 1.0 - beta;
    p_out[0] = 0.0;
    for (int64_t k = 1; k <= kmax; ++k) {
        q = q * b;
        p_out[k] = 1.0 - q;
    }
}

/* ------------------------------------------------------------------ */
/* Compartment model of the SIR/SEIR workloads.                         */
/* ------------------------------------------------------------------ */
typedef struct {
    int32_t n_states;        /* 3 (SIR) or 4 (SEIR) */
    int32_t s_code;          /* susceptible code (0) */
    int32_t inf_code;        /* the state neighbours must be in to infect */
    int32_t s_target;        /* state an infected S agent moves to */
    double beta;             /* per-infectious-neighbour probability */
    double trans_p[4];       /* per-state spontaneous transition probability */
    int32_t trans_next[4];   /* ...and its target state */
    uint64_t seed;
} or_model;

#define OR_HIST_W 32

/* Append `who` to a padded history row: first -1 slot, else drop the oldest
 * entry and write at the end. */
static void hist_append(int32_t *row, int32_t who)
{
    for (int j = 0; j < OR_HIST_W; ++j)
        if (row[j] == -1) { row[j] = who; return; }
    memmove(row, row + 1, (OR_HIST_W - 1) * sizeof(int32_t));
    row[OR_HIST_W - 1] = who;
}

static int cmp_pair(const void *a, const void *b)
{
    const int64_t *x = (const int64_t *)a, *y = (const int64_t *)b;
    return (x[0] > y[0]) - (x[0] < y[0]);
}

/*
 * Run `ticks` ticks of the compartment model over a CSR in the caller's
 * numbering.  Slots [0, n_local) are stepped; [n_local, n_total) are ghosts
 * whose states come from the halo each tick (model.py:1414-1417 runs the
 * exchange before the launch; gpu_kernels.py:1030-1036 writes ghosts into
 * both buffers).  Double buffering as in model.py:2302-2408: every agent
 * reads the tick-start snapshot; unchanged agents carry their state over.
 *
 * An S agent counts its neighbours in `inf_code` (k); with one draw u per
 * agent per tick it moves to s_target when u < p_k.  Other states move to
 * trans_next[s] when u < trans_p[s].  u = (word & 0x7FFFFF) / 2^23, the
 * reference's conversion (math_utils.py:189), compared in double.
 *
 * history (optional, n_local x 32 int32, -1 padded) records the first
 * infectious neighbour in CSR order of every agent that gets infected.
 *
 * halo: ghost_ids[n_ghost] names the ghost slots (slot n_local + j has id
 * ghost_ids[j]); halo_ids[n_halo] lists the ids carried by each halo record
 * (any order) and halo_vals[ticks][n_halo] their states for each tick.
 *
 * trace (optional): states of all n_total slots after each tick.
 * Returns 0, or -1 when a halo id names no ghost.
 */
int or_sir_run(const or_model *md, int64_t n_local, int64_t n_total,
               const int64_t *row_ptr, const int32_t *col_idx,
               const int64_t *rng_id, int8_t *state, int32_t *history,
               int64_t tick0, int64_t ticks,
               const int64_t *ghost_ids, const int64_t *halo_ids,
               const int8_t *halo_vals, int64_t n_halo,
               int8_t *trace, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int64_t n_ghost = n_total - n_local;
    int64_t maxdeg = 0;
    for (int64_t i = 0; i < n_local; ++i) {
        int64_t d = row_ptr[i + 1] - row_ptr[i];
        if (d > maxdeg) maxdeg = d;
    }
    double *pk = (double *)malloc(sizeof(double) * (size_t)(maxdeg + 1));
    or_infection_probs(md->beta, maxdeg, pk);

    /* halo record -> slot, resolved by sorting (id, slot) pairs */
    int64_t *halo_slot = NULL;
    if (n_halo > 0) {
        int64_t *pairs = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)n_ghost);
        for (int64_t j = 0; j < n_ghost; ++j) { pairs[2*j] = ghost_ids[j]; pairs[2*j+1] = n_local + j; }
        qsort(pairs, (size_t)n_ghost, 2 * sizeof(int64_t), cmp_pair);
        halo_slot = (int64_t *)malloc(sizeof(int64_t) * (size_t)n_halo);
        for (int64_t h = 0; h < n_halo; ++h) {
            int64_t key[2] = {halo_ids[h], 0};
            int64_t *hit = (int64_t *)bsearch(key, pairs, (size_t)n_ghost, 2 * sizeof(int64_t), cmp_pair);
            if (!hit) { free(pairs); free(halo_slot); free(pk); return -1; }
            halo_slot[h] = hit[1];
        }
        free(pairs);
    }

    int8_t *cur = state;
    int8_t *nxt = (int8_t *)malloc((size_t)n_total);
    memcpy(nxt, cur, (size_t)n_total);

    for (int64_t k = 0; k < ticks; ++k) {
        int64_t t = tick0 + k;
        if (n_halo > 0) {
            const int8_t *hv = halo_vals + (size_t)k * (size_t)n_halo;
            for (int64_t h = 0; h < n_halo; ++h) { cur[halo_slot[h]] = hv[h]; nxt[halo_slot[h]] = hv[h]; }
        }
        #pragma omp parallel for schedule(dynamic, 4096)
        for (int64_t i = 0; i < n_local; ++i) {
            int s = cur[i];
            int ns = s;
            double u = (double)(or_sir_word(md->seed, t, rng_id[i]) & 0x7FFFFFu) / 8388608.0;
            if (s == md->s_code) {
                int64_t kcnt = 0, first = -1;
                for (int64_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
                    int32_t nb = col_idx[e];
                    if (cur[nb] == md->inf_code) { if (first < 0) first = nb; ++kcnt; }
                }
                if (kcnt > 0 && u < pk[kcnt]) {
                    ns = md->s_target;
                    if (history) hist_append(history + (size_t)i * OR_HIST_W, (int32_t)first);
                }
            } else if (s >= 0 && s < 4 && md->trans_p[s] > 0.0) {
                if (u < md->trans_p[s]) ns = md->trans_next[s];
            }
            nxt[i] = (int8_t)ns;
        }
        for (int64_t g = n_local; g < n_total; ++g) nxt[g] = cur[g];
        /* the new buffer becomes the snapshot of the next tick */
        memcpy(cur, nxt, (size_t)n_total);
        if (trace) memcpy(trace + (size_t)k * (size_t)n_total, cur, (size_t)n_total);
    }
    free(nxt);
    free(pk);
    free(halo_slot);
    return 0;
}

/*
 * or_sir_run with the work a CPU implementation would skip taken out; the
 * results are the same bit for bit (tests/test_oracle_fast.py runs both on
 * the same inputs).  It is the tuned CPU baseline, next to or_sir_run's
 * naive one:
 *   - u < p compares in integers: u = m / 2^23 with m = word & 0x7FFFFF, so
 *     u < p exactly when m < ceil(p * 2^23) (scaling by 2^23 is exact);
 *   - an S agent draws first; p_k grows with k, so when m is not below the
 *     threshold of its whole degree no count of infectious neighbours can
 *     infect it and its row is not read; otherwise the scan stops at the
 *     first k whose threshold exceeds m (that neighbour set already holds the
 *     first infectious neighbour in CSR order, which the history records);
 *   - states with no spontaneous transition draw nothing;
 *   - the two state buffers swap instead of being copied each tick.
 * Same arguments and return value as or_sir_run.
 */
int or_sir_run_fast(const or_model *md, int64_t n_local, int64_t n_total,
                    const int64_t *row_ptr, const int32_t *col_idx,
                    const int64_t *rng_id, int8_t *state, int32_t *history,
                    int64_t tick0, int64_t ticks,
                    const int64_t *ghost_ids, const int64_t *halo_ids,
                    const int8_t *halo_vals, int64_t n_halo,
                    int8_t *trace, int nthreads)
{
    /* p_k grows with k only for beta in [0, 1]; elsewhere the skip and the
     * early stop are not valid, and the plain restatement runs instead */
    if (!(md->beta >= 0.0 && md->beta <= 1.0))
        return or_sir_run(md, n_local, n_total, row_ptr, col_idx, rng_id, state, history, tick0, ticks,
                          ghost_ids, halo_ids, halo_vals, n_halo, trace, nthreads);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int64_t n_ghost = n_total - n_local;
    int64_t maxdeg = 0;
    for (int64_t i = 0; i < n_local; ++i) {
        int64_t d = row_ptr[i + 1] - row_ptr[i];
        if (d > maxdeg) maxdeg = d;
    }
    double *pk = (double *)malloc(sizeof(double) * (size_t)(maxdeg + 1));
    uint32_t *thr = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(maxdeg + 1));
    or_infection_probs(md->beta, maxdeg, pk);
    for (int64_t k = 0; k <= maxdeg; ++k) {
        double x = ceil(pk[k] * 8388608.0);
        thr[k] = x <= 0.0 ? 0u : x >= 8388608.0 ? 8388608u : (uint32_t)x;
    }
    thr[0] = 0u;                       /* k = 0 never infects */
    uint32_t tthr[4];
    for (int s = 0; s < 4; ++s) {
        double x = md->trans_p[s] > 0.0 ? ceil(md->trans_p[s] * 8388608.0) : 0.0;
        tthr[s] = x >= 8388608.0 ? 8388608u : (uint32_t)x;
    }

    int64_t *halo_slot = NULL;
    if (n_halo > 0) {
        int64_t *pairs = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)n_ghost);
        for (int64_t j = 0; j < n_ghost; ++j) { pairs[2*j] = ghost_ids[j]; pairs[2*j+1] = n_local + j; }
        qsort(pairs, (size_t)n_ghost, 2 * sizeof(int64_t), cmp_pair);
        halo_slot = (int64_t *)malloc(sizeof(int64_t) * (size_t)n_halo);
        for (int64_t h = 0; h < n_halo; ++h) {
            int64_t key[2] = {halo_ids[h], 0};
            int64_t *hit = (int64_t *)bsearch(key, pairs, (size_t)n_ghost, 2 * sizeof(int64_t), cmp_pair);
            if (!hit) { free(pairs); free(halo_slot); free(pk); free(thr); return -1; }
            halo_slot[h] = hit[1];
        }
        free(pairs);
    }

    int8_t *buf = (int8_t *)malloc((size_t)n_total);
    memcpy(buf, state, (size_t)n_total);
    int8_t *cur = state, *nxt = buf;
    const int s_code = md->s_code, inf = md->inf_code, s_target = md->s_target;

    for (int64_t k = 0; k < ticks; ++k) {
        int64_t t = tick0 + k;
        if (n_halo > 0) {
            const int8_t *hv = halo_vals + (size_t)k * (size_t)n_halo;
            for (int64_t h = 0; h < n_halo; ++h) { cur[halo_slot[h]] = hv[h]; nxt[halo_slot[h]] = hv[h]; }
        }
        #pragma omp parallel for schedule(dynamic, 4096)
        for (int64_t i = 0; i < n_local; ++i) {
            int s = cur[i];
            int ns = s;
            if (s == s_code) {
                int64_t b = row_ptr[i], e = row_ptr[i + 1];
                uint32_t m = or_sir_word(md->seed, t, rng_id[i]) & 0x7FFFFFu;
                if (m < thr[e - b]) {
                    uint32_t kcnt = 0;
                    int32_t first = -1;
                    for (int64_t j = b; j < e; ++j) {
                        int32_t nb = col_idx[j];
                        if (cur[nb] == inf) {
                            if (first < 0) first = nb;
                            if (m < thr[++kcnt]) {
                                ns = s_target;
                                if (history) hist_append(history + (size_t)i * OR_HIST_W, first);
                                break;
                            }
                        }
                    }
                }
            } else if (s >= 0 && s < 4 && tthr[s] != 0u) {
                if ((or_sir_word(md->seed, t, rng_id[i]) & 0x7FFFFFu) < tthr[s]) ns = md->trans_next[s];
            }
            nxt[i] = (int8_t)ns;
        }
        if (n_ghost > 0) memcpy(nxt + n_local, cur + n_local, (size_t)n_ghost);
        int8_t *tmp = cur; cur = nxt; nxt = tmp;
        if (trace) memcpy(trace + (size_t)k * (size_t)n_total, cur, (size_t)n_total);
    }
    if (cur != state) memcpy(state, cur, (size_t)n_total);
    free(buf);
    free(thr);
    free(pk);
    free(halo_slot);
    return 0;
}

/*
 * The reference's own SIR test step (tests/test_simulation.py:24-43) as the
 * generated fused kernel runs it (model.py:2302-2408): float32 state
 * {1=S, 2=I, 3=R}, write buffer W == R at tick start, one
 * rand_uniform_philox(seed, t, logical_id, 1) draw per agent, write-back of
 * [0, n) after the tick.  logical_id is the float32 id of model.py:1085-1096.
 */
void or_ref_sir_run(int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                    const float *logical_id, float *state, int32_t seed,
                    int64_t tick0, int64_t ticks, double p_infection, double p_recovery)
{
    float *w = (float *)malloc(sizeof(float) * (size_t)n);
    memcpy(w, state, sizeof(float) * (size_t)n);
    for (int64_t k = 0; k < ticks; ++k) {
        int64_t t = tick0 + k;
        for (int64_t r = 0; r < n; ++r) {
            int s = (int)state[r];
            double u = or_rand_uniform_philox(seed, t, (int64_t)logical_id[r], 1);
            if (s == 2 && u < p_recovery) {
                w[r] = 3.0f;
            } else if (s == 1) {
                for (int64_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
                    int ns = (int)state[col_idx[e]];
                    if (ns == 2 && u < p_infection) w[r] = 2.0f;
                }
            }
        }
        memcpy(state, w, sizeof(float) * (size_t)n);
    }
    free(w);
}
