"""Pure-Python big-integer Philox restatements.  TEST INFRASTRUCTURE ONLY.

These follow the published algorithm with Python's unbounded integers and its
floor-division/modulo semantics, i.e. the arithmetic the reference writes in
sagesim/math_utils.py:106-189 (`product // 4294967296`, `% 4294967296`).
They share no code with oracle/sir_oracle.c or the HIP kernels and are used to
pin both on edge-case counters (zero, all-ones, 2^32 wrap of tick*1000003).
"""

MOD32 = 4294967296


def rand_uniform_philox_words(seed, tick, agent_index, salt):
    """Word 0 of Philox-2x32-10 as math_utils.py:106-186 computes it."""
    c0 = (int(tick) * 1000003 + int(salt)) % MOD32
    c1 = int(agent_index) % MOD32
    key = int(seed) % MOD32
    for rnd in range(10):
        product = c0 * 3529668003
        hi = (product // MOD32) % MOD32
        lo = product % MOD32
        c0 = (hi ^ key ^ c1) % MOD32
        if rnd == 9:
            break
        c1 = lo
        key = (key + 2654435769) % MOD32
    return c0


def rand_uniform_philox(seed, tick, agent_index, salt):
    """math_utils.py:189: float(c0 % 8388608) / 8388608.0"""
    return float(rand_uniform_philox_words(seed, tick, agent_index, salt) % 8388608) / 8388608.0


def philox4x32_10(ctr, key):
    """Philox-4x32-10 (Salmon et al., SC'11) on a 4-word counter and 2-word key."""
    c = [int(x) % MOD32 for x in ctr]
    k0, k1 = int(key[0]) % MOD32, int(key[1]) % MOD32
    for _ in range(10):
        p0 = 0xD2511F53 * c[0]
        p1 = 0xCD9E8D57 * c[2]
        c = [((p1 // MOD32) ^ c[1] ^ k0) % MOD32, p1 % MOD32,
             ((p0 // MOD32) ^ c[3] ^ k1) % MOD32, p0 % MOD32]
        k0 = (k0 + 0x9E3779B9) % MOD32
        k1 = (k1 + 0xBB67AE85) % MOD32
    return c


def sir_word(seed, tick, agent_id):
    """Draw of the SIR workloads: key = seed, counter = (tick, agent_id) as 64-bit halves."""
    seed, tick, agent_id = int(seed), int(tick), int(agent_id)
    return philox4x32_10([tick % MOD32, (tick >> 32) % MOD32,
                          agent_id % MOD32, (agent_id >> 32) % MOD32],
                         [seed % MOD32, (seed >> 32) % MOD32])[0]
