"""Writes tests/golden/rng_vectors.json from the pure-Python big-int Philox
transcriptions (oracle/philox_py.py), which follow math_utils.py:106-189's
integer arithmetic literally.  The fixture freezes those words so that any
later change to the oracle, the Python transcription or the device code is
caught against the same numbers.

Run: python tests/golden/make_rng_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import philox_py as P  # noqa: E402

MOD32 = 1 << 32

# (seed, tick, agent_index, salt) for the reference draw: zero counters,
# all-ones words, tick*1000003 + salt crossing 2^32, ids above 2^32 and
# float32-rounded ids, negative-free seeds as np.random.randint(0, 2**31) draws.
REF_CASES = [
    (0, 0, 0, 0), (0, 0, 0, 1), (42, 0, 0, 1), (42, 1, 0, 1), (42, 5, 9, 1),
    (MOD32 - 1, MOD32 - 1, MOD32 - 1, MOD32 - 1),
    (999, 4294, 7, 1), (999, 4295, 7, 1),           # 4295 * 1000003 > 2^32
    (123456789, 10**6, 2**24, 3), (2**31 - 1, 2**33 + 5, 2**40 + 17, 2),
    (7, 3, 16777217, 1), (2024, 99, 10**7 - 1, 1),
]

# (seed, tick, agent_id) for the Philox-4x32-10 draw of the SIR workloads
SIR_CASES = [
    (0, 0, 0), (7, 0, 0), (7, 5, 123), (2024, 199, 9_999_999), (1, 1, 1),
    (2**64 - 1, 2**32 - 1, 2**64 - 1), (2**63 + 12345, 2**32 + 7, 2**40 + 3),
    (31, 149, 4_000_000 - 1), (5, 0, 2**32), (0xDEADBEEFCAFEF00D, 300, 77),
]


def main():
    ref = [{"seed": s, "tick": t, "id": i, "salt": c,
            "word": P.rand_uniform_philox_words(s, t, i, c),
            "u": P.rand_uniform_philox(s, t, i, c)} for (s, t, i, c) in REF_CASES]
    sir = [{"seed": s, "tick": t, "id": i, "word0": P.sir_word(s, t, i),
            "words": P.philox4x32_10([t % MOD32, t >> 32, i % MOD32, i >> 32], [s % MOD32, s >> 32])}
           for (s, t, i) in SIR_CASES]
    with open(os.path.join(HERE, "rng_vectors.json"), "w") as f:
        json.dump({"philox2x32_ref": ref, "philox4x32_sir": sir}, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
