"""Pins the oracle's RNG functions.

* Philox-2x32-10 of the reference (math_utils.py:106-189): the C oracle
  against the pure-Python big-int transcription (Python `//` and `%`, the
  arithmetic the reference source writes) and the committed golden vectors.
* Philox-4x32-10 of the SIR workloads: the C oracle against two independent
  source oracles compiled on the host, torch's ATen/core/PhiloxRNGEngine.h and
  rocRAND's rocrand_philox4x32_10.h, and against the big-int transcription.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from oracle import philox_py as P

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
MOD32 = 1 << 32


def _golden():
    with open(os.path.join(GOLD, "rng_vectors.json")) as f:
        return json.load(f)


def test_ref_philox_golden():
    L = O.lib()
    for c in _golden()["philox2x32_ref"]:
        assert P.rand_uniform_philox_words(c["seed"], c["tick"], c["id"], c["salt"]) == c["word"]
        u = L.or_rand_uniform_philox(c["seed"], c["tick"], c["id"], c["salt"])
        assert u == c["u"], c
        assert u == (c["word"] % 8388608) / 8388608.0


def test_ref_philox_random_counters_match_bigint():
    L = O.lib()
    rng = np.random.default_rng(1)
    for _ in range(300):
        seed = int(rng.integers(0, 2**31))
        tick = int(rng.integers(0, 2**40))
        aid = int(rng.integers(0, 2**34))
        salt = int(rng.integers(0, 16))
        want = P.rand_uniform_philox_words(seed, tick, aid, salt)
        c0 = (tick * 1000003 + salt) % MOD32
        assert L.or_philox2x32_ref(seed % MOD32, c0, aid % MOD32) == want
        assert L.or_rand_uniform_philox(seed, tick, aid, salt) == P.rand_uniform_philox(seed, tick, aid, salt)


def test_ref_uniform_range():
    """u is in [0, 1) and never reaches 1: p = 1 always passes (the reference's
    known-answer tests with p_infection = p_recovery = 1 rely on it)."""
    L = O.lib()
    us = [L.or_rand_uniform_philox(7, t, i, 1) for t in range(50) for i in range(40)]
    assert min(us) >= 0.0 and max(us) <= (2**23 - 1) / 2**23 < 1.0
    # different salts give different streams (tests/test_math_utils.py:101-108)
    a = [L.or_rand_uniform_philox(7, 3, i, 1) for i in range(100)]
    b = [L.or_rand_uniform_philox(7, 3, i, 2) for i in range(100)]
    assert a != b


def test_sir_philox_golden():
    L = O.lib()
    for c in _golden()["philox4x32_sir"]:
        s, t, i = c["seed"], c["tick"], c["id"]
        assert P.sir_word(s, t, i) == c["word0"]
        assert L.or_sir_word(s, t, i) == c["word0"]
        ctr = [t % MOD32, t >> 32, i % MOD32, i >> 32]
        assert O.philox4x32_10(ctr, [s % MOD32, s >> 32]).tolist() == c["words"]


def _cases():
    rng = np.random.default_rng(5)
    cases = [(0, 0, 0), (7, 5, 123), (2**64 - 1, 2**32 - 1, 2**64 - 1), (5, 0, 2**32)]
    for _ in range(200):
        cases.append((int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2)),
                      int(rng.integers(0, 2**33)), int(rng.integers(0, 2**63))))
    return cases


def _run_pin(exe, cases):
    inp = "".join(f"{s} {t} {i}\n" for s, t, i in cases)
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    return [tuple(int(x) for x in line.split()) for line in out if line.strip()]


@pytest.fixture(scope="module")
def pins(tmp_path_factory):
    d = tmp_path_factory.mktemp("pins")
    exes = {}
    import importlib.util
    spec = importlib.util.find_spec("torch")      # its headers only; torch is not imported
    tinc = os.path.join(spec.submodule_search_locations[0], "include")
    t_exe = str(d / "pin_torch")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-I", tinc, os.path.join(GOLD, "pin_torch_philox.cpp"),
                           "-o", t_exe])
    exes["torch"] = t_exe
    if shutil.which("hipcc"):
        r_exe = str(d / "pin_rocrand")
        subprocess.check_call(["hipcc", "-O1", "-std=c++17", os.path.join(GOLD, "pin_rocrand_philox.cpp"),
                               "-o", r_exe], stderr=subprocess.DEVNULL)
        exes["rocrand"] = r_exe
    return exes


def test_sir_philox_matches_torch_and_rocrand(pins):
    cases = _cases()
    L = O.lib()
    mine = [L.or_sir_word(s, t, i) for s, t, i in cases]
    for name, exe in pins.items():
        got = _run_pin(exe, cases)
        assert [g[0] for g in got] == mine, name
    # word 1 of the same block, through the full 4-word oracle
    got = _run_pin(pins["torch"], cases[:20])
    for (s, t, i), g in zip(cases[:20], got):
        w = O.philox4x32_10([t % MOD32, t >> 32, i % MOD32, i >> 32], [s % MOD32, s >> 32])
        assert int(w[1]) == g[1]


def test_infection_probabilities():
    """p_k = 1 - (1-beta)^k by repeated multiplication: closed forms at beta 0 and 1,
    monotone, and within a few ulps of the exact power."""
    assert O.infection_probs(0.0, 5).tolist() == [0.0] * 6
    assert O.infection_probs(1.0, 3).tolist() == [0.0, 1.0, 1.0, 1.0]
    p = O.infection_probs(0.03, 400)
    assert (np.diff(p) >= 0).all()
    from fractions import Fraction
    b = Fraction(1) - Fraction(0.03)
    for k in (1, 2, 10, 100, 400):
        exact = 1 - b ** k
        assert abs(Fraction(p[k]) - exact) < Fraction(1, 2**45)
