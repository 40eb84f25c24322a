"""The device runtime of compiled step functions (sagesim_amd/csrc/step_rt.h)
and the wrapper kernels Model.generate_source writes, built for the host with
g++ (`__device__`/`__global__` defined away, the launch grid walked by a
loop), against the golden RNG vectors and the oracle.  No GPU is needed, so
the semantics of the general step-function path are pinned on the CPU."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from oracle import philox_py as P
from sagesim_amd import graphs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RT = os.path.join(ROOT, "sagesim_amd", "csrc", "step_rt.h")
GOLD = os.path.join(ROOT, "tests", "golden", "rng_vectors.json")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is needed")

_HOST_PRELUDE = """
#include <cstdio>
#include <cstdint>
#include <vector>
#define __device__
#define __global__
#define __restrict__
struct SSDim { unsigned x = 0; };
static SSDim blockIdx, blockDim, threadIdx;
"""


def _build(tmp_path, body, name):
    src = tmp_path / f"{name}.cpp"
    src.write_text(body)
    exe = tmp_path / name
    r = subprocess.run(["g++", "-std=c++20", "-O1", "-I", str(tmp_path), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return str(exe)


def _rt_copy(tmp_path):
    shutil.copy(RT, tmp_path / "sagesim_rt.h")


def test_runtime_rng_matches_golden_vectors(tmp_path):
    _rt_copy(tmp_path)
    with open(GOLD) as f:
        gold = json.load(f)
    rng = np.random.default_rng(5)
    cases = [(c["seed"], c["tick"], c["id"], c["salt"]) for c in gold["philox2x32_ref"]]
    cases += [(int(rng.integers(0, 2**31)), int(rng.integers(0, 2**40)), int(rng.integers(0, 2**34)),
               int(rng.integers(0, 16))) for _ in range(200)]
    # the runtime keys Philox-4x32 with (seed, 0): vectors with a 32-bit seed
    sir = [c for c in gold["philox4x32_sir"] if c["seed"] < 2**32]
    assert len(sir) >= 5
    lines = [_HOST_PRELUDE, '#include "sagesim_rt.h"', "int main() {"]
    for s, t, i, salt in cases:
        # a seed above 2^31 is an int32 in the reference (model.py:1104): its low 32 bits count
        lines.append(f"  std::printf(\"%.17g\\n\", ss_rt_uniform((int)(unsigned){s}u, {t}LL, {i}LL, {salt}LL));")
    for c in sir:
        lines.append(f"  {{ AgentIndex a{{0, 0.f, (int){c['seed']}u, (long long){c['id']}ULL}}; "
                     f"std::printf(\"%u\\n\", rand_u23_philox4x32((long long){c['tick']}ULL, a)); }}")
    lines.append("  return 0; }")
    out = subprocess.run([_build(tmp_path, "\n".join(lines), "rng")], capture_output=True, text=True,
                         check=True).stdout.split()
    for (s, t, i, salt), got in zip(cases, out[:len(cases)]):
        assert float(got) == P.rand_uniform_philox(s, t, i, salt), (s, t, i, salt)
    for c, got in zip(gold["philox2x32_ref"], out):
        assert float(got) == c["u"]
    for c, got in zip(sir, out[len(cases):]):
        assert int(got) == c["word0"] & 0x7FFFFF, c


def _sir_model(g, init, p_inf, p_rec, seed):
    from sagesim_amd import Breed, Model, NetworkSpace
    from sagesim_amd.steps import reference_sir_step
    b = Breed("SIR")
    b.register_property("state", 1)
    b.register_step_func(reference_sir_step(), priority=0)
    m = Model(NetworkSpace())
    m.register_breed(b)
    m.register_global_property("p_infection", p_inf)
    m.register_global_property("p_recovery", p_rec)
    m.set_seed(seed)
    for i in range(g.n):
        m.create_agent_of_breed(b, state=int(init[i]))
    u = np.repeat(np.arange(g.n), np.diff(g.row_ptr))
    m.get_space().connect_agents_bulk(u, g.col_idx, directed=True)
    return m


def test_generated_wrapper_matches_oracle_on_host(tmp_path):
    """The wrapper Model.generate_source writes around the reference SIR step
    (tests/test_simulation.py:24-43), run tick by tick on the host with the
    double-buffer write-back, equals the oracle's restatement bit for bit."""
    _rt_copy(tmp_path)
    g = graphs.erdos_renyi(3000, 0.004, 12)
    rng = np.random.default_rng(1)
    init = np.where(rng.random(g.n) < 0.02, 2, 1)
    m = _sir_model(g, init, 0.3, 0.15, 2024)
    order, rp, ci, logical, groups = m._host_layout()
    src, names = m.generate_source(groups)
    assert names == ["ss_prio_0"]
    assert src.startswith("#define SS_RT_OFF32 1")     # 32-bit property offsets (arrays below 2 GiB)
    (tmp_path / "gen.hip").write_text(src)
    rp.astype(np.int64).tofile(tmp_path / "rp.bin")
    ci.astype(np.int32).tofile(tmp_path / "ci.bin")
    st0 = m._host_props["state"].astype(np.float32)
    st0.tofile(tmp_path / "st.bin")
    n, ticks = g.n, 30
    main = _HOST_PRELUDE + f"""
#include "gen.hip"
static void rd(const char *p, void *d, size_t b) {{ FILE *f = std::fopen(p, "rb"); if (std::fread(d, 1, b, f) != b) throw 1; std::fclose(f); }}
int main() {{
    const long long n = {n};
    std::vector<long long> rp(n + 1); rd("{tmp_path}/rp.bin", rp.data(), 8 * (n + 1));
    std::vector<int> ci(rp[n]); rd("{tmp_path}/ci.bin", ci.data(), 4 * rp[n]);
    std::vector<float> r(n), w(n), ids(n), br(n, 0.f);
    std::vector<long long> ids64(n);
    rd("{tmp_path}/st.bin", r.data(), 4 * n); w = r;
    for (long long i = 0; i < n; ++i) {{ ids[i] = (float)i; ids64[i] = i; }}
    double glob[2] = {{0.3, 0.15}};
    SSProgArgs A{{}};
    A.n_agents = n; A.row_ptr = rp.data(); A.col_idx = ci.data(); A.agent_ids = ids.data();
    A.logical_ids = ids.data(); A.breeds = br.data(); A.globals = glob; A.seed = 2024; A.logical_ids64 = ids64.data();
    A.prop_r[0] = r.data(); A.prop_w[0] = w.data();
    blockDim.x = 1;
    for (long long t = 0; t < {ticks}; ++t) {{
        for (long long i = 0; i < n; ++i) {{ blockIdx.x = (unsigned)i; ss_prio_0(&A, t, 0, n, nullptr); }}
        r = w;                                   // write-back after the tick
        A.prop_r[0] = r.data();
    }}
    for (long long i = 0; i < n; ++i) std::printf("%g\\n", r[i]);
}}
"""
    exe = _build(tmp_path, main, "wrap")
    got = np.array(subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split(), np.float32)
    want = O.ref_sir_run(g.row_ptr, g.col_idx, np.arange(n, dtype=np.float32), st0, 2024, ticks, 0.3, 0.15)
    assert np.array_equal(got, want), int((got != want).sum())
    assert 0 < (want == 3).sum() < n
