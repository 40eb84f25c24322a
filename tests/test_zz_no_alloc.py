"""ss_halo_pack_noalloc (the halo pack that allocates nothing) against
ss_halo_pack through the C ABI: two calls in a row, the second smaller, then
a size beyond the fixed staging refused on the host."""
import ctypes

import numpy as np
import pytest

from sagesim_amd import Engine, Replica, graphs
from sagesim_amd._lib import ptr, step_lib
from sagesim_amd.engine import sir_model

pytestmark = pytest.mark.gpu

MAX_RECORDS = 1 << 20


def say(*a):
    print("[case]", *a, flush=True)


def test_noalloc_pack_matches_pack():
    L = step_lib()
    g = graphs.barabasi_albert(300000, 4, 21)
    rng = np.random.default_rng(5)
    st = (rng.random(g.n) < 0.01).astype(np.int8)
    with Engine(0) as eng:
        rep = Replica(eng, g, st, sir_model(0.2, 0.1), 77)
        eng.simulate(12)
        eng.sync()
        for n in (250000, 3001):
            say("noalloc pack", n)
            slots = np.ascontiguousarray(rng.integers(0, g.n, size=n).astype(np.int32))
            want = eng.halo_pack(rep.rep, slots)
            got = np.full(n, -7, np.int8)
            rc = L.ss_halo_pack_noalloc(eng._h, rep.rep, ptr(slots), n, ptr(got))
            assert rc == 0, L.ss_last_error(eng._h).decode()
            assert np.array_equal(got, want)
            assert set(np.unique(got).tolist()) <= {0, 1, 2} and (got == 1).any()
        say("noalloc pack: size beyond the staging")
        big = np.zeros(MAX_RECORDS + 1, np.int32)
        out = np.zeros(MAX_RECORDS + 1, np.int8)
        rc = L.ss_halo_pack_noalloc(eng._h, rep.rep, ptr(big), big.shape[0], ptr(out))
        assert rc == -1 and "exceeds" in L.ss_last_error(eng._h).decode()
        bad = np.array([0, g.n + 5], np.int32)
        rc = L.ss_halo_pack_noalloc(eng._h, rep.rep, ptr(bad), 2, ctypes.cast(out.ctypes.data, ctypes.c_void_p))
        assert rc == -4
        rep.remove()
