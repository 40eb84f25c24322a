"""The C ABI library loads and exports every entry point include/sagestep.h
declares; without a GPU every call fails cleanly with a status and a message."""
import os
import re
import subprocess

import pytest

from sagesim_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "sagestep.h")) as f:
        txt = f.read()
    return set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*\*?(ss_[a-z0-9_]+)\s*\(", txt, re.M))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("ss_create", "ss_destroy", "ss_last_error", "ss_replica_create", "ss_simulate",
                 "ss_state_download", "ss_map_build", "ss_halo_unpack", "ss_simulate_halo", "ss_timer_stop"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = _lib.step_lib()            # loads libsagestep.so (no device needed to load)
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.STEP_LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (ss_[a-z0-9_]+)", out))
    missing = _declared() - exported
    assert not missing, missing
    for name in _declared():
        assert hasattr(lib, name)


def _declared_arity():
    """name -> number of parameters of every ss_* declaration in the header."""
    with open(os.path.join(ROOT, "include", "sagestep.h")) as f:
        txt = re.sub(r"//[^\n]*", "", f.read())
    out = {}
    for m in re.finditer(r"(?:int|int64_t|const char \*)\s*\*?(ss_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", txt):
        params = m.group(2).strip()
        out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_ctypes_prototypes_match_header():
    """Every prototype the Python side declares takes as many arguments as the
    C declaration (a mismatch would pass garbage through the ABI)."""
    lib = _lib.step_lib()
    arity = _declared_arity()
    assert len(arity) >= 40
    checked = 0
    for name, n in arity.items():
        fn = getattr(lib, name)
        if fn.argtypes is not None:
            assert len(fn.argtypes) == n, (name, len(fn.argtypes), n)
            checked += 1
    assert checked == len(arity), sorted(set(arity) - {k for k in arity if getattr(lib, k).argtypes is not None})


def test_graph_library_exports():
    lib = _lib.graph_lib()
    for name in ("sg_er", "sg_ba", "sg_ws", "sg_sbm_part", "sg_from_edges", "sg_degree_sort_csr",
                 "sg_scramble_ids", "sg_free"):
        assert hasattr(lib, name)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present: context creation would succeed")
def test_no_device_fails_cleanly():
    from sagesim_amd import Engine
    with pytest.raises(_lib.NativeError) as ex:
        Engine(0)
    assert ex.value.code in (-1, -2)


def test_footprint_counts_push_arrays_only_without_ghosts():
    """A ghost-free replica holds two 16-bit count arrays for the push direction
    (4 bytes per agent); a replica with ghost slots has none."""
    L = _lib.step_lib()
    no_ghost = L.ss_replica_footprint(1000, 0, 10_000, 0, 50)
    with_ghost = L.ss_replica_footprint(1000, 1, 10_000, 0, 50)
    # one ghost slot adds 8 (row_ptr) + 3 (states) bytes and at most one bitmap word per buffer
    assert no_ghost - with_ghost >= 4 * 1000 - 32


def test_footprint_is_positive_and_grows():
    L = _lib.step_lib()
    a = L.ss_replica_footprint(1000, 0, 10_000, 0, 50)
    b = L.ss_replica_footprint(2000, 0, 20_000, 0, 50)
    c = L.ss_replica_footprint(2000, 0, 20_000, 1, 50)
    assert 0 < a < b < c


def test_2bit_halo_layout():
    """Record h sits in bits 2 (h % 4) and 2 (h % 4) + 1 of byte h // 4; a
    ragged tail is zero-padded; unpack inverts pack along the last axis."""
    import numpy as np
    from sagesim_amd import pack_2bit, unpack_2bit
    # 1 | 2 << 2 | 3 << 4 | 0 << 6 = 57; the fifth record alone in byte 1
    assert pack_2bit(np.array([1, 2, 3, 0, 1], np.int8)).tolist() == [57, 1]
    assert pack_2bit(np.array([3, 3, 3, 3], np.int8)).tolist() == [255]
    assert pack_2bit(np.zeros(0, np.int8)).shape == (0,)
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 4, 5, 1023):
        v = rng.integers(0, 4, size=(3, n)).astype(np.int8)
        p = pack_2bit(v)
        assert p.shape == (3, (n + 3) // 4) and p.dtype == np.uint8
        assert np.array_equal(unpack_2bit(p, n), v)
    with pytest.raises(ValueError):
        pack_2bit(np.array([0, 4], np.int8))
