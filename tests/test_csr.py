"""Port of the reference's tests/test_csr.py:10-47 (golden CSR arrays)."""
import numpy as np
from numpy.testing import assert_array_equal

from sagesim_amd.model import build_csr_from_ragged, build_csr_values_only


def test_build_csr_from_lists():
    offsets, values = build_csr_from_ragged([[5, 2], [8, 3, 1], [], [7, 4, 9, 6]])
    assert_array_equal(offsets, [0, 2, 5, 5, 9])
    assert_array_equal(values, [5, 2, 8, 3, 1, 7, 4, 9, 6])


def test_build_csr_from_sets():
    offsets, values = build_csr_from_ragged([{5, 2}, {8}])
    assert_array_equal(offsets, [0, 2, 3])
    assert set(values[:2]) == {5, 2} and values[2] == 8


def test_build_csr_empty():
    offsets, values = build_csr_from_ragged([])
    assert_array_equal(offsets, [0])
    assert len(values) == 0


def test_build_csr_single_agent():
    offsets, values = build_csr_from_ragged([[1, 2, 3]])
    assert_array_equal(offsets, [0, 3])
    assert_array_equal(values, [1, 2, 3])


def test_build_csr_mixed_empty():
    offsets, values = build_csr_from_ragged([[1], [], [2, 3]])
    assert_array_equal(offsets, [0, 1, 1, 3])
    assert_array_equal(values, [1, 2, 3])


def test_build_csr_values_only():
    offsets, _ = build_csr_from_ragged([[10, 20], [30]])
    assert_array_equal(build_csr_values_only([[100, 200], [300]], offsets), [100, 200, 300])


def test_generators_build_the_same_csr_as_ragged_lists():
    from sagesim_amd import graphs
    rng = np.random.default_rng(0)
    u, v = rng.integers(0, 50, 200), rng.integers(0, 50, 200)
    g = graphs.from_edges(50, u, v)
    rows = [set() for _ in range(50)]
    for a, b in zip(u.tolist(), v.tolist()):
        if a != b:
            rows[a].add(b)
            rows[b].add(a)
    offsets, values = build_csr_from_ragged(rows)
    assert_array_equal(g.row_ptr, offsets)
    # sets are written in their own iteration order, the generator's rows ascending
    for i in range(50):
        row = values[offsets[i]:offsets[i + 1]]
        assert row.tolist() == list(rows[i])
        assert_array_equal(np.sort(row), g.col_idx[g.row_ptr[i]:g.row_ptr[i + 1]])
