"""Port of the reference's tests/test_space.py:7-63 (NetworkSpace semantics)."""
from sagesim_amd import NetworkSpace


def _space(n=5, ordered=False):
    s = NetworkSpace(ordered=ordered)
    for i in range(n):
        s.add_agent(i)
    return s


def test_add_agents():
    s = _space(3)
    assert s.get_location(0) == set() and s.num_agents == 3


def test_connect_undirected():
    s = _space()
    s.connect_agents(0, 1)
    assert 1 in s.get_neighbors(0) and 0 in s.get_neighbors(1)


def test_connect_directed():
    s = _space()
    s.connect_agents(0, 1, directed=True)
    assert 1 in s.get_neighbors(0) and 0 not in s.get_neighbors(1)


def test_no_duplicates():
    s = _space()
    s.connect_agents(0, 1)
    s.connect_agents(0, 1)
    assert len(s.get_neighbors(0)) == 1


def test_disconnect():
    s = _space()
    s.connect_agents(0, 1)
    s.disconnect_agents(0, 1)
    assert 1 not in s.get_neighbors(0) and 0 not in s.get_neighbors(1)
    s.disconnect_agents(2, 3)           # absent edge: no error


def test_disconnect_directed_keeps_reverse():
    s = _space()
    s.connect_agents(0, 1)
    s.disconnect_agents(0, 1, directed=True)
    assert 1 not in s.get_neighbors(0) and 0 in s.get_neighbors(1)


def test_ordered_keeps_insertion_order():
    s = _space(6, ordered=True)
    for j in (4, 2, 5, 2, 1):
        s.connect_agents(0, j)
    assert s.get_neighbors(0) == [4, 2, 5, 1]
    s.disconnect_agents(0, 5)
    assert s.get_neighbors(0) == [4, 2, 1] and 0 not in s.get_neighbors(5)


def test_bulk_add():
    s = NetworkSpace()
    s.bulk_add_agents(4)
    s.connect_agents(3, 0)
    assert s.get_neighbors(0) == {3}
