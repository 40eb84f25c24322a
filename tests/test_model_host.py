"""Host half of Model.setup, no GPU: CSR rows from per-edge and bulk
connections, neighbour visibility, breed-local array registration, and the
wrapper source generate_source writes for breed-local arrays and the
kernel-extras hooks (model.py:398-409, :620-651, :892-921, :2302-2353)."""
import numpy as np
import pytest

from sagesim_amd import Breed, Model, NetworkSpace, StepFunction

_STEP = StepFunction.from_source("""
__device__ void s_step(auto tick, auto agent_index, auto agent_ids, auto breeds, auto locations, auto x) {
    x[agent_index] = x[agent_index] + 1;
}""")


def _model(ordered=True, n=6):
    b = Breed("A")
    b.register_property("x", 0.0)
    b.register_step_func(_STEP, priority=0)
    m = Model(NetworkSpace(ordered=ordered))
    m.register_breed(b)
    for _ in range(n):
        m.create_agent_of_breed(b)
    return m, b


def test_rows_keep_container_order_then_bulk_edges():
    m, _ = _model(ordered=True)
    sp = m.get_space()
    sp.connect_agents(0, 4)
    sp.connect_agents(0, 2)
    sp.connect_agents(3, 3)                         # a self-loop, as connect_agents keeps it
    sp.connect_agents_bulk([0, 5, 1], [5, 1, 0])     # undirected: 0->5, 5->0, 5->1, 1->5, 1->0, 0->1
    _, rp, ci, _, _ = m._host_layout()
    rows = [ci[rp[i]:rp[i + 1]].tolist() for i in range(6)]
    assert rows[0] == [4, 2, 5, 1]                   # per-edge first, in insertion order, then bulk in order
    assert rows[1] == [5, 0]
    assert rows[2] == [0]
    assert rows[3] == [3]
    assert rows[4] == [0]
    assert rows[5] == [0, 1]
    # an edge given again in bulk is not listed twice
    sp.connect_agents_bulk([0], [4])
    _, rp, ci, _, _ = m._host_layout()
    assert ci[rp[0]:rp[1]].tolist() == [4, 2, 5, 1]


def test_set_rows_in_set_order():
    m, _ = _model(ordered=False, n=70)
    sp = m.get_space()
    for j in (65, 3, 40, 9, 64):
        sp.connect_agents(0, j, directed=True)
    _, rp, ci, _, _ = m._host_layout()
    assert ci[rp[0]:rp[1]].tolist() == list(sp.get_location(0))      # the set's own iteration order


def test_neighbor_visibility_rules():
    m, b = _model()
    assert m._prop_visible["breed"] is True and m._prop_visible["locations"] is False
    assert m._prop_visible["x"] is True
    m.set_property_neighbor_visible("breed", False)
    assert m._prop_visible["breed"] is False
    with pytest.raises(ValueError):
        m.set_property_neighbor_visible("nope", True)
    m._is_setup = True
    with pytest.raises(RuntimeError):
        m.set_property_neighbor_visible("x", False)


def test_breed_local_array_registration_checks():
    m, b = _model()
    assert m.register_breed_local_array("buf", b, (3,)) == 0
    assert m.register_breed_local_array("mat", b, (2, 4)) == 1
    for bad in [("buf", b, (1,)), ("bad name", b, (1,)), ("z", b, (1, 2, 3)), ("z", b, (0,)),
                ("z", Breed("other"), (1,))]:
        with pytest.raises(ValueError):
            m.register_breed_local_array(*bad)


class _Extras(Model):
    def _get_extra_kernel_config(self):
        return {"extra_kernel_params": ["int *hits", "buf"],
                "post_breed_step_code": [(["hits[_real_idx] += 1;"], False),
                                         (["buf[breed_id] += 1.0f;"], True),
                                         (["buf[7] = 0.0f;"], False, 1)]}


def test_generated_wrapper_with_blas_and_extras():
    a, c = Breed("A"), Breed("C")
    for b in (a, c):
        b.register_property("x", 0.0)
    a.register_step_func(_STEP, priority=0)
    c.register_step_func(_STEP, priority=0)
    a.register_step_func(_STEP, priority=3)
    m = _Extras(NetworkSpace())
    m.register_breed(a)
    m.register_breed(c)
    m.create_agent_of_breed(a)
    m.create_agent_of_breed(c)
    m.register_breed_local_array("mem", a, (4,))
    m._host_layout()
    m._bla_ids = [(1, 2, 1)]
    src, names = m.generate_source()
    assert names == ["ss_prio_0", "ss_prio_1"]
    # property, then the breed-local array pair, in the step call
    assert "s_step(tick, agent_index, agent_ids, breeds, locations, P0, B0, B0_idx);" in src
    assert "const BlaRows1<float> B0{(const float *)A->prop_r[1], (float *)A->prop_w[1], 1, 4};" in src
    assert "const IdxMap B0_idx{(const int *)A->prop_r[2]};" in src
    # own-slot write of the double-buffered scalar property, stored before post code
    assert "OwnSlot<float> O0{};" in src and "ss_flush(P0, O0);" in src
    assert src.index("ss_flush(P0, O0);") < src.index("hits[_real_idx] += 1;")
    assert "int * const hits = (int *)A->extra[0];" in src and "float * const buf = (float *)A->extra[1];" in src
    # once per breed: breed 0 in group 0 and breed 1 in group 0, not breed 0 again in group 1
    assert src.count("buf[breed_id] += 1.0f;") == 2
    # only priority index 1
    assert src.count("buf[7] = 0.0f;") == 1 and src.index("buf[7] = 0.0f;") > src.index("ss_prio_1")
    assert src.count("hits[_real_idx] += 1;") == 3
