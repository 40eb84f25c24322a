"""Port of the reference's tests/test_model_lifecycle.py:39-87 (host-side
lifecycle; the setup case runs on the GPU)."""
import pytest

from sagesim_amd import Breed, Model, NetworkSpace, StepFunction

_INC = StepFunction.from_source("""
__device__ void _increment_step(auto tick, auto agent_index, auto agent_ids, auto breeds, auto locations,
                                auto counter) {
    counter[agent_index] = counter[agent_index] + 1;
}""")


class MinimalBreed(Breed):
    def __init__(self):
        super().__init__("Minimal")
        self.register_property("counter", 0)
        self.register_step_func(_INC, __file__, priority=0)


class MinimalModel(Model):
    def __init__(self):
        super().__init__(NetworkSpace())
        self._breed = MinimalBreed()
        self.register_breed(self._breed)

    def create_agent(self, counter=0):
        return self.create_agent_of_breed(self._breed, counter=counter)


def test_register_breed_and_create_agent():
    m = MinimalModel()
    assert m.create_agent() == 0 and m.num_agents == 1


def test_create_multiple_agents():
    m = MinimalModel()
    assert [m.create_agent() for _ in range(5)] == [0, 1, 2, 3, 4] and m.num_agents == 5


def test_set_get_agent_property():
    m = MinimalModel()
    a = m.create_agent(counter=0)
    m.set_agent_property_value(a, "counter", 42)
    assert m.get_agent_property_value(a, "counter") == 42


def test_register_global_property():
    m = MinimalModel()
    m.register_global_property("rate", 0.5)
    assert m.get_global_property_value("rate") == pytest.approx(0.5)


def test_set_global_property():
    m = MinimalModel()
    m.register_global_property("rate", 0.5)
    m.set_global_property_value("rate", 0.9)
    assert m.get_global_property_value("rate") == pytest.approx(0.9)


def test_set_seed():
    m = MinimalModel()
    m.set_seed(12345)
    assert m._seed == 12345


def test_breed_after_agents_is_an_error():
    m = MinimalModel()
    m.create_agent()
    with pytest.raises(Exception):
        m.register_breed(Breed("Late"))


def test_unknown_property_is_an_error():
    m = MinimalModel()
    with pytest.raises(ValueError):
        m.create_agent_of_breed(m._breed, nonexistent=3)
    m.create_agent()
    with pytest.raises(ValueError):
        m.set_agent_property_value(0, "nonexistent", 1)


def test_generated_source_has_reference_argument_order():
    m = MinimalModel()
    m.register_global_property("rate", 0.5)
    m.register_global_property("table", [1.0, 2.0, 3.0])
    m.create_agent()
    m._order = [0]
    m._prop_width = {"counter": 1}
    src, names = m.generate_source()
    assert names == ["ss_prio_0"]
    assert "_increment_step(tick, agent_index, G0, G1, agent_ids, breeds, locations, P0)" in src
    assert "const double G0 = A->globals[0];" in src and "GlobalArray G1{A->globals + 1, 3}" in src


@pytest.mark.gpu
def test_model_setup():
    m = MinimalModel()
    m.create_agent()
    m.setup(use_gpu=True)
    assert m._is_setup
    m.simulate(3)
    assert m.get_agent_property_value(0, "counter") == 3
