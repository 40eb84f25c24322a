"""Port of the reference's tests/test_breed.py:15-54 (registry semantics)."""
from pathlib import Path

import pytest

from sagesim_amd import Breed, StepFunction

_DUMMY = StepFunction.from_source(
    "__device__ void _dummy_step(auto tick, auto agent_index, auto agent_ids, auto breeds, auto locations, auto prop_a) {}")


def test_register_property_scalar():
    breed = Breed("TestBreed")
    breed.register_property("health", 100)
    assert "health" in breed.properties and breed.properties["health"] == 100


def test_register_property_list_default():
    breed = Breed("TestBreed")
    breed.register_property("history", [0.0] * 10)
    assert breed.properties["history"] == [0.0] * 10 and len(breed.properties["history"]) == 10


def test_register_step_func():
    breed = Breed("TestBreed")
    breed.register_property("prop_a", 0)
    breed.register_step_func(_DUMMY, __file__, priority=0)
    assert 0 in breed.step_funcs
    func, path = breed.step_funcs[0]
    assert func is _DUMMY and func.name == "_dummy_step"
    assert path == str(Path(__file__).resolve())


def test_register_step_func_no_double_buffer():
    breed = Breed("TestBreed")
    breed.register_property("counter", 0)
    breed.register_property("result", 0)
    breed.register_step_func(_DUMMY, __file__, priority=0, no_double_buffer=["counter"])
    assert "counter" in breed.no_double_buffer_props and "result" not in breed.no_double_buffer_props


def test_property_neighbor_visible():
    breed = Breed("TestBreed")
    breed.register_property("state", 1, neighbor_visible=True)
    breed.register_property("internal", 0, neighbor_visible=False)
    assert breed.prop2neighbor_visible["state"] and not breed.prop2neighbor_visible["internal"]


def test_step_function_forms(tmp_path):
    src = "__device__ void my_step(auto tick, auto agent_index) {}\n"
    assert StepFunction.from_source(src).name == "my_step"
    f = tmp_path / "steps.hip"
    f.write_text(src)
    b = Breed("B")
    b.register_step_func("my_step", str(f), priority=2)
    assert b.step_funcs[2][0].name == "my_step"
    b.register_step_func(src, priority=2)          # same priority replaces (breed.py:95)
    assert len(b.step_funcs) == 1
    with pytest.raises(ValueError):
        StepFunction.from_file(str(f), "absent_step")
    with pytest.raises(TypeError):
        b.register_step_func(42)
