"""Breed: per-breed property and step-function registry (breed.py:8-99).

Step functions are HIP __device__ functions with the reference's argument
order (tick, agent_index, globals..., agent_ids, breeds, locations,
properties...).  `register_step_func` takes a `StepFunction`, HIP source
text, or a function name plus the path of a .hip file holding it (the
reference's (callable, module path) pair, breed.py:68-99).
"""
from collections import OrderedDict
from math import nan
from pathlib import Path
import re
from typing import Any, Dict, List, Optional, Set, Union

_DEVICE_FN = re.compile(r"__device__\s+[\w:<>\s\*&]*?\b([A-Za-z_]\w*)\s*\(")


class StepFunction:
    """A HIP __device__ step function: its name and the source that defines it
    (the source may hold helpers too).  `fused` names a built-in kernel that
    computes exactly the same transition, used instead of the compiled source
    when a model qualifies (see Model.setup)."""

    def __init__(self, name: str, source: str, fused=None):
        self.name = name
        self.source = source
        self.fused = fused

    @property
    def __name__(self):
        return self.name

    @classmethod
    def from_source(cls, source: str, name: Optional[str] = None):
        if name is None:
            found = _DEVICE_FN.findall(source)
            if not found:
                raise ValueError("no __device__ function found in the step source")
            name = found[-1]
        return cls(name, source)

    @classmethod
    def from_file(cls, path, name: str):
        src = Path(path).read_text()
        if not re.search(r"\b" + re.escape(name) + r"\s*\(", src):
            raise ValueError(f"{name} is not defined in {path}")
        return cls(name, src)

    def __repr__(self):
        return f"StepFunction({self.name!r})"


def as_step_function(step_func, module_fpath=None) -> StepFunction:
    if isinstance(step_func, StepFunction):
        return step_func
    if hasattr(step_func, "hip_source"):
        return StepFunction(getattr(step_func, "name", None) or step_func.__name__, step_func.hip_source,
                            getattr(step_func, "fused", None))
    if isinstance(step_func, str):
        if module_fpath is not None and re.fullmatch(r"[A-Za-z_]\w*", step_func):
            return StepFunction.from_file(module_fpath, step_func)
        return StepFunction.from_source(step_func)
    raise TypeError("a step function is a StepFunction, HIP source text, or a function name with its .hip file")


class Breed:
    def __init__(self, name: str) -> None:
        self._properties: "OrderedDict[str, Any]" = OrderedDict()
        self._prop2pos: Dict[str, int] = {}
        self._name: str = name
        self._step_funcs: Dict[int, Any] = {}
        self._breedidx: int = -1
        self._num_properties: int = 0
        self._prop2maxdims: Dict[str, List[int]] = {}
        self._prop2neighbor_visible: Dict[str, bool] = {}
        self._prop2dtype: Dict[str, str] = {}
        self._no_double_buffer_props: Set[str] = set()

    @property
    def name(self) -> str:
        return self._name

    @property
    def properties(self) -> Dict[str, Any]:
        return self._properties

    @property
    def step_funcs(self) -> Dict[int, Any]:
        return self._step_funcs

    @property
    def prop2neighbor_visible(self) -> Dict[str, bool]:
        return self._prop2neighbor_visible

    @property
    def no_double_buffer_props(self) -> Set[str]:
        return self._no_double_buffer_props

    def register_property(self, name: str, default: Union[int, float, List] = nan,
                          max_dims: Optional[List[int]] = None, neighbor_visible: bool = True,
                          dtype: str = "float32") -> None:
        """As breed.py:40-61.  Properties live on the device as float32, the
        reference's storage (internal_utils.py:189-230), unless `dtype` names
        another of float64, int64, int32, int16, int8, uint8."""
        self._properties[name] = default
        self._prop2pos[name] = self._num_properties
        self._num_properties += 1
        self._prop2maxdims[name] = max_dims
        self._prop2neighbor_visible[name] = neighbor_visible
        self._prop2dtype[name] = dtype

    def register_step_func(self, step_func, module_fpath: str = None, priority: int = 0,
                           no_double_buffer: Optional[List[str]] = None):
        """As breed.py:68-99: one step function per priority (a second one at the
        same priority replaces the first); lower priorities run first;
        `no_double_buffer` properties are written in place, so later
        priorities of the same tick see the new values."""
        fn = as_step_function(step_func, module_fpath)
        path = str(Path(module_fpath).resolve()) if module_fpath is not None else "<source>"
        self._step_funcs[priority] = (fn, path)
        if no_double_buffer:
            self._no_double_buffer_props.update(no_double_buffer)
