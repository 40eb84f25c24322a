"""sagesim_amd: MI355X-native per-tick agent step path for SAGESim models.

Native code: libsagestep.so (HIP kernels for gfx950 behind the C ABI in
include/sagestep.h) and libsagegraph.so (host graph generators).  Build them
with `python -m sagesim_amd.build`.

The reference's API (Model, Breed, NetworkSpace) runs user step functions
written as HIP device source (compiled with hipRTC) or, for the built-in
compartment steps, the hand-written fused tick kernel.
"""
from .engine import Engine, Replica, sir_model, seir_model, pack_2bit, unpack_2bit  # noqa: F401
from . import graphs  # noqa: F401
from .breed import Breed, StepFunction  # noqa: F401
from .space import NetworkSpace  # noqa: F401
from .model import Model  # noqa: F401

__version__ = "0.1.0"
__all__ = ["Model", "Breed", "NetworkSpace", "StepFunction", "Engine", "Replica", "sir_model", "seir_model",
           "pack_2bit", "unpack_2bit"]
