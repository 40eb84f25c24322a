"""sagesim_amd: MI355X-native per-tick agent step path for SAGESim models.

Native code: libsagestep.so (HIP kernels for gfx950 behind the C ABI in
include/sagestep.h) and libsagegraph.so (host graph generators).  Build them
with `python -m sagesim_amd.build`.
"""
from .engine import Engine, Replica, sir_model, seir_model  # noqa: F401
from . import graphs  # noqa: F401

__version__ = "0.1.0"
