"""MI355X-native DEM time step of rusty_particles.

`Simulation`, `set_num_threads` and `hello` mirror the reference's Python
module (src/lib.rs:15-35); the time step runs in the HIP engine
librp_hip.so through the C ABI of include/rp_abi.h.
"""
from .simulation import Ensemble, Simulation, hello, set_num_threads  # noqa: F401
from . import abi  # noqa: F401

__all__ = ["Simulation", "Ensemble", "set_num_threads", "hello", "abi"]
