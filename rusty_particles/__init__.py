"""The reference's Python module name (src/lib.rs:24-35, `#[pymodule] fn
rusty_particles`): `Simulation`, `set_num_threads` and `hello`, so that
scripts written against the reference (`import rusty_particles`,
`rusty_particles.Simulation.create(...)`) run unchanged on the MI355X engine.
The implementation lives in rusty_particles_amd; nothing here adds behaviour.
"""
from rusty_particles_amd.simulation import Simulation, hello, set_num_threads

__all__ = ["Simulation", "set_num_threads", "hello"]
