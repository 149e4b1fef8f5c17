"""CPU oracle for the DEM step (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package.  It wraps liboracle.so (rp_oracle.cpp), a C++
restatement of the reference's Simulation::step_cpu, and mirrors the
reference's Python `Simulation` surface so the reference's own tests can be
restated against it.
"""
from .sim import OracleSimulation, build, lib, read_stl_numpy  # noqa: F401
