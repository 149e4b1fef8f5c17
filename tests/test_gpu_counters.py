"""Step counters against the oracle: the excessive-overlap flag
(simulation.rs:314-315, 361-370, 469-473) and rejected JKR roots
(physics.rs:193-199), step by step; and the reference's warning line, once
per step with excessive overlap (simulation.rs:528-530), from the Python
Simulation."""
import numpy as np
import pytest

from helpers import dense_scene, params, run_engine, run_oracle
from rusty_particles_amd import abi, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("normal,solver", [
    (abi.RP_NORMAL_HERTZ, None), (abi.RP_NORMAL_JKR, None), (abi.RP_NORMAL_JKR, "shortcut"),
    (abi.RP_NORMAL_SJKR, None), (abi.RP_NORMAL_LINEAR, None)], ids=["hertz", "jkr", "jkr-shortcut", "sjkr", "linear"])
def test_overlap_flag_and_jkr_failures_match_oracle(normal, solver):
    # spacing 0.0085 against radii ~0.005: overlaps up to ~0.0015 (> 0.2 r for
    # the closest pairs after jitter), so the flag is set in early steps
    sc = dense_scene(counts=(7, 7, 7), spacing=0.0085, jitter=4e-4, seed=21, vel=0.3)
    p = params(2e-6, (-0.05, -0.05, -0.05), (0.05, 0.05, 0.05), normal=normal,
               tangential=abi.RP_TANGENTIAL_COULOMB, gamma=0.05 if normal in (abi.RP_NORMAL_JKR,
                                                                              abi.RP_NORMAL_SJKR) else 0.0)
    meshes = [scenes.box_mesh((-0.03, -0.03, -0.03), (0.03, 0.03, 0.03))]
    print(f"case normal={normal} solver={solver} n={sc['x'].shape[0]}", flush=True)
    o = run_oracle(p, sc, meshes, steps=0, record=False)
    e = run_engine(p, sc, meshes, steps=0, record=False, solver=solver)
    flags, fails = [], []
    prev = 0
    for s in range(6):
        o.step(1)
        e.step(1)
        st = e.stats()
        cum = o.jkr_failures()
        assert st["excessive_overlap"] == o.overlap_flag(), s
        assert st["jkr_failures"] == cum - prev, s
        prev = cum
        flags.append(st["excessive_overlap"])
        fails.append(st["jkr_failures"])
    assert any(flags)
    assert e.stats()["overlap_steps"] == sum(flags)
    if normal == abi.RP_NORMAL_JKR:
        assert sum(fails) > 0


def test_simulation_prints_the_reference_warning_per_step(capfd):
    from rusty_particles_amd import Simulation
    from rusty_particles_amd.simulation import OVERLAP_WARNING
    sim = Simulation.create(1e-6, -0.1, -0.1, -0.1, 0.1, 0.1, 0.1, 2)
    sim.set_contact_models("hertz", "coulomb")
    m = float(scenes.sphere_mass(np.array([0.01]), 2500.0)[0])
    sim.add_particle(0.0, 0.0, 0.0, 0.01, m)
    sim.add_particle(0.0, 0.015, 0.0, 0.01, m)  # overlap 0.005 > 0.2 r
    for _ in range(4):
        sim.step()
    err = [ln for ln in capfd.readouterr().err.splitlines() if ln.strip()]
    assert err.count(OVERLAP_WARNING) == 4
    assert sim.stats()["overlap_steps"] == 4
