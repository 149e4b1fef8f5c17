"""The Python surface (src/lib.rs:15-277) without a GPU: argument handling,
error types, host-side state and output format."""
import numpy as np
import pytest

import rusty_particles_amd as rp
from rusty_particles_amd import abi, native, simulation


def test_hello_and_threads():
    assert rp.hello() == "Hello from Rust!"
    simulation._threads_set = False
    rp.set_num_threads(4)
    with pytest.raises(RuntimeError):
        rp.set_num_threads(2)


def test_model_names_case_insensitive_and_invalid():
    sim = rp.Simulation.create(1e-3, -1, -1, -1, 1, 1, 1, 0)
    for nm in ("Hertz", "hertzian", "LINEAR", "spring_dashpot", "hysteretic", "JKR", "sJKR", "simplified_jkr"):
        for tm in ("Mindlin", "coulomb", "Linear", "spring_coulomb"):
            sim.set_contact_models(nm, tm)
    with pytest.raises(ValueError):
        sim.set_contact_models("invalid", "coulomb")
    with pytest.raises(ValueError):
        sim.set_contact_models("hertz", "nope")


def test_particles_and_index_errors():
    sim = rp.Simulation.create(1e-3, -1, -1, -1, 1, 1, 1, 2)
    sim.add_particle(0.1, 0.2, 0.3, 0.05, 1.0)
    sim.add_particle(0.0, 0.0, 0.0, 0.05, 2.0, fixed=True)
    assert sim.get_particle_count() == 2
    assert sim.get_particle_position(0) == (0.1, 0.2, 0.3)
    sim.set_particle_velocity(1, 1.0, 2.0, 3.0)
    sim.set_particle_angular_velocity(1, 4.0, 5.0, 6.0)
    assert sim.get_particle_velocity(1) == (1.0, 2.0, 3.0)
    assert sim.get_particle_angular_velocity(1) == (4.0, 5.0, 6.0)
    for f in (sim.get_particle_position, sim.get_particle_velocity, sim.get_particle_angular_velocity):
        with pytest.raises(IndexError):
            f(2)
    with pytest.raises(IndexError):
        sim.set_particle_velocity(5, 0, 0, 0)


def test_step_without_device_raises_runtime_error():
    if native.device_count() > 0:
        pytest.skip("a HIP device is present")
    sim = rp.Simulation.create(1e-3, -1, -1, -1, 1, 1, 1, 1)
    sim.add_particle(0, 0, 0, 0.1, 1.0)
    with pytest.raises(RuntimeError):
        sim.step()
    with pytest.raises(RuntimeError):
        sim.enable_gpu()


def test_run_step_count_uses_f32(monkeypatch):
    """steps = ceil(f32(duration) / f32(dt)) (src/lib.rs:183)."""
    seen = []
    sim = rp.Simulation.create(1e-4, -1, -1, -1, 1, 1, 1, 1)
    monkeypatch.setattr(sim, "_steps", lambda n: seen.append(n))
    sim.run(1.0)
    sim.run(0.1)
    assert seen == [int(np.ceil(np.float32(1.0) / np.float32(1e-4))),
                    int(np.ceil(np.float32(0.1) / np.float32(1e-4)))]


def test_add_mesh_io_error(tmp_path):
    sim = rp.Simulation.create(1e-3, -1, -1, -1, 1, 1, 1, 1)
    with pytest.raises(OSError):
        sim.add_mesh(str(tmp_path / "nope.stl"))


def test_write_vtk_layout(tmp_path):
    """Legacy ASCII POLYDATA with velocity, radius and id (src/vtk.rs:5-44)."""
    sim = rp.Simulation.create(1e-3, -1, -1, -1, 1, 1, 1, 2)
    sim.add_particle(0.5, 1.0, -0.25, 0.1, 1.0)
    sim.add_particle(0.0, 0.0, 0.0, 0.2, 1.0)
    sim.set_particle_velocity(0, 1.5, 0.0, 0.0)
    path = tmp_path / "out.vtk"
    sim.write_vtk(str(path))
    lines = path.read_text().splitlines()
    assert lines[:5] == ["# vtk DataFile Version 3.0", "DEM Simulation Data", "ASCII", "DATASET POLYDATA",
                         "POINTS 2 float"]
    assert lines[5] == "0.5 1 -0.25" and lines[6] == "0 0 0"
    assert lines[7:10] == ["POINT_DATA 2", "VECTORS velocity float", "1.5 0 0"]
    assert lines[11:15] == ["SCALARS radius float 1", "LOOKUP_TABLE default", "0.1", "0.2"]
    assert lines[15:19] == ["SCALARS id int 1", "LOOKUP_TABLE default", "0", "1"]


def test_ensemble_rejects_mismatched_members():
    from rusty_particles_amd import Ensemble, Simulation
    a = Simulation.create(1e-5, -0.1, -0.1, -0.1, 0.1, 0.1, 0.1, 0)
    b = Simulation.create(2e-5, -0.1, -0.1, -0.1, 0.1, 0.1, 0.1, 0)
    with pytest.raises(ValueError):
        Ensemble([a, b])
    c = Simulation.create(1e-5, -0.1, -0.1, -0.1, 0.1, 0.1, 0.1, 0)
    c.set_contact_models("jkr", "coulomb")
    with pytest.raises(ValueError):
        Ensemble([a, c])
    d = Simulation.create(1e-5, -0.1, -0.1, -0.1, 0.1, 0.1, 0.1, 0)
    d.set_particle_material(2e7, 0.3, 2500.0, 0.2, 0.5, 0.01)  # materials may differ
    Ensemble([a, d])
    with pytest.raises(ValueError):
        Ensemble([])


def test_jkr_solver_names():
    from rusty_particles_amd import Simulation
    s = Simulation.create(1e-5, -0.1, -0.1, -0.1, 0.1, 0.1, 0.1, 0)
    s.set_jkr_solver("shortcut")
    with pytest.raises(ValueError):
        s.set_jkr_solver("newton")
    with pytest.raises(ValueError):
        s.set_jkr_solver("surrogate")  # needs a trained network


def test_set_mesh_motion_checks_its_arguments():
    """Extension for rotating drums (the reference's walls are at rest)."""
    from rusty_particles_amd import scenes
    sim = rp.Simulation.create(1e-3, -1, -1, -1, 1, 1, 1, 0)
    with pytest.raises(IndexError):
        sim.set_mesh_motion(0, (0, 0, 0), (0, 0, 1))
    sim.add_mesh_arrays(*scenes.box_mesh((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5)))
    sim.set_mesh_motion(0, (0, 0, 0), (0, 0, 1))
    with pytest.raises(ValueError):
        sim.set_mesh_motion(0, (0, 0), (0, 0, 1))
    with pytest.raises(IndexError):
        sim.set_mesh_motion(1, (0, 0, 0), (0, 0, 1))
    lib = native.lib()
    o = (native.C.c_double * 3)()
    assert lib.rp_engine_set_mesh_motion(None, 0, o, o) == rp.abi.RP_ERR_INVALID


def test_reference_module_name():
    """pydem scripts import the reference's module name (src/lib.rs:24-35)."""
    import rusty_particles
    assert rusty_particles.Simulation is rp.Simulation
    assert rusty_particles.hello() == "Hello from Rust!"
    assert rusty_particles.set_num_threads is rp.set_num_threads
    sim = rusty_particles.Simulation.create(1e-3, -1, -1, -1, 1, 1, 1, 1)
    sim.add_particle(0.0, 0.5, 0.0, 0.01, 1.0)
    assert sim.get_particle_count() == 1


def test_set_mesh_motion_validates_before_any_engine():
    sim = rp.Simulation.create(1e-3, -1, -1, -1, 1, 1, 1, 0)
    from rusty_particles_amd import scenes
    for _ in range(6):
        sim.add_mesh_arrays(*scenes.box_mesh((-1, -1, -1), (1, 1, 1)))
    with pytest.raises(ValueError):
        sim.set_mesh_motion(0, (0, 0, 0), (0, float("nan"), 0))
    with pytest.raises(ValueError):
        sim.set_mesh_motion(0, (float("inf"), 0, 0), (0, 1, 0))
    for k in range(4):
        sim.set_mesh_motion(k, (0, 0, 0), (0, 1, 0))
    with pytest.raises(ValueError):
        sim.set_mesh_motion(4, (0, 0, 0), (0, 1, 0))
    sim.set_mesh_motion(4, (0, 0, 0), (0, 0, 0))  # at rest: allowed
    sim.set_mesh_motion(0, (0, 0, 0), (0, 0, 0))  # stopping one frees a slot
    sim.set_mesh_motion(5, (0, 0, 0), (1, 0, 0))
    with pytest.raises(IndexError):
        sim.set_mesh_motion(6, (0, 0, 0), (1, 0, 0))


def test_every_mutator_bumps_the_version():
    sim = rp.Simulation.create(1e-3, -1, -1, -1, 1, 1, 1, 0)
    from rusty_particles_amd import scenes
    calls = [
        lambda: sim.add_particle(0, 0, 0, 0.01, 1.0),
        lambda: sim.add_particles([[0.1, 0, 0]], 0.01, 1.0),
        lambda: sim.set_particle_velocity(0, 1, 2, 3),
        lambda: sim.set_particle_angular_velocity(0, 1, 2, 3),
        lambda: sim.set_particle_material(1e7, 0.3, 2500, 0.5, 0.5),
        lambda: sim.set_wall_material(1e9, 0.3, 7800, 0.5, 0.5),
        lambda: sim.set_periodic_boundary(True, False, False),
        lambda: sim.set_contact_models("hertz", "coulomb"),
        lambda: sim.set_gravity(0, -1, 0),
        lambda: sim.set_engine_tuning(0.0, 0.0),
        lambda: sim.set_jkr_solver("shortcut"),
        lambda: sim.add_mesh_arrays(*scenes.box_mesh((-1, -1, -1), (1, 1, 1))),
        lambda: sim.set_mesh_motion(0, (0, 0, 0), (0, 1, 0)),
        lambda: sim.set_physics_fix(True),
    ]
    for f in calls:
        v = sim._version
        f()
        assert sim._version > v


def test_overlap_warning_once_per_counted_step(capsys):
    class FakeEngine:
        def __init__(self):
            self.k = 0

        def stats(self):
            return {"overlap_steps": self.k}

    sim = rp.Simulation.create(1e-3, -1, -1, -1, 1, 1, 1, 0)
    eng = FakeEngine()
    eng.k = 3
    sim._warn_overlap(eng)
    eng.k = 4
    sim._warn_overlap(eng)
    sim._warn_overlap(eng)
    err = capsys.readouterr().err.strip().splitlines()
    assert err == [simulation.OVERLAP_WARNING] * 4
    assert err[0] == "Warning: Excessive particle overlap detected (>20% of radius). Simulation may be unstable."


def test_history_bookkeeping_without_an_engine():
    """The physics-fix history read with the host state: kept only in that
    mode with a spring law, and particles added later join without entries
    (their offsets repeat the last one)."""
    sim = rp.Simulation.create(1e-3, -1, -1, -1, 1, 1, 1, 0)
    sim.add_particles([[0, 0, 0], [0.01, 0, 0], [0.02, 0, 0]], 0.006, 1.0)
    sim.set_contact_models("hertz", "coulomb")
    assert not sim._keeps_history()
    sim.set_physics_fix(True)
    assert not sim._keeps_history()  # Coulomb friction has no history
    sim.set_contact_models("hertz", "mindlin")
    assert sim._keeps_history()
    assert sim._params.mode == abi.RP_MODE_PHYSICS_FIX
    sim._hist = (np.array([0, 1, 3, 4]), np.array([1, 0, 2, 1]), np.zeros((4, 3)))
    sim.add_particle(0.5, 0, 0, 0.006, 1.0)
    sim.add_particles([[0.6, 0, 0], [0.7, 0, 0]], 0.006, 1.0)
    offs, part, xi = sim._hist
    assert offs.tolist() == [0, 1, 3, 4, 4, 4, 4]
    assert part.tolist() == [1, 0, 2, 1] and xi.shape == (4, 3)
    sim.set_physics_fix(False)
    assert sim._params.mode == abi.RP_MODE_PARITY and not sim._keeps_history()
