"""The reference's pydem tests restated against the HIP engine through the
product's Simulation class (the drop-in for the pyo3 module)."""
import pytest

import pydem_cases as pc
import rusty_particles_amd as rp

pytestmark = pytest.mark.gpu


def create(*args):
    return rp.Simulation.create(*args)


def test_free_fall():
    pc.free_fall(create)


def test_hertzian_compression():
    pc.hertzian_compression(create)


def test_rolling_incline(tmp_path):
    pc.rolling_incline(create, tmp_path)


def test_jkr_pulloff():
    pc.jkr_pulloff(create)


def test_sjkr_pulloff():
    pc.sjkr_pulloff(create)


def test_contact_models():
    pc.contact_models(create)


def test_periodic_column():
    pc.periodic_column(create)


def test_overlapping_column_explodes_and_recovers():
    """1000 spheres stacked 1 mm apart with r = 5 cm fly apart in the first steps
    (faster than the Verlet skin): the engine re-runs the affected steps after
    rebuilding instead of failing."""
    sim = pc.overlapping_column(create)
    st = sim.stats()
    assert st["steps"] >= 100
    assert st["rebuilds"] > 1


def test_gpu_smoke():
    """pydem/test_gpu.py:13-36: enable_gpu + 100 steps on 1000 particles."""
    sim = rp.Simulation.create(0.01, -1.0, 0.0, -1.0, 1.0, 5.0, 1.0, 1000)
    for i in range(1000):
        sim.add_particle(0.0, 2.0 + i * 0.01, 0.0, 0.05, 1.0)
    sim.enable_gpu()
    for _ in range(100):
        sim.step()
    assert sim.get_particle_count() <= 1000
