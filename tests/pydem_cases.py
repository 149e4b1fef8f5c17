"""The reference's own Python tests (pydem/), restated as scenario functions.

Each function takes `create`, a factory with the signature of
Simulation.create (src/lib.rs:57-66), runs the scenario of the cited test
and asserts the same closed form with the same tolerance.  They run against
the CPU oracle (tests/test_oracle_pydem.py) and the HIP engine
(tests/test_gpu_pydem.py).  The reference's random inputs are seeded here.
"""
import os
import random
import struct

import numpy as np
import pytest


def free_fall(create):
    """pydem/test_analytical.py:13-48."""
    sim = create(1e-6, -10.0, 0.0, -10.0, 10.0, 20.0, 10.0, 10)
    sim.add_particle(0.0, 10.0, 0.0, 0.1, 1.0)
    sim.run(1.0)
    _, y, _ = sim.get_particle_position(0)
    assert abs(y - (10.0 - 0.5 * 9.81)) < 1e-2
    return y


def hertzian_compression(create):
    """pydem/test_analytical.py:51-112: static overlap (3mg/(4E*sqrt(R*)))^(2/3)."""
    radius, mass, E, nu, g = 0.1, 1.0, 1e7, 0.3, 9.81
    sim = create(1e-4, -1.0, -1.0, -1.0, 1.0, 2.0, 1.0, 10)
    sim.set_particle_material(E, nu, 2500.0, 0.5, 0.5)
    sim.add_particle(0.0, 0.0, 0.0, radius, 1e9, fixed=True)
    sim.add_particle(0.0, 2.0 * radius, 0.0, radius, mass)
    sim.run(1.0)
    _, y1, _ = sim.get_particle_position(0)
    _, y2, _ = sim.get_particle_position(1)
    overlap = 2.0 * radius - (y2 - y1)
    e_star = E / (2.0 * (1.0 - nu ** 2))
    expect = (3.0 * mass * g / (4.0 * e_star * np.sqrt(radius / 2.0))) ** (2.0 / 3.0)
    assert abs(overlap - expect) < 1e-5
    return overlap


def _write_floor(path):
    def tri(f, v1, v2, v3):
        n = np.cross(v2 - v1, v3 - v1)
        n = n / np.linalg.norm(n)
        f.write(struct.pack("<3f", *n))
        f.write(struct.pack("<3f", *v1))
        f.write(struct.pack("<3f", *v2))
        f.write(struct.pack("<3f", *v3))
        f.write(struct.pack("<H", 0))
    with open(path, "wb") as f:
        f.write(b"\x00" * 80)
        f.write(struct.pack("<I", 2))
        v1, v2, v3 = np.array([-10.0, 0.0, -10.0]), np.array([10.0, 0.0, -10.0]), np.array([-10.0, 0.0, 10.0])
        tri(f, v1, v3, v2)
        v1, v2, v3 = np.array([10.0, 0.0, -10.0]), np.array([10.0, 0.0, 10.0]), np.array([-10.0, 0.0, 10.0])
        tri(f, v1, v3, v2)


def rolling_incline(create, tmpdir):
    """pydem/test_analytical_rotation.py:13-173: time to pure rolling 2 v0 / (7 mu g)."""
    radius, mass, v0, mu, g, dt = 0.05, 1.0, 1.0, 0.1, 9.81, 1e-4
    expected = 2 * v0 / (7 * mu * g)
    sim = create(dt, -10.0, 0.0, -10.0, 10.0, 10.0, 10.0, 1)
    sim.set_particle_material(1e8, 0.3, 2500.0, mu, 0.5)
    sim.set_wall_material(1e8, 0.3, 2500.0, mu, 0.5)
    sim.set_contact_models("Hertzian", "Coulomb")
    path = os.path.join(str(tmpdir), "floor.stl")
    _write_floor(path)
    sim.add_mesh(path)
    sim.add_particle(0.0, radius, 0.0, radius, mass)
    sim.set_particle_velocity(0, v0, 0.0, 0.0)
    t = 0.0
    for _ in range(10000):
        sim.step()
        t += dt
        vx, _, _ = sim.get_particle_velocity(0)
        _, _, wz = sim.get_particle_angular_velocity(0)
        if vx + wz * radius <= 0.0:
            break
    else:
        pytest.fail("Did not reach saturation within max steps")
    assert abs(t - expected) / expected < 0.05
    return t


def jkr_pulloff(create):
    """pydem/test_jkr.py:11-90: minimum force -(2/3) pi gamma R* within 10%."""
    radius, mass, gamma, E, nu, dt = 0.005, 1.0, 0.1, 1e7, 0.3, 1e-5
    r_star = radius / 2.0
    expected = -(2.0 / 3.0) * np.pi * gamma * r_star
    e_star = 1.0 / (2.0 * (1.0 - nu ** 2) / E)
    a_eq = (9.0 * np.pi * gamma * r_star ** 2 / (2.0 * e_star)) ** (1.0 / 3.0)
    d_eq = a_eq ** 2 / r_star - np.sqrt(8.0 * np.pi * gamma * a_eq / e_star)
    sim = create(dt, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 2)
    sim.set_particle_material(E, nu, 2500.0, 0.5, 0.5, gamma)
    sim.set_contact_models("JKR", "Coulomb")
    sim.add_particle(0.0, 0.0, 0.0, radius, mass, True)
    sim.add_particle(2.0 * radius - d_eq, 0.0, 0.0, radius, mass, False)
    sim.set_particle_velocity(1, 0.002, 0.0, 0.0)
    fmin = 0.0
    for _ in range(50000):
        v0, _, _ = sim.get_particle_velocity(1)
        sim.step()
        v1, _, _ = sim.get_particle_velocity(1)
        fmin = min(fmin, mass * (v1 - v0) / dt)
        x, _, _ = sim.get_particle_position(1)
        if 2.0 * radius - x < -radius:
            break
    assert abs(fmin - expected) / abs(expected) < 0.1
    return fmin


def sjkr_pulloff(create):
    """pydem/test_sjkr.py:12-109: |minimum force| 1.5 pi gamma R* within 10%."""
    radius, mass, gamma, E, nu, dt = 0.005, 1.0, 0.5, 1e5, 0.3, 1e-5
    r_star = radius / 2.0
    expected = 1.5 * np.pi * gamma * r_star
    e_star = 1.0 / (2.0 * (1.0 - nu ** 2) / E)
    delta_c = -np.sqrt(3.0 * np.pi ** 2 * gamma ** 2 * r_star / e_star ** 2)
    sim = create(dt, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 2)
    sim.set_particle_material(E, nu, 2500.0, 0.5, 0.5, gamma)
    sim.set_contact_models("sJKR", "Coulomb")
    sim.add_particle(0.0, 0.0, 0.0, radius, mass, True)
    sim.add_particle(2.0 * radius - radius * 0.01, 0.0, 0.0, radius, mass, False)
    sim.step()
    sim.set_particle_velocity(1, 0.01, 0.0, 0.0)
    fmin = 0.0
    for _ in range(50000):
        v0, _, _ = sim.get_particle_velocity(1)
        sim.step()
        v1, _, _ = sim.get_particle_velocity(1)
        fmin = min(fmin, mass * (v1 - v0) / dt)
        x, _, _ = sim.get_particle_position(1)
        if 2.0 * radius - x < delta_c * 1.1:
            break
    assert abs(abs(fmin) - expected) / expected < 0.1
    return fmin


def contact_models(create):
    """pydem/test_models.py:12-33."""
    sim = create(0.001, -1.0, 0.0, -1.0, 1.0, 5.0, 1.0, 100)
    sim.add_particle(0.0, 1.0, 0.0, 0.05, 1.0)
    sim.run(0.1)
    sim.set_contact_models("linear", "coulomb")
    sim.run(0.1)
    with pytest.raises(ValueError):
        sim.set_contact_models("invalid", "coulomb")
    with pytest.raises(ValueError):
        sim.set_contact_models("hertz", "invalid")


def periodic_column(create):
    """pydem/test_periodic.py:12-47 (seeded): 101 particles, periodic Y, 100 steps."""
    rnd = random.Random(12345)
    sim = create(0.01, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1)
    sim.set_periodic_boundary(False, True, False)
    sim.add_particle(0.9, 0.0, 0.0, 0.1, 1.0)
    for _ in range(100):
        sim.add_particle(rnd.uniform(-0.4, 0.4), rnd.uniform(1.0, 4.0), rnd.uniform(-0.4, 0.4),
                         rnd.uniform(0.02, 0.03), 1.0)
    for _ in range(100):
        sim.step()
    n = sim.get_particle_count()
    for i in range(n):
        _, y, _ = sim.get_particle_position(i)
        assert -1.0 - 1.0 <= y <= 1.0 + 1.0
    return sim


def overlapping_column(create):
    """pydem/test_multithreading.py:12-34: 1000 heavily overlapping spheres, 0.1 s."""
    sim = create(0.001, -1.0, 0.0, -1.0, 1.0, 5.0, 1.0, 5000)
    for i in range(1000):
        sim.add_particle(0.0, 2.0 + i * 0.001, 0.0, 0.05, 1.0)
    sim.run(0.1)
    return sim
