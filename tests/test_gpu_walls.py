"""Moving walls on the GPU against the oracle: a rotating drum.

The reference's walls are at rest (simulation.rs:475-479); the engine adds
a mesh surface velocity (rp_engine_set_mesh_motion), the rigid rotation
omega x (c - origin) at the contact point, with the triangles left in place,
which is exact for a drum turning about its own axis.  The oracle carries
the same extension.  Checked here, 500 steps of spheres thrown against the
wall of a turning drum: pair lists bit-exact and states equal to the bit
(Hertz and sJKR, whose unloading branch reads the relative velocity); the
turning wall drags the spheres round (their angular momentum about the axis
follows the sign of omega), and omega = 0 equals a static drum bit for bit.
"""
import numpy as np
import pytest

from rusty_particles_amd import abi, scenes

from helpers import assert_pairs_equal, assert_state_close, params, run_engine, run_oracle

pytestmark = pytest.mark.gpu

RD, LD = 0.02, 0.012


def _drum_scene(seed=11):
    rng = np.random.default_rng(seed)
    s = 2.3e-3
    pts = []
    for z in (-s, 0.0, s):
        for y in np.arange(-RD + 1.6e-3, -0.002, s):
            for x in np.arange(-RD, RD, s):
                if x * x + y * y <= (RD - 1.6e-3) ** 2:
                    pts.append((x, y, z))
    x = np.array(pts) + rng.uniform(-5e-5, 5e-5, (len(pts), 3))
    n = x.shape[0]
    r = rng.uniform(0.9e-3, 1.1e-3, n)
    v = rng.normal(0.0, 0.05, (n, 3))
    v[:, 1] -= 0.2  # thrown at the lower wall
    return dict(x=x, radius=r, mass=scenes.sphere_mass(r, 2500.0), v=v, w=np.zeros((n, 3)),
                fixed=np.zeros(n, np.uint8))


def _setup(normal, gamma):
    p = params(1e-5, (-0.025, -0.025, -0.01), (0.025, 0.025, 0.01), normal=normal,
               tangential=abi.RP_TANGENTIAL_COULOMB, E=1e7, gamma=gamma)
    return p, _drum_scene(), [scenes.drum_mesh(RD, LD, nseg=24, nlen=2, ncap=2)]


def _lz(st):
    return float(np.sum(st["mass"] * (st["x"][:, 0] * st["v"][:, 1] - st["x"][:, 1] * st["v"][:, 0])))


@pytest.mark.parametrize("normal,gamma", [(abi.RP_NORMAL_HERTZ, 0.0), (abi.RP_NORMAL_SJKR, 0.01)],
                         ids=["hertz", "sjkr"])
def test_rotating_drum_matches_oracle(normal, gamma):
    p, sc, meshes = _setup(normal, gamma)
    print(f"case normal={normal} n={sc['x'].shape[0]}", flush=True)
    omega = (0.0, 0.0, 10.0)
    o = run_oracle(p, sc, meshes, steps=0)
    o.set_mesh_motion(0, (0.0, 0.0, 0.0), omega)
    o.step(500)
    e = run_engine(p, sc, meshes, steps=0)
    e.set_mesh_motion(0, (0.0, 0.0, 0.0), omega)
    e.set_record(True)
    e.step(500)
    assert e.stats()["wall_contacts"] > 0
    assert assert_pairs_equal(o, e) > 0
    assert_state_close(o, e, 0.0, 0.0, 0.0)


def test_turning_wall_drags_and_zero_omega_is_static():
    p, sc, meshes = _setup(abi.RP_NORMAL_HERTZ, 0.0)
    runs = {}
    for w in (None, 0.0, 10.0, -10.0):
        e = run_engine(p, sc, meshes, steps=0)
        if w is not None:
            e.set_mesh_motion(0, (0.0, 0.0, 0.0), (0.0, 0.0, w))
        e.step(500)
        runs[w] = e.download()
    for k in ("x", "v", "w"):
        assert np.array_equal(runs[None][k], runs[0.0][k])
    l0 = _lz(runs[None])
    assert _lz(runs[10.0]) > l0 and _lz(runs[-10.0]) < l0



def test_simulation_api_applies_mesh_motion():
    """Simulation.set_mesh_motion before the engine exists (it is created at
    the first step) gives, bit for bit, the run of the engine-level call."""
    import rusty_particles_amd as rp
    p, sc, meshes = _setup(abi.RP_NORMAL_HERTZ, 0.0)
    omega = (0.0, 0.0, 10.0)
    sim = rp.Simulation(p.dt, tuple(p.bounds_min), tuple(p.bounds_max))
    sim.set_contact_models("hertz", "coulomb")
    m = p.particle
    sim.set_particle_material(m.youngs_modulus, m.poissons_ratio, m.density, m.friction_coefficient,
                              m.restitution_coefficient, m.surface_energy)
    sim.add_particles(sc["x"], sc["radius"], sc["mass"])
    for k in range(sc["x"].shape[0]):
        sim.set_particle_velocity(k, *sc["v"][k])
    sim.add_mesh_arrays(*meshes[0])
    sim.set_mesh_motion(0, (0.0, 0.0, 0.0), omega)
    for _ in range(200):
        sim.step()
    got = sim.state()
    e = run_engine(sim._params, sc, meshes, steps=0)
    e.set_mesh_motion(0, (0.0, 0.0, 0.0), omega)
    e.step(200)
    ref = e.download()
    for k in ("x", "v", "w"):
        assert np.array_equal(got[k], ref[k]), k


def test_mesh_motion_limits_and_errors():
    """At most 4 meshes move at once; a fifth, an unknown mesh or a non-finite
    value is refused with the previous motions kept, and stepping goes on."""
    from rusty_particles_amd import native
    p, sc, meshes = _setup(abi.RP_NORMAL_HERTZ, 0.0)
    e = run_engine(p, sc, meshes * 5, steps=0)
    for k in range(4):
        e.set_mesh_motion(k, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0 + k))
    with pytest.raises(native.EngineError):
        e.set_mesh_motion(4, (0.0, 0.0, 0.0), (0.0, 0.0, 5.0))
    with pytest.raises(native.EngineError):
        e.set_mesh_motion(5, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0))
    with pytest.raises(native.EngineError):
        e.set_mesh_motion(0, (0.0, 0.0, 0.0), (0.0, float("nan"), 1.0))
    e.set_mesh_motion(3, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0))  # frees a slot
    e.set_mesh_motion(4, (0.0, 0.0, 0.0), (0.0, 0.0, 5.0))
    e.step(20)
    st = e.download()
    assert np.all(np.isfinite(st["x"]))
