"""Batched ensemble: several independent instances in one engine (one set of
launches per step) must each evolve exactly as a single engine given that
instance alone -- bitwise, since the per-instance arithmetic and summation
order are the same."""
import numpy as np
import pytest

from helpers import copy_params, dense_scene, params
from rusty_particles_amd import abi, native, scenes

pytestmark = pytest.mark.gpu


def _instances(k, seed):
    rng = np.random.default_rng(seed)
    inst = []
    for q in range(k):
        c = tuple(int(v) for v in rng.integers(3, 9, 3))
        sc = dense_scene(counts=c, seed=seed * 100 + q, centre=tuple(rng.uniform(-0.02, 0.02, 3)))
        mat = abi.RpMaterial(1e7, 0.3, 2500.0, float(rng.uniform(0.1, 0.8)), float(rng.uniform(0.2, 0.9)),
                             float(rng.uniform(0.0, 0.02)))
        inst.append((sc, mat))
    return inst


def _single(p, sc, mat, meshes, steps, solver=None, motion=None):
    q = copy_params(p)
    q.particle = mat
    e = native.Engine(q)
    e.upload(sc["x"], sc["radius"], sc["mass"], sc["fixed"], v=sc["v"], w=sc["w"])
    for v9, n3 in meshes:
        e.add_mesh(v9, n3)
    if motion:
        e.set_mesh_motion(0, *motion)
    if solver:
        e.set_jkr_solver(solver)
    e.step(steps)
    return e.download()


def _batched(p, inst, meshes, steps, solver=None, chunks=None, motion=None):
    e = native.Engine(copy_params(p))
    cat = lambda key: np.concatenate([sc[key] for sc, _ in inst])
    e.upload_batch([sc["x"].shape[0] for sc, _ in inst], [m for _, m in inst], cat("x"), cat("radius"),
                   cat("mass"), cat("fixed"), v=cat("v"), w=cat("w"))
    for v9, n3 in meshes:
        e.add_mesh(v9, n3)
    if motion:
        e.set_mesh_motion(0, *motion)
    if solver:
        e.set_jkr_solver(solver)
    for c in (chunks or [steps]):
        e.step(c)
    return e, e.download()


@pytest.mark.parametrize("normal,tangential,solver", [
    (abi.RP_NORMAL_SJKR, abi.RP_TANGENTIAL_COULOMB, None),
    (abi.RP_NORMAL_HERTZ, abi.RP_TANGENTIAL_COULOMB, None),
    (abi.RP_NORMAL_JKR, abi.RP_TANGENTIAL_COULOMB, "shortcut"),
    (abi.RP_NORMAL_LINEAR, abi.RP_TANGENTIAL_MINDLIN, None),
])
def test_instances_match_single_runs(normal, tangential, solver):
    inst = _instances(7, seed=normal * 10 + tangential)
    p = params(1e-5, (-0.06, -0.06, -0.06), (0.06, 0.06, 0.06), normal=normal, tangential=tangential)
    v9, n3 = scenes.box_mesh((-0.05, -0.05, -0.05), (0.05, 0.05, 0.05))
    meshes = [(v9, n3)]
    eng, got = _batched(p, inst, meshes, 60, solver, chunks=[25, 35])
    counts = eng.instance_counts()
    off = 0
    for q, (sc, mat) in enumerate(inst):
        ref = _single(p, sc, mat, meshes, 60, solver)
        n = ref["x"].shape[0]
        assert counts[q] == n
        for key in ("x", "v", "w", "a", "tau", "rx", "rv", "rw", "radius", "mass"):
            assert np.array_equal(got[key][off:off + n], ref[key]), (q, key)
        off += n
    assert off == got["x"].shape[0]


def test_instances_match_single_runs_with_a_moving_wall():
    """The box's surface moves (the moving-wall extension); wall contacts
    occur in several instances, and each instance still evolves as it does
    alone, bit for bit."""
    inst = _instances(7, seed=77)
    p = params(1e-5, (-0.06, -0.06, -0.06), (0.06, 0.06, 0.06), normal=abi.RP_NORMAL_SJKR,
               tangential=abi.RP_TANGENTIAL_COULOMB)
    meshes = [scenes.box_mesh((-0.05, -0.05, -0.05), (0.05, 0.05, 0.05))]
    motion = ((0.01, 0.0, -0.02), (0.3, 1.0, -0.5))
    eng, got = _batched(p, inst, meshes, 60, chunks=[25, 35], motion=motion)
    assert eng.stats()["wall_contacts"] > 0
    off = 0
    for q, (sc, mat) in enumerate(inst):
        ref = _single(p, sc, mat, meshes, 60, motion=motion)
        n = ref["x"].shape[0]
        for key in ("x", "v", "w", "a", "tau", "rx", "rv", "rw"):
            assert np.array_equal(got[key][off:off + n], ref[key]), (q, key)
        off += n
    assert off == got["x"].shape[0]


def test_removal_counts_per_instance():
    """Particles that leave the bounds are removed per instance; the others
    keep stepping and the per-instance counts follow."""
    inst = _instances(4, seed=3)
    for q, (sc, _) in enumerate(inst):
        if q % 2 == 0:
            # removal happens 1 m below bounds_min.y (simulation.rs:200-214): start just above that line
            sc["x"][:5, 1] = -0.08 - 0.9995
            sc["v"][:5] = [0.0, -10.0, 0.0]
    p = params(1e-4, (-0.08, -0.08, -0.08), (0.08, 0.08, 0.08))
    eng, got = _batched(p, inst, [], 30)
    counts = eng.instance_counts()
    off = 0
    for q, (sc, mat) in enumerate(inst):
        ref = _single(p, sc, mat, [], 30)
        n = ref["x"].shape[0]
        assert counts[q] == n and (n < sc["x"].shape[0]) == (q % 2 == 0)
        assert np.array_equal(got["x"][off:off + n], ref["x"]), q
        off += n


def test_batched_rejects_record():
    inst = _instances(2, seed=5)
    p = params(1e-5, (-0.06,) * 3, (0.06,) * 3, normal=abi.RP_NORMAL_JKR)
    e, _ = _batched(p, inst, [], 1)
    with pytest.raises(native.EngineError):
        e.set_record(True)


def test_batched_surrogate_matches_single_runs():
    """JKR dimer instances with different surface energies (one outside the
    envelope) in one batched engine with the surrogate: each instance ends
    exactly where a single engine with the surrogate puts it."""
    from helpers import jkr_dimers
    from rusty_particles_amd import surrogate
    e_star = 1.0 / (2 * (1 - 0.3 ** 2) / 1e8)
    env = surrogate.envelope(-0.40, 30.0, (5e-6, 1.2e-4), e_star, (0.02, 0.08))
    net = surrogate.train(env, nsamples=300_000, epochs=30, batch=2048, lr=3e-3, seed=5)
    rng = np.random.default_rng(8)
    gammas = [0.03, 0.05, 0.075, 0.2]   # the last lies outside the envelope: exact solve
    inst = []
    for q, g in enumerate(gammas):
        sc = jkr_dimers(rng.uniform(-0.5, 20.0, 500 + 300 * q), vel=0.2, seed=20 + q, gamma=g)
        sc["fixed"] = np.zeros(sc["x"].shape[0], np.uint8)
        inst.append((sc, abi.RpMaterial(1e8, 0.3, 1500.0, 0.3, 0.5, g)))
    p = params(1e-8, (-0.01,) * 3, (0.01,) * 3, normal=abi.RP_NORMAL_JKR, E=1e8, nu=0.3, mu=0.3, gamma=0.05)
    e = native.Engine(copy_params(p))
    cat = lambda key: np.concatenate([sc[key] for sc, _ in inst])
    e.upload_batch([sc["x"].shape[0] for sc, _ in inst], [m for _, m in inst], cat("x"), cat("radius"),
                   cat("mass"), cat("fixed"), v=cat("v"), w=cat("w"))
    e.set_jkr_solver("surrogate", net)
    e.step(40)
    got = e.download()
    assert e.jkr_stats()["served"] > 0
    off = 0
    for sc, mat in inst:
        q = copy_params(p)
        q.particle = mat
        s1 = native.Engine(q)
        s1.upload(sc["x"], sc["radius"], sc["mass"], sc["fixed"], v=sc["v"], w=sc["w"])
        s1.set_jkr_solver("surrogate", net)
        s1.step(40)
        ref = s1.download()
        n = ref["x"].shape[0]
        for key in ("x", "v", "w", "a", "tau"):
            assert np.array_equal(got[key][off:off + n], ref[key]), key
        off += n


def test_ensemble_of_simulations_matches_members_alone():
    """Simulation objects stepped together by an Ensemble (one batched
    engine) end where each would alone -- including a member that had
    already been stepped (its accelerations and Kahan residuals carry over)."""
    from rusty_particles_amd import Ensemble, Simulation

    def make(seed, mu, gamma, counts):
        sc = dense_scene(counts=counts, seed=seed)
        s = Simulation.create(1e-5, -0.06, -0.06, -0.06, 0.06, 0.06, 0.06, 0)
        s.set_contact_models("sjkr", "coulomb")
        s.set_particle_material(1e7, 0.3, 2500.0, mu, 0.5, gamma)
        s.add_particles(sc["x"], sc["radius"], sc["mass"])
        for i in range(sc["x"].shape[0]):
            s.set_particle_velocity(i, *sc["v"][i])
            s.set_particle_angular_velocity(i, *sc["w"][i])
        return s

    spec = [(1, 0.3, 0.01, (5, 6, 7)), (2, 0.6, 0.0, (8, 4, 5)), (3, 0.1, 0.02, (3, 3, 9))]
    alone = [make(*a) for a in spec]
    together = [make(*a) for a in spec]
    alone[1].run(2e-4)        # 20 steps before the ensemble forms
    together[1].run(2e-4)
    for s in alone:
        s.run(5e-4)
    ens = Ensemble(together)
    ens.run(5e-4)
    for a, b in zip(alone, together):
        sa, sb = a.state(), b.state()
        for k in ("x", "v", "w", "a", "tau", "rx", "rv", "rw"):
            assert np.array_equal(sa[k], sb[k]), k


def test_ensemble_reuploads_members_changed_between_runs():
    """A member edited between two Ensemble runs (velocity set, a particle
    added, material changed, or stepped on its own) is stepped from its new
    state: the ensemble ends where the members end when each is stepped alone
    with the same edits."""
    from rusty_particles_amd import Ensemble, Simulation

    def make(seed, mu, gamma, counts):
        sc = dense_scene(counts=counts, seed=seed)
        s = Simulation.create(1e-5, -0.06, -0.06, -0.06, 0.06, 0.06, 0.06, 0)
        s.set_contact_models("sjkr", "coulomb")
        s.set_particle_material(1e7, 0.3, 2500.0, mu, 0.5, gamma)
        s.add_particles(sc["x"], sc["radius"], sc["mass"])
        return s

    spec = [(4, 0.3, 0.01, (4, 5, 6)), (5, 0.6, 0.0, (6, 4, 4)), (6, 0.2, 0.02, (3, 4, 7))]
    alone = [make(*a) for a in spec]
    together = [make(*a) for a in spec]
    for s in alone:
        s.run(2e-4)
    ens = Ensemble(together)
    ens.run(2e-4)

    def edit(group):
        group[0].set_particle_velocity(3, 0.1, -0.2, 0.05)
        group[1].add_particle(0.0, 0.04, 0.0, 0.005, scenes.sphere_mass(np.array([0.005]), 2500.0)[0])
        group[2].set_particle_material(1e7, 0.3, 2500.0, 0.7, 0.5, 0.015)
        group[2].run(1e-4)  # 10 steps on its own engine

    edit(alone)
    edit(together)
    for s in alone:
        s.run(3e-4)
    ens.run(3e-4)
    for a, b in zip(alone, together):
        sa, sb = a.state(), b.state()
        for k in ("x", "v", "w", "a", "tau", "rx", "rv", "rw"):
            assert np.array_equal(sa[k], sb[k]), k
