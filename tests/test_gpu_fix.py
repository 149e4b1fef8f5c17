"""The physics-fix mode (RP_MODE_PHYSICS_FIX) on the GPU against the oracle.

* Dense scenes with walls, several model pairs: pair lists, states and the
  tangential history tables (partner, displacement) equal to the bit after
  many steps, with a thin skin so the lists are rebuilt -- and the particles
  re-sorted -- several times (the tables move with their owners).
* The closed form of the oracle's tests: a sphere resting on a fixed 10 m
  sphere, pushed tangentially, oscillates with period 2 pi sqrt(m / (3.5 kt)).
* Wall contacts through Eberly's closest point in every branch.
* The capacity run as the benchmark specifies it: JKR (surrogate with exact
  fallback) + Mindlin with history in the turning drum, at oracle size.
* What the mode refuses: batched ensembles, the autotuner with a history.
"""
import math

import numpy as np
import pytest

from closest_point_cases import CASES
from helpers import assert_forces_close, copy_params, dense_scene, params
from oracle import OracleSimulation
from rusty_particles_amd import abi, native, scenes, surrogate

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "w", "a", "tau", "rx", "rv", "rw")


def _pair(p, sc, meshes=(), motions=(), solver=None, net=None, record=True):
    e = native.Engine(copy_params(p))
    e.upload(sc["x"], sc["radius"], sc["mass"], sc.get("fixed"), v=sc.get("v"), w=sc.get("w"))
    for v9, n3 in meshes:
        e.add_mesh(v9, n3)
    for k, org, om in motions:
        e.set_mesh_motion(k, org, om)
    if solver:
        e.set_jkr_solver(solver, net)
    e.set_record(record)
    o = OracleSimulation(p.dt, None, None, params=copy_params(p))
    o.set_state(sc["x"], sc["radius"], sc["mass"], sc.get("fixed"), v=sc.get("v"), w=sc.get("w"))
    for v9, n3 in meshes:
        o.add_mesh_arrays(v9, n3)
    for k, org, om in motions:
        o.set_mesh_motion(k, org, om)
    o.set_record(record)
    return e, o


def _equal(o, e, history=True):
    so, se = o.state(), e.download()
    for k in FIELDS:
        assert np.array_equal(so[k], se[k]), k
    if history:
        oo, op_, ox = o.history()
        eo, ep_, ex = e.history()
        assert np.array_equal(oo, eo) and np.array_equal(op_, ep_), "history partners differ"
        assert np.array_equal(ox, ex), "history displacements differ"
        assert e.stats()["history_contacts"] == op_.size
        return op_.size
    return 0


@pytest.mark.parametrize("normal,tangential,solver,gamma", [
    (abi.RP_NORMAL_HERTZ, abi.RP_TANGENTIAL_MINDLIN, None, 0.0),
    (abi.RP_NORMAL_HERTZ, abi.RP_TANGENTIAL_LSC, None, 0.0),
    (abi.RP_NORMAL_LINEAR, abi.RP_TANGENTIAL_MINDLIN, None, 0.0),
    (abi.RP_NORMAL_JKR, abi.RP_TANGENTIAL_MINDLIN, "exact", 0.05),
    (abi.RP_NORMAL_JKR, abi.RP_TANGENTIAL_MINDLIN, "shortcut", 0.05),
    (abi.RP_NORMAL_SJKR, abi.RP_TANGENTIAL_LSC, None, 0.05),
    (abi.RP_NORMAL_HERTZ, abi.RP_TANGENTIAL_COULOMB, None, 0.0),
], ids=["hertz-mindlin", "hertz-lsc", "linear-mindlin", "jkr-mindlin-exact", "jkr-mindlin-shortcut",
        "sjkr-lsc", "hertz-coulomb"])
def test_dense_scene_matches_oracle_with_history(normal, tangential, solver, gamma):
    sc = dense_scene(counts=(9, 9, 9), spacing=0.0095, seed=23, vel=0.3, omega=30.0)
    p = params(2e-6, (-0.06, -0.06, -0.06), (0.06, 0.06, 0.06), normal=normal, tangential=tangential,
               gamma=gamma)
    p.mode = abi.RP_MODE_PHYSICS_FIX
    p.skin = 2e-5  # thin: the lists are rebuilt (and the particles re-sorted) several times
    meshes = [scenes.box_mesh((-0.045, -0.045, -0.045), (0.045, 0.045, 0.045))]
    print(f"case fix normal={normal} tangential={tangential} solver={solver}", flush=True)
    e, o = _pair(p, sc, meshes, solver=solver)
    for chunk in (1, 60, 139):
        o.step(chunk)
        e.step(chunk)
        oo, oj = o.pairs()
        eo, ej = e.pairs()
        assert np.array_equal(oo, eo) and np.array_equal(oj, ej), chunk
        nh = _equal(o, e, history=tangential != abi.RP_TANGENTIAL_COULOMB)
    if tangential != abi.RP_TANGENTIAL_COULOMB:
        assert nh > 100
    assert e.stats()["rebuilds"] >= 3
    assert np.isfinite(e.download()["v"]).all()


def test_tangential_oscillation_period():
    R_A, r, E, nu, rho, dt = 10.0, 5e-3, 1e7, 0.3, 2500.0, 1e-6
    p = params(dt, (-20.0, -20.0, -20.0), (20.0, 20.0, 20.0), normal=abi.RP_NORMAL_HERTZ,
               tangential=abi.RP_TANGENTIAL_MINDLIN, E=E, nu=nu, mu=0.5)
    p.mode = abi.RP_MODE_PHYSICS_FIX
    m = rho * 4 / 3 * math.pi * r ** 3
    estar = 1 / (2 * (1 - nu * nu) / E)
    rstar = R_A * r / (R_A + r)
    d0 = (m * 9.81 / (4 / 3 * estar * math.sqrt(rstar))) ** (2 / 3)
    gstar = 1 / (2 * (2 - nu) / (E / (2 * (1 + nu))))
    kt = 8 * gstar * math.sqrt(rstar * d0)
    period = 2 * math.pi * math.sqrt(m / (3.5 * kt))
    sc = dict(x=np.array([[0.0, -R_A, 0.0], [0.0, r - d0, 0.0]]), radius=np.array([R_A, r]),
              mass=np.array([1e12, m]), v=np.array([[0.0, 0.0, 0.0], [2e-3, 0.0, 0.0]]), w=np.zeros((2, 3)),
              fixed=np.array([1, 0], np.uint8))
    e, o = _pair(p, sc, record=False)
    nsteps = int(2.2 * period / dt)
    vc = []
    for _ in range(nsteps // 50):
        e.step(50)
        st = e.download()
        vc.append(st["v"][1, 0] + st["w"][1, 2] * r)
    o.step(50 * (nsteps // 50))
    _equal(o, e)
    vc = np.array(vc)
    cross = np.nonzero((vc[:-1] > 0) & (vc[1:] <= 0))[0]
    assert cross.size >= 2
    t = [(c + vc[c] / (vc[c] - vc[c + 1])) * 50 * dt for c in cross[:2]]
    assert t[1] - t[0] == pytest.approx(period, rel=5e-3)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_wall_contact_eberly(case):
    name, tri, pt, want, true = case
    q = np.asarray(true if true is not None else want)
    dist = float(np.linalg.norm(np.subtract(pt, q)))
    p = params(1e-6, (-40.0, -40.0, -40.0), (40.0, 40.0, 40.0), normal=abi.RP_NORMAL_HERTZ,
               tangential=abi.RP_TANGENTIAL_MINDLIN, E=1e7)
    p.mode = abi.RP_MODE_PHYSICS_FIX
    rr = np.array([1.01 * dist])
    sc = dict(x=np.array([pt], float), radius=rr, mass=scenes.sphere_mass(rr, 2500.0),
              v=np.array([[0.01, -0.02, 0.03]]), w=np.array([[0.5, 0.0, -0.5]]), fixed=np.zeros(1, np.uint8))
    meshes = [(np.array([np.concatenate(tri)], np.float32), np.array([[0.0, 0.0, 1.0]], np.float32))]
    e, o = _pair(p, sc, meshes)
    o.step(1)
    e.step(1)
    assert_forces_close(o, e, p)
    assert e.stats()["wall_contacts"] == 1


def _drum_env(p, r):
    rr = r[: min(r.size, 4000)]
    rs = rr[:, None] * rr[None, :200] / (rr[:, None] + rr[None, :200])
    e_star = 1.0 / (2 * (1 - p.particle.poissons_ratio ** 2) / p.particle.youngs_modulus)
    return surrogate.envelope(-0.40, 30.0, (0.9 * rs.min(), 1.1 * rs.max()), e_star, p.particle.surface_energy)


@pytest.fixture(scope="module")
def drum_fix():
    p, x, r, m, meshes = scenes.drum_test_workload(20_000)
    p.tangential_model = abi.RP_TANGENTIAL_MINDLIN
    p.mode = abi.RP_MODE_PHYSICS_FIX
    net = surrogate.train(_drum_env(p, r), nsamples=400_000, epochs=30, batch=2048, lr=3e-3, seed=6)
    return p, dict(x=x, radius=r, mass=m), meshes, net


@pytest.mark.parametrize("solver", ["shortcut", "surrogate"])
def test_capacity_drum_jkr_mindlin(drum_fix, solver):
    """The capacity configuration as specified -- JKR (gamma 0.01) with the
    surrogate and exact fallback, Mindlin tangential with history, turning
    drum -- on 20,000 spheres; shortcut: bit-exact with the oracle, history
    included; surrogate: forces within 1e-6 of the per-particle force scale
    at the first step, served contacts within the 1e-3 check bound, states
    within 1e-6 after 200 steps."""
    p, sc, meshes, net = drum_fix
    motions = [(0, scenes.DRUM_ORIGIN, scenes.DRUM_OMEGA)]
    print(f"case fix drum solver={solver}", flush=True)
    e, o = _pair(p, sc, meshes, motions, solver=solver, net=net, record=False)
    if solver == "shortcut":
        o.step(200)
        e.step(200)
        # the chains start at rest: only contacts that slid carry a displacement
        assert _equal(o, e) > 100
        assert np.isfinite(e.download()["v"]).all()
        return
    e.set_jkr_solver("surrogate", net, check=True)
    x0 = sc["x"]
    o.step(1)
    e.step(1)
    assert_forces_close(o, e, p, rtol=1e-6)
    js = e.jkr_stats()
    assert js["served"] > 0 and js["err_max"] <= 1e-3, js
    o.step(199)
    e.step(199)
    so, se = o.state(), e.download()
    disp = np.abs(so["x"] - x0).max()
    assert np.abs(so["x"] - se["x"]).max() <= 1e-6 * disp
    assert np.abs(so["v"] - se["v"]).max() <= 1e-6 * np.abs(so["v"]).max()
    assert e.stats()["history_contacts"] > 100


def test_refusals():
    p = params(1e-5, (-0.05,) * 3, (0.05,) * 3, tangential=abi.RP_TANGENTIAL_MINDLIN)
    p.mode = abi.RP_MODE_PHYSICS_FIX
    sc = dense_scene(counts=(4, 4, 4), seed=2)
    e = native.Engine(copy_params(p))
    with pytest.raises(native.EngineError):
        e.upload_batch([32, 32], [p.particle, p.particle], sc["x"], sc["radius"], sc["mass"])
    e.upload(sc["x"], sc["radius"], sc["mass"])
    with pytest.raises(native.EngineError):
        e.autotune(5)
    e.step(3)  # still usable
    assert np.isfinite(e.download()["x"]).all()
