"""The benchmark's five configurations on the GPU against the oracle, each
with its own generator, materials, geometry and code paths, at sizes the
oracle follows (its fast enumeration, pinned to the literal walk by
tests/test_oracle_fast.py).

Bars: pair lists bit-exact; cell ids and the sort permutation bit-exact
(assert_cells_equal); where the whole domain lies in one reference cell the
engine sums contacts in the reference's order, so states are compared with
array_equal; across several reference cells forces agree within 1e-10 of
the per-particle sum of contact-force magnitudes (assert_forces_close) and
states to the rounding that order change leaves; the JKR surrogate's forces
within its check-mode bound (1e-3 relative; measured ~1e-9).
"""
import numpy as np
import pytest

from helpers import assert_cells_equal, assert_forces_close, copy_params
from oracle import OracleSimulation
from rusty_particles_amd import abi, native, scenes, surrogate

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "w", "a", "tau", "rx", "rv", "rw")


def _oracle(p, x, r, m, v=None, meshes=(), motions=(), steps=0):
    o = OracleSimulation(p.dt, None, None, params=copy_params(p))
    o.set_state(x, r, m, v=v)
    for v9, n3 in meshes:
        o.add_mesh_arrays(v9, n3)
    for k, org, om in motions:
        o.set_mesh_motion(k, org, om)
    o.set_fast(True)
    o.set_record(True)
    if steps:
        o.step(steps)
    return o


def _engine(p, x, r, m, v=None, meshes=(), motions=(), solver=None, net=None):
    e = native.Engine(copy_params(p))
    e.upload(x, r, m, v=v)
    for v9, n3 in meshes:
        e.add_mesh(v9, n3)
    for k, org, om in motions:
        e.set_mesh_motion(k, org, om)
    if solver and solver != "exact":
        e.set_jkr_solver(solver, net)
    e.set_record(True)
    return e


def _pairs_equal(o, e):
    oo, oj = o.pairs()
    eo, ej = e.pairs()
    assert np.array_equal(oo, eo) and np.array_equal(oj, ej), "pair lists differ"
    return oj.size


def _states_equal(o, e):
    so, se = o.state(), e.download()
    for k in FIELDS:
        assert np.array_equal(so[k], se[k]), k


def _counters_equal(o, e, fails_before):
    st = e.stats()
    assert st["excessive_overlap"] == o.overlap_flag()
    assert st["jkr_failures"] == o.jkr_failures() - fails_before
    return st


# ---- 1: the reference's CPU-runnable configuration, all 10,000 sites ----------

@pytest.mark.parametrize("stirred", [False, True], ids=["as-specified", "stirred"])
def test_config1_box_settling_all_sites(stirred):
    """10,000 spheres (r 5 mm) on the jittered 11 mm lattice in the STL box,
    Hertz + Coulomb, dt 1e-5 s, all 2,000 steps: bit-identical.  As specified
    (at rest) the spheres fall ~2 mm in those 0.02 s and meet neither each
    other nor the floor; the stirred variant starts them with velocities
    ~N(0, 1.5) m/s (and spins ~N(0, 20) rad/s), so they collide with each
    other and the walls throughout."""
    p, x, r, m, meshes = scenes.box_settle_workload(10_000)
    v = w = None
    if stirred:
        rng = np.random.default_rng(1)
        v = rng.normal(0.0, 1.5, x.shape)
        w = rng.normal(0.0, 20.0, x.shape)
    print(f"case config1 stirred={stirred} n={x.shape[0]}", flush=True)
    e = native.Engine(copy_params(p))
    e.upload(x, r, m, v=v, w=w)
    for v9, n3 in meshes:
        e.add_mesh(v9, n3)
    e.set_record(True)
    o = OracleSimulation(p.dt, None, None, params=copy_params(p))
    o.set_state(x, r, m, v=v, w=w)
    for v9, n3 in meshes:
        o.add_mesh_arrays(v9, n3)
    o.set_fast(True)
    o.set_record(True)
    assert_cells_equal(e, o)
    o.step(1999)
    f0 = o.jkr_failures()
    o.step(1)
    e.step(2000)
    npairs = _pairs_equal(o, e)
    _states_equal(o, e)
    st = _counters_equal(o, e, f0)
    assert st["steps"] == 2000
    if stirred:
        assert npairs > 100 and st["wall_contacts"] > 0
    assert_cells_equal(e, o)


# ---- 2: the hopper cone, a bed resting on it and leaving through the outlet ------

@pytest.mark.parametrize("ref_cs", [1.0, 0.031], ids=["one-ref-cell", "ref-cells-3cm"])
def test_config2_hopper_cone_bed(ref_cs):
    """20,000 spheres of the hopper configuration (r ~ U[0.5, 1.5] mm, Hertz +
    Mindlin, E 5e7, 4,096-triangle cone, dt 2e-6 s) packed in the cone and
    moving down at 0.5 m/s: wall contacts on the cone, pair contacts in the
    bed, spheres passing the outlet.  With 3 cm reference cells the domain
    spans several of them (the summation order then differs)."""
    p, x, r, m, meshes = scenes.hopper_test_workload(20_000)
    p.ref_cell_size = ref_cs
    rng = np.random.default_rng(2)
    v = np.zeros_like(x)
    v[:, 1] = -0.5
    v += rng.normal(0.0, 0.05, v.shape)
    print(f"case config2 ref_cs={ref_cs} n={x.shape[0]}", flush=True)
    e = _engine(p, x, r, m, v=v, meshes=meshes)
    o = _oracle(p, x, r, m, v=v, meshes=meshes)
    assert_cells_equal(e, o)
    o.step(1)
    e.step(1)
    assert _pairs_equal(o, e) > 1000
    assert_forces_close(o, e, p)
    st = e.stats()
    assert st["wall_contacts"] > 100 and st["contacts"] > 1000
    below0 = int(np.sum(o.state()["x"][:, 1] < 0.0))
    o.step(399)
    f0 = o.jkr_failures()
    o.step(1)
    e.step(400)
    _pairs_equal(o, e)
    if ref_cs == 1.0:
        _states_equal(o, e)
    else:
        so, se = o.state(), e.download()
        assert np.abs(so["x"] - se["x"]).max() <= 1e-13
        assert np.abs(so["v"] - se["v"]).max() <= 1e-9 * np.abs(so["v"]).max()
    _counters_equal(o, e, f0)
    assert int(np.sum(e.download()["x"][:, 1] < 0.0)) > below0  # spheres left through the outlet
    assert_cells_equal(e, o)


# ---- 3: periodic JKR packing, pencil-numbered cells, three solvers ---------------

def _jkr_env(p, r):
    rr = r[: min(r.size, 4000)]
    rs = rr[:, None] * rr[None, :200] / (rr[:, None] + rr[None, :200])
    e_star = 1.0 / (2 * (1 - p.particle.poissons_ratio ** 2) / p.particle.youngs_modulus)
    return surrogate.envelope(-0.40, 30.0, (0.9 * rs.min(), 1.1 * rs.max()), e_star, p.particle.surface_energy)


@pytest.fixture(scope="module")
def jkr3():
    p, x, r, m, v = scenes.jkr_periodic_workload(20_000)
    p.cell_order = abi.RP_CELL_ORDER_PENCIL
    net = surrogate.train(_jkr_env(p, r), nsamples=400_000, epochs=30, batch=2048, lr=3e-3, seed=4)
    return p, x, r, m, v, net


@pytest.mark.parametrize("solver", ["exact", "shortcut", "surrogate"])
def test_config3_jkr_packing_pencil_cells(jkr3, solver):
    """20,000 spheres of the JKR configuration (r lognormal, median 50 um,
    sigma 0.3; E 1e8, gamma 0.05, Coulomb 0.3; periodic box at solid
    fraction 0.35, Gaussian velocities 0.05 m/s, dt 1e-8 s), cells numbered
    in pencils (forced through rp_params.cell_order), 120 steps."""
    p, x, r, m, v, net = jkr3
    print(f"case config3 solver={solver} n={x.shape[0]}", flush=True)
    e = _engine(p, x, r, m, v=v, solver=solver, net=net)
    o = _oracle(p, x, r, m, v=v)
    c = assert_cells_equal(e, o)
    assert int(c["grid"]["shift"].max()) == 3  # pencils of 8 x 8 cells
    if solver == "surrogate":
        _surrogate_against_exact(o, e, p, net, 120)
    else:
        o.step(119)
        f0 = o.jkr_failures()
        o.step(1)
        e.step(120)
        assert _pairs_equal(o, e) > 10_000
        _states_equal(o, e)
        _counters_equal(o, e, f0)
    assert_cells_equal(e, o)


def _surrogate_against_exact(o, e, p, net, steps):
    """The surrogate form against the oracle's exact solve.  Pair recording
    is off while the surrogate serves (a recording step uses the exact
    solve): the first step's forces within 1e-6 of the per-particle sum of
    contact-force magnitudes, with contacts served and check mode's relative
    force error of every served contact at most 1e-3 (measured ~1e-9); after
    `steps` steps the states still agree to 1e-6 of the largest displacement
    and velocity; one more, recorded, step gives the oracle's pair list."""
    x0 = e.download()["x"]
    e.set_record(False)
    e.set_jkr_solver("surrogate", net, check=True)
    o.step(1)
    e.step(1)
    assert_forces_close(o, e, p, rtol=1e-6)
    walls = e.stats()["wall_contacts"]
    js = e.jkr_stats()
    assert js["served"] > 0 and js["err_max"] <= 1e-3, js
    o.step(steps - 1)
    e.step(steps - 1)
    js = e.jkr_stats()
    assert js["served"] > 0 and js["err_max"] <= 1e-3, js
    so, se = o.state(), e.download()
    disp = np.abs(so["x"] - x0).max()
    assert np.abs(so["x"] - se["x"]).max() <= 1e-6 * disp
    assert np.abs(so["v"] - se["v"]).max() <= 1e-6 * np.abs(so["v"]).max()
    e.set_record(True)
    o.step(1)
    e.step(1)
    assert _pairs_equal(o, e) > 10_000
    return walls


# ---- 4: batched sJKR ensemble, each instance against the oracle -----------------

def test_config4_ensemble_instances_match_oracle():
    """16 instances of the batched configuration (N_k ~ U{500..2500}, r ~
    U[1, 3] mm, per-instance friction, restitution and surface energy,
    sJKR + Coulomb, box STL, dt 5e-6 s; lattice spacing 5 mm so many spheres
    start in contact, random velocities), stepped by one batched engine with
    the autotuner and the HIP-graph loop: each instance equals the oracle run
    with that instance's material, bit for bit."""
    p, counts, mats, x, r, m, meshes = scenes.batch_ensemble_workload(16, nmin=500, nmax=2500, spacing=0.005)
    rng = np.random.default_rng(9)
    v = rng.normal(0.0, 0.2, x.shape)
    n = x.shape[0]
    print(f"case config4 instances={len(counts)} n={n}", flush=True)
    e = native.Engine(copy_params(p))
    e.upload_batch(counts, mats, x, r, m, v=v)
    for v9, n3 in meshes:
        e.add_mesh(v9, n3)
    e.step(20)
    e.autotune(10)
    e.step(280)
    got = e.download()
    assert list(e.instance_counts()) == counts
    off = 0
    for k, (cnt, mat) in enumerate(zip(counts, mats)):
        q = copy_params(p)
        q.particle = mat
        sl = slice(off, off + cnt)
        o = _oracle(q, x[sl], r[sl], m[sl], v=v[sl], meshes=meshes, steps=300)
        so = o.state()
        for f in FIELDS:
            assert np.array_equal(so[f], got[f][sl]), (k, f)
        off += cnt
    assert e.stats()["contacts"] > 0


# ---- 5: JKR bed in the turning drum ----------------------------------------------

@pytest.fixture(scope="module")
def drum5():
    p, x, r, m, meshes = scenes.drum_test_workload(20_000)
    net = surrogate.train(_jkr_env(p, r), nsamples=400_000, epochs=30, batch=2048, lr=3e-3, seed=5)
    return p, x, r, m, meshes, net


@pytest.mark.parametrize("solver", ["shortcut", "surrogate"])
def test_config5_turning_drum_jkr_bed(drum5, solver):
    """20,000 spheres of the capacity run (r ~ U[0.4, 0.8] mm, JKR gamma
    0.01 + Coulomb, 16,384-triangle drum turning at 30 rpm, dt 2e-6 s) in a
    6 cm drum: chains of touching spheres (overlaps of a settled bed, JKR
    roots) whose lowest members press on the turning side; 300 steps."""
    p, x, r, m, meshes, net = drum5
    motions = [(0, scenes.DRUM_ORIGIN, scenes.DRUM_OMEGA)]
    print(f"case config5 solver={solver} n={x.shape[0]}", flush=True)
    e = _engine(p, x, r, m, meshes=meshes, motions=motions, solver=solver, net=net)
    o = _oracle(p, x, r, m, meshes=meshes, motions=motions)
    assert_cells_equal(e, o)
    if solver == "surrogate":
        walls = _surrogate_against_exact(o, e, p, net, 300)
    else:
        o.step(1)
        e.step(1)
        walls = e.stats()["wall_contacts"]
        o.step(298)
        f0 = o.jkr_failures()
        o.step(1)
        e.step(299)
        assert _pairs_equal(o, e) > 10_000
        _states_equal(o, e)
        _counters_equal(o, e, f0)
    assert walls > 100  # the pressed chains' first step on the turning side
    assert_cells_equal(e, o)
