"""The captured step loop (HIP graph replays of 20 launches) must give the
same state, bit for bit, as launching every kernel individually -- across
rebuilds, skin re-issues, removals and the JKR surrogate."""
import numpy as np
import pytest

from helpers import copy_params, dense_scene, jkr_dimers, params
from rusty_particles_amd import abi, native, scenes

pytestmark = pytest.mark.gpu


def _run(p, sc, graph, chunks, meshes=(), solver=None, net=None):
    e = native.Engine(copy_params(p))
    e.set_graph(graph)
    e.upload(sc["x"], sc["radius"], sc["mass"], sc.get("fixed"), v=sc.get("v"), w=sc.get("w"))
    for v9, n3 in meshes:
        e.add_mesh(v9, n3)
    if solver:
        e.set_jkr_solver(solver, net)
    for c in chunks:
        e.step(c)
    return e.download(), e.stats()


def _same(a, b):
    for k in ("x", "v", "w", "a", "tau", "rx", "rv", "rw"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("normal", [abi.RP_NORMAL_HERTZ, abi.RP_NORMAL_SJKR])
def test_graph_matches_plain_launches(normal):
    sc = dense_scene(counts=(12, 12, 12), seed=21, vel=0.3)
    p = params(1e-5, (-0.08,) * 3, (0.08,) * 3, normal=normal, gamma=0.01 if normal == abi.RP_NORMAL_SJKR else 0.0)
    v9, n3 = scenes.box_mesh((-0.07,) * 3, (0.07,) * 3)
    chunks = [7, 61, 133]
    a, sa = _run(p, sc, True, chunks, [(v9, n3)])
    b, sb = _run(p, sc, False, chunks, [(v9, n3)])
    _same(a, b)
    assert sa["steps"] == sb["steps"] == sum(chunks) and sa["rebuilds"] == sb["rebuilds"]


def test_graph_with_removals_and_reissues():
    sc = dense_scene(counts=(10, 10, 10), seed=22, vel=0.05)
    sc["x"][:10, 1] = -0.05 - 0.9995  # removed after the first step (1 m below bounds_min.y)
    sc["v"][:10] = [0.0, -10.0, 0.0]
    sc["v"][10:20] = [3.0, 0.0, 0.0]  # fast: skin violations and re-issued launches
    p = params(1e-5, (-0.05,) * 3, (0.05,) * 3)
    a, _ = _run(p, sc, True, [150])
    b, _ = _run(p, sc, False, [150])
    assert a["x"].shape[0] == 990
    _same(a, b)


def test_graph_with_surrogate():
    from rusty_particles_amd import surrogate
    e_star = 1.0 / (2 * (1 - 0.3 ** 2) / 1e8)
    env = surrogate.envelope(-0.40, 30.0, (5e-6, 1.2e-4), e_star, 0.05)
    net = surrogate.train(env, nsamples=200_000, epochs=20, batch=2048, lr=3e-3, seed=3)
    sc = jkr_dimers(np.random.default_rng(1).uniform(-0.5, 20.0, 2000), vel=0.2, seed=4)
    p = params(1e-8, (-0.01,) * 3, (0.01,) * 3, normal=abi.RP_NORMAL_JKR, E=1e8, nu=0.3, mu=0.3, gamma=0.05)
    p.particle.density = 1500.0
    a, _ = _run(p, sc, True, [3, 95], solver="surrogate", net=net)
    b, _ = _run(p, sc, False, [3, 95], solver="surrogate", net=net)
    _same(a, b)


def test_autotune_keeps_state_and_results():
    """The autotuner restores the state it tried candidates on, and the
    skin / cell size it picks change no result."""
    sc = dense_scene(counts=(12, 12, 12), seed=23, vel=0.2)
    p = params(1e-5, (-0.08,) * 3, (0.08,) * 3)
    v9, n3 = scenes.box_mesh((-0.07,) * 3, (0.07,) * 3)
    e = native.Engine(copy_params(p))
    e.upload(sc["x"], sc["radius"], sc["mass"], v=sc["v"], w=sc["w"])
    e.add_mesh(v9, n3)
    e.step(13)
    before = e.download()
    skin, cell = e.autotune(10)
    assert skin > 0.0
    _same(before, e.download())
    e.step(120)
    ref, _ = _run(p, sc, True, [13, 120], [(v9, n3)])
    _same(ref, e.download())
    assert e.stats()["steps"] == 133


def test_timing_counts_every_force_launch():
    """Timing mode: every force launch is counted once (its dispatch records
    its own start and end), the rebuild chain at least at the call's start,
    the timed force work fits in the call's wall time, and the state is the
    one an untimed run reaches."""
    import time
    sc = dense_scene(counts=(12, 12, 12), seed=23, vel=0.3)
    p = params(1e-5, (-0.08,) * 3, (0.08,) * 3)
    e = native.Engine(copy_params(p))
    e.upload(sc["x"], sc["radius"], sc["mass"], sc.get("fixed"), v=sc.get("v"), w=sc.get("w"))
    e.step(3)
    e.set_timing(True)
    e.synchronize()
    t0 = time.perf_counter()
    e.step(25)
    e.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3
    tm = e.timing()
    e.set_timing(False)
    assert tm["force"][1] == 25
    assert tm["rebuild"][1] >= 1
    assert 0.0 < tm["force"][0] <= wall_ms
    # timing does not change the state: the same 28 steps without it
    f = native.Engine(copy_params(p))
    f.upload(sc["x"], sc["radius"], sc["mass"], sc.get("fixed"), v=sc.get("v"), w=sc.get("w"))
    f.step(28)
    _same(e.download(), f.download())


def test_tuned_skin_keeps_up_with_falling_particles():
    """A lattice released at rest under gravity: at rest every skin looks
    cheap, but the particles speed up after tuning.  The autotuner keeps the
    skin above the fall over 1,000 steps and the engine widens a tuned skin
    that rebuilds at more than a quarter of the checks, so the last 500
    steps of a long call rebuild at fewer than a third of the checks -- and
    the result equals an untuned run's bit for bit (skin and cell size never
    change one)."""
    p, counts, mats, x, r, m, meshes = scenes.batch_ensemble_workload(ninst=1, nmin=6000, nmax=6000)
    p.particle = mats[0]
    sc = dict(x=x, radius=r, mass=m, fixed=np.zeros(x.shape[0], np.uint8))

    def make():
        e = native.Engine(copy_params(p))
        e.upload(sc["x"], sc["radius"], sc["mass"], sc["fixed"])
        for v9, n3 in meshes:
            e.add_mesh(v9, n3)
        return e

    tuned = make()
    tuned.step(2)
    skin, cell = tuned.autotune(20)
    print(f"autotune: skin {skin:.3g} m, cell {cell:.3g} m (r_max {r.max():.3g} m)")
    tuned.step(2998)
    r0 = tuned.stats()["rebuilds"]
    tuned.step(500)
    late = tuned.stats()["rebuilds"] - r0
    print(f"rebuilds in the last 500 steps: {late}")
    assert late < 17
    plain = make()
    plain.step(3500)
    _same(tuned.download(), plain.download())


@pytest.mark.parametrize("normal,solver,batched", [(abi.RP_NORMAL_SJKR, None, False),
                                                   (abi.RP_NORMAL_JKR, "shortcut", False),
                                                   (abi.RP_NORMAL_SJKR, None, True)],
                         ids=["sjkr", "jkr-shortcut", "sjkr-batched"])
def test_tilings_of_the_zero_force_pass_agree(normal, solver, batched):
    """One lane per particle and the wave-dealt zero-force pass give the same
    state bit for bit (integer counts, forces added in list order)."""
    sc = dense_scene(counts=(12, 12, 12), seed=31, vel=0.3)
    p = params(1e-5, (-0.08,) * 3, (0.08,) * 3, normal=normal, gamma=0.05)
    v9, n3 = scenes.box_mesh((-0.07,) * 3, (0.07,) * 3)
    out = {}
    for tiling in ("lane", "wave"):
        e = native.Engine(copy_params(p))
        if batched:
            n = sc["x"].shape[0]
            counts = [n // 3, n // 3, n - 2 * (n // 3)]
            mats = [abi.RpMaterial(1e7, 0.3, 2500.0, 0.2 + 0.2 * k, 0.5, 0.02 * k) for k in range(3)]
            e.upload_batch(counts, mats, sc["x"], sc["radius"], sc["mass"], sc["fixed"], v=sc["v"], w=sc["w"])
        else:
            e.upload(sc["x"], sc["radius"], sc["mass"], sc.get("fixed"), v=sc.get("v"), w=sc.get("w"))
        e.add_mesh(v9, n3)
        if solver:
            e.set_jkr_solver(solver)
        e.set_tiling(tiling)
        assert e.tiling() == tiling
        for c in (7, 61, 52):
            e.step(c)
        out[tiling] = (e.download(), e.stats())
    _same(out["lane"][0], out["wave"][0])
    for k in ("contacts", "wall_contacts", "jkr_failures", "rebuilds"):
        assert out["lane"][1][k] == out["wave"][1][k], k
