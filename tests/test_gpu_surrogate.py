"""JKR surrogate on the GPU: MFMA layout, inference and gradients against the
FP64 numpy reference (oracle/mlp.py), labels against the CPU oracle's Newton
solve, training to a polishable accuracy, and the engine's three JKR solvers
(exact, shortcut, surrogate) against each other on the same scenes."""
import numpy as np
import pytest

from helpers import copy_params, jkr_dimers, params, run_engine
from oracle import mlp
from oracle import sim as osim
from rusty_particles_amd import abi, native, surrogate

pytestmark = pytest.mark.gpu

E_STAR = 1.0 / (2 * (1 - 0.3 ** 2) / 1e8)  # E = 1e8, nu = 0.3 for both particles
ENV = surrogate.envelope(-0.40, 30.0, (5e-6, 1.2e-4), E_STAR, 0.05)


@pytest.fixture(scope="module")
def net():
    return surrogate.train(ENV, nsamples=400_000, epochs=40, batch=2048, lr=3e-3, seed=3)


def test_mfma_tile_layout():
    rng = np.random.default_rng(3)
    a = rng.uniform(-1, 1, (16, 16)).astype(np.float32)
    b = rng.uniform(-1, 1, (16, 16)).astype(np.float32)
    c = surrogate.mfma_probe(a, b)
    ref = mlp.bf16(a) @ mlp.bf16(b)
    assert np.allclose(c, ref, rtol=1e-5, atol=1e-5)


def test_fused_inference_matches_reference():
    rng = np.random.default_rng(4)
    p = rng.normal(0, 0.3, abi.RP_MLP_PARAMS).astype(np.float32)
    x = rng.uniform(-1, 1, (1000, 4)).astype(np.float32)  # not a multiple of 16: tail rows
    net = surrogate.JkrSurrogate(p, ENV)
    y = net.infer_scaled(x)
    ref = mlp.forward(p.astype(np.float64), x, bf=True)
    assert np.abs(y - ref).max() < 2e-3 * max(1.0, np.abs(ref).max())
    # a row's output does not depend on where it sits in the batch
    perm = rng.permutation(1000)
    assert np.array_equal(net.infer_scaled(x[perm]), y[perm])


def test_gradient_matches_fp64_backprop():
    """GPU forward + backward (BF16 operands, FP32 accumulation) vs the FP64
    backprop with the same BF16 rounding points: per-tensor relative L2 error."""
    rng = np.random.default_rng(5)
    p = rng.normal(0, 0.3, abi.RP_MLP_PARAMS).astype(np.float32)
    x = rng.uniform(-1, 1, (256, 4)).astype(np.float32)
    t = rng.uniform(-1, 1, (256, 2)).astype(np.float32)
    loss, g = surrogate.gradient(p, x, t)
    rloss, rg, _ = mlp.forward_backward(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), bf=True)
    assert abs(loss - rloss) < 1e-4 * rloss
    gu, ru = mlp.unpack(g), mlp.unpack(rg)
    for name, _ in mlp.SIZES:
        err = np.linalg.norm(gu[name] - ru[name]) / np.linalg.norm(ru[name])
        assert err < 1e-2, (name, err)


def test_labels_match_oracle_newton():
    x, t, raw = surrogate.samples(ENV, 2000, seed=11)
    assert raw[:, 7].min() == 1.0  # every sample converged and lies in the served set
    d = raw[:, 6]
    assert np.all((d <= 0.0) | (d >= surrogate.GAP_HI)) and d.min() >= -0.40 and d.max() <= 30.0
    n = np.array([0.0, 1.0, 0.0])
    for row in raw[::40]:
        delta, R, E, g, a = row[:5]
        _, ao, conv = osim.normal_law(64, abi.RP_NORMAL_JKR, delta, n, n * 0.0, E, R, gamma=g)
        assert conv == 1 and abs(ao - a) <= 1e-12 * a


def test_training_reaches_polishable_accuracy(net):
    """The predicted alpha is close enough that the engine's four FP64 Newton
    steps land on the root; the network's own force output is reported, not
    used."""
    assert net.epoch_loss[-1] < net.epoch_loss[0] * 0.05
    x, t, raw = surrogate.samples(ENV, 20000, seed=99)
    in_lo, in_hi, out_lo, out_hi = net.scale()
    y = net.infer_scaled(x).astype(np.float64)
    alpha_hat = out_lo[0] + 0.5 * (y[:, 0] + 1.0) * (out_hi[0] - out_lo[0])
    d = raw[:, 6]
    s = (raw[:, 1] ** 2 * 8 * np.pi * raw[:, 3] / raw[:, 2]) ** (1 / 3)
    alpha = raw[:, 4] / s
    rel = np.abs(alpha_hat - alpha) / alpha
    print("alpha relative error: median %.2e  max %.2e" % (np.median(rel), rel.max()))
    assert np.median(rel) < 1e-2 and rel.max() < 0.15
    a = np.clip(alpha_hat, 0.45, 1e4)
    for _ in range(4):
        a = a - (a * a - np.sqrt(a) - d) / (2 * a - 0.5 / np.sqrt(a))
    assert np.max(np.abs(a - alpha) / alpha) < 1e-8


def _jkr_params(sc, dt=1e-8, periodic=(0, 0, 0), box=0.01):
    p = params(dt, (-box,) * 3, (box,) * 3, normal=abi.RP_NORMAL_JKR, tangential=abi.RP_TANGENTIAL_COULOMB,
               E=1e8, nu=0.3, mu=0.3, e=0.5, gamma=0.05, periodic=periodic)
    p.particle.density = 1500.0
    return p


def _run(p, sc, solver, net=None, steps=1, check=False):
    e = native.Engine(copy_params(p))
    e.upload(sc["x"], sc["radius"], sc["mass"], sc["fixed"], v=sc["v"], w=sc["w"])
    e.set_jkr_solver(solver, net, check=check)
    e.step(steps)
    return e


def _dimer_d(n=6000, seed=7):
    """Scaled overlaps covering every regime: no root, the clip, the served
    set on both sides of the gap, the gap itself and beyond the envelope."""
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.uniform(-3.0, -0.3, n // 6), rng.uniform(-0.47, 0.0, n // 6),
                           rng.uniform(0.0, 0.25, n // 6), rng.uniform(0.2, 30.0, n // 3),
                           rng.uniform(30.0, 60.0, n // 6)])


def test_shortcut_identical_to_exact():
    sc = jkr_dimers(_dimer_d(), vel=0.05)
    p = _jkr_params(sc)
    a = _run(p, sc, "exact", steps=20)
    b = _run(p, sc, "shortcut", steps=20)
    da, db = a.download(), b.download()
    for k in ("x", "v", "w", "a", "tau", "rx", "rv", "rw"):
        assert np.array_equal(da[k], db[k]), k
    assert a.stats()["jkr_failures"] == b.stats()["jkr_failures"] > 0


def test_surrogate_forces_match_exact(net):
    d = _dimer_d()
    sc = jkr_dimers(d)
    p = _jkr_params(sc)
    ex = _run(p, sc, "exact")
    su = _run(p, sc, "surrogate", net, check=True)
    st = su.jkr_stats()
    served = ((d >= -0.40) & (d <= 0.0)) | ((d >= 0.2) & (d <= 30.0))
    print(st)
    assert st["served"] == 2 * served.sum() and st["fallback"] == 0
    assert st["err_max"] < 1e-8
    ae, asu = ex.download()["a"], su.download()["a"]
    g = np.array([0.0, -9.81, 0.0])
    fe, fs = (ae - g) * sc["mass"][:, None], (asu - g) * sc["mass"][:, None]
    rstar = np.repeat(sc["rstar"], 2)
    scale = np.maximum(np.linalg.norm(fe, axis=1), 2 / 3 * np.pi * 0.05 * rstar)
    err = np.linalg.norm(fs - fe, axis=1) / scale
    assert err.max() < 1e-8, err.max()
    assert ex.stats()["jkr_failures"] == su.stats()["jkr_failures"]


def test_surrogate_tilings_agree(net):
    """The surrogate form with the zero-force pass in one lane per particle
    and dealt over the wavefront: same state bit for bit, same served rows."""
    d = np.random.default_rng(13).uniform(-0.5, 20.0, 2000)
    sc = jkr_dimers(d, vel=0.2, seed=13)
    p = _jkr_params(sc)
    out = {}
    for tiling in ("lane", "wave"):
        e = native.Engine(copy_params(p))
        e.upload(sc["x"], sc["radius"], sc["mass"], sc["fixed"], v=sc["v"], w=sc["w"])
        e.set_jkr_solver("surrogate", net)
        e.set_tiling(tiling)
        e.step(57)
        out[tiling] = (e.download(), e.jkr_stats())
    for k in ("x", "v", "w", "a", "tau"):
        assert np.array_equal(out["lane"][0][k], out["wave"][0][k]), k
    assert out["lane"][1]["served"] == out["wave"][1]["served"] > 0


def test_surrogate_trajectory_tracks_exact(net):
    """Dimers with approach velocities cross every regime during 300 steps;
    the surrogate run stays within 1e-9 of the particle radius."""
    d = np.random.default_rng(9).uniform(-0.5, 20.0, 3000)
    sc = jkr_dimers(d, vel=0.2, seed=9)
    p = _jkr_params(sc)
    ex = _run(p, sc, "exact", steps=300)
    su = _run(p, sc, "surrogate", net, steps=300)
    xe, xs = ex.download()["x"], su.download()["x"]
    assert np.max(np.abs(xe - xs)) < 1e-9 * sc["radius"].min()
    st = su.jkr_stats()
    print(st)
    assert st["served"] > 0 and st["fallback"] <= st["served"] // 100


# A wide envelope: stiff and weakly adhesive pairs down to s / R* = 1e-3,
# where the reference's solve starts far from the root.
WIDE = surrogate.envelope(-0.40, 30.0, (5e-6, 1.2e-4), (1e7, 1e10), (0.005, 0.1))


def _newton_steps(d, a0, cap=20):
    """FP64 Newton steps on alpha^2 - sqrt(alpha) = d from a0 (clamped as the
    engine's polish clamps it) until the engine's residual test passes;
    cap + 1 when it never does within cap steps."""
    a = np.clip(a0, 0.45, 1e4)
    steps = np.full(d.shape, cap + 1)
    for k in range(cap + 1):
        ok = (a > 0.45) & (np.abs(a * a - np.sqrt(a) - d) <= 1e-10 * (1.0 + np.abs(d)))
        steps[ok & (steps > cap)] = k
        a = a - (a * a - np.sqrt(a) - d) / (2 * a - 0.5 / np.sqrt(a))
    return steps


def test_network_gain_over_exact_solve():
    """What the trained network gains over the solve without it, on 20,000
    served contacts drawn over a wide envelope.  Without the network every
    contact runs the reference's Newton loop (physics.rs:171-189) from
    sqrt(R* delta) or R*/2, a data-dependent number of iterations (a
    wavefront waits for its slowest lane).  With it, the engine runs a fixed
    four FP64 Newton steps from the predicted radius: the test shows that
    those four steps reach the root for the served contacts, that the
    reference needs more iterations than that, and that an undertrained
    network's start does not get there."""
    net = surrogate.train(WIDE, nsamples=400_000, epochs=40, batch=2048, lr=3e-3, seed=5)
    weak = surrogate.train(WIDE, nsamples=4_096, epochs=1, batch=2048, lr=3e-3, seed=5)
    x, t, raw = surrogate.samples(WIDE, 20_000, seed=123)
    delta, rstar, estar, gamma, a_lab, _, d, ok = raw.T
    assert ok.min() == 1.0
    s = (rstar ** 2 * 8 * np.pi * gamma / estar) ** (1 / 3)
    a_ref, its_ref, conv = osim.jkr_solve(delta, estar, rstar, gamma)
    assert conv.all() and np.allclose(a_ref, a_lab, rtol=1e-12, atol=0)
    steps_net = _newton_steps(d, net.predict(d, rstar, estar, gamma)[:, 0])
    steps_weak = _newton_steps(d, weak.predict(d, rstar, estar, gamma)[:, 0])
    # a wavefront of 64 contacts waits for its slowest lane
    wave_ref = its_ref[: its_ref.size // 64 * 64].reshape(-1, 64).max(1)
    neg = d <= 0.0
    print("\nNewton iterations per served contact (20,000 contacts, s/R* %.1e..%.2f):" % ((s / rstar).min(),
                                                                                        (s / rstar).max()))
    print("  reference solve:    mean %.2f  (d<=0: %.2f, d>0: %.2f)  max %d  per 64-lane wave mean %.2f"
          % (its_ref.mean(), its_ref[neg].mean(), its_ref[~neg].mean(), its_ref.max(), wave_ref.mean()))
    print("  trained network:    steps to the root mean %.2f  max %d  within 4: %.4f"
          % (steps_net.mean(), steps_net.max(), np.mean(steps_net <= 4)))
    print("  undertrained:       within 4: %.4f" % np.mean(steps_weak <= 4))
    assert np.mean(steps_net <= 4) >= 0.999      # the polish is exact for the served contacts
    assert its_ref.mean() >= 1.5 * steps_net.mean()
    assert wave_ref.mean() > 4.0                 # fixed four steps beat the wave's slowest lane
    assert np.mean(steps_weak <= 4) < 0.9        # the gain comes from training
