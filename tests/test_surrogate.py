"""CPU checks of the surrogate's reference pieces (no GPU).

* the FP64 backward pass of oracle/mlp.py against central finite differences;
* the BF16 rounding helper against known values;
* the scaled JKR form the surrogate learns: for any R*, E*, gamma the oracle's
  Newton solve (physics.rs:132-219 restated) returns a = alpha s for
  delta = (alpha^2 - sqrt(alpha)) t, and F = pi gamma R* ((32/3) alpha^3 -
  8 alpha^1.5) (unless clipped at pull-off)."""
import numpy as np

from oracle import mlp
from oracle import sim as osim
from rusty_particles_amd import abi


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, mlp.NPARAMS)
    x = rng.uniform(-1, 1, (32, 4))
    t = rng.uniform(-1, 1, (32, 2))
    loss, g, _ = mlp.forward_backward(p, x, t)
    idx = rng.choice(mlp.NPARAMS, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp.forward_backward(p + e, x, t)[0] - mlp.forward_backward(p - e, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3], dtype=np.float32)
    r = mlp.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -7        # above the tie rounds up
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def _scales(R, E, g):
    s = (R * R * 8 * np.pi * g / E) ** (1 / 3)
    return s, s * s / R


def _draw(rng):
    while True:
        R = 10 ** rng.uniform(-6, -2)
        E = 10 ** rng.uniform(5, 9)
        g = 10 ** rng.uniform(-3, 0)
        s, t = _scales(R, E, g)
        if 1e-3 <= s / R <= 0.9:
            return R, E, g, s, t


N_Y = np.array([0.0, 1.0, 0.0])


def test_reference_newton_reaches_the_larger_root_where_served():
    """Served set (d in [-0.40, 0] or d >= 0.2, s/R* in [1e-3, 0.9]): the
    reference's iteration converges to the larger root of alpha^2 - sqrt(alpha)
    = d, and F / (pi gamma R*) = (32/3) alpha^3 - 8 alpha^1.5 unless clipped."""
    rng = np.random.default_rng(1)
    for k in range(300):
        R, E, g, s, t = _draw(rng)
        alpha = rng.uniform(0.622, 1.0) if k % 3 == 0 else rng.uniform(1.14, 6.0)  # d <= 0 or d >= 0.2
        d = alpha ** 2 - np.sqrt(alpha)
        assert d <= 0.0 or d >= 0.2
        f, a, conv = osim.normal_law(64, abi.RP_NORMAL_JKR, d * t, N_Y, N_Y * 0.0, E, R, gamma=g)
        assert conv == 1
        # the iteration stops once a step is below 1e-6 R*: alpha to ~(1e-6 R*/s)^2
        assert abs(a / s - alpha) < 1e-9 + 4.0 * (1e-6 * R / s) ** 2, (R, E, g, alpha, a / s)
        fhat = (32 / 3) * alpha ** 3 - 8 * alpha ** 1.5
        if fhat >= -2 / 3 + 1e-6:
            assert abs(f[1] / (np.pi * g * R) - fhat) < 1e-6 * max(1.0, abs(fhat))
        else:
            assert f[1] == 0.0


def test_reference_newton_misses_the_root_for_small_positive_overlap():
    """0 < d < 0.1575: the start alpha0 = sqrt(d) lies left of the minimum at
    4^(-2/3), so the iteration never reaches the larger root; the surrogate
    leaves these contacts (0 < d < 0.2) to the exact solve."""
    rng = np.random.default_rng(2)
    for _ in range(100):
        R, E, g, s, t = _draw(rng)
        d = rng.uniform(0.005, 0.15)
        f, a, conv = osim.normal_law(64, abi.RP_NORMAL_JKR, d * t, N_Y, N_Y * 0.0, E, R, gamma=g)
        assert a / s < 0.25 ** (2 / 3)


def test_no_root_shortcut_threshold():
    """delta < t d_min - 1e-3 R* - 1e-8 (R* + t + |delta|) (the engine's
    shortcut): the oracle's 20 iterations end with a rejected root and zero
    force.  Just above d_min (a root exists) they converge."""
    from rusty_particles_amd.surrogate import D_MIN
    assert abs(D_MIN - (0.25 ** (4 / 3) - 0.25 ** (1 / 3))) < 1e-16
    a_c = 0.25 ** (2 / 3)
    assert abs(D_MIN - (a_c * a_c - np.sqrt(a_c))) < 1e-15
    rng = np.random.default_rng(3)
    for _ in range(300):
        R, E, g, s, t = _draw(rng)
        thr = D_MIN * t - 1e-3 * R
        delta = thr - abs(thr) * 10 ** rng.uniform(-7, 1) - 1e-8 * (R + t + abs(thr))
        f, a, conv = osim.normal_law(64, abi.RP_NORMAL_JKR, delta, N_Y, N_Y * 0.0, E, R, gamma=g)
        assert conv == 0 and not np.any(f), (R, E, g, delta / t)
        delta = (D_MIN + 0.05) * t
        f, a, conv = osim.normal_law(64, abi.RP_NORMAL_JKR, delta, N_Y, N_Y * 0.0, E, R, gamma=g)
        assert conv == 1


def test_far_prefilter_bounds_the_shortcut():
    """The engine skips the per-contact cube root below the shortcut threshold
    of the largest possible R* (r_max / 2): the threshold falls with R*, so
    that bound implies the shortcut for every smaller R*."""
    from rusty_particles_amd.surrogate import D_MIN
    rng = np.random.default_rng(4)
    for _ in range(200):
        E, g = 10 ** rng.uniform(5, 9), 10 ** rng.uniform(-3, 0)
        C = 8 * np.pi * g / E
        rmax = 10 ** rng.uniform(-6, -2)
        rm = 0.5 * rmax
        tm = np.cbrt(rm) * C ** (2 / 3)
        far = (D_MIN * tm - 1e-3 * rm - 1e-8 * (rm + tm)) * (1 + 1e-12)
        R = rm * rng.uniform(1e-3, 1.0, 50)
        t = np.cbrt(R * R * C) ** 2 / R
        ov = far * rng.uniform(1.0, 50.0, 50)
        ov = ov - 1e-8 * np.abs(ov)
        assert np.all(ov < D_MIN * t - 1e-3 * R - 1e-8 * (R + t + np.abs(ov)))
