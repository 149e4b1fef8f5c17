"""The benchmark configurations at their full sizes on the GPU, with checks
that do not need the oracle to follow every step (simulation.rs:173-562).

* 1M spheres of the hopper configuration standing in columns in its
  4,096-triangle cone (~2M pair and ~14k wall contacts from the first step;
  the domain lies in one reference cell): the first step against the oracle's fast enumeration,
  pair lists and state bit for bit, cell ids and sort permutation included;
  then 30 more steps: count conserved, finite state, wall contacts persist.
* 2M spheres in the periodic JKR crystal (pencil-numbered cells at this
  size): the first step bit for bit against the oracle; then 40 steps of the
  shortcut solver, after which the total linear momentum has changed by
  exactly M g t up to the rounding of the per-particle force sums.  Pair
  forces are antisymmetric to the bit (same overlap, R*, |F|; normal and
  relative velocity negated exactly), so a dropped, doubled or one-sided
  contact shows as an imbalance of about |F| dt, several orders of magnitude
  above the bound.
* 6.7M spheres in 256 instances of the ensemble configuration: one batched
  step, then the smallest, the largest and two random instances bit for bit
  against the oracle run with their own material; 30 more steps: per-instance
  counts conserved, finite state.
* 16M spheres in the turning drum: 20 steps, then a recorded step: the
  partner sets of 3,000 sampled spheres equal a KD-tree rebuild on the host
  with the reference's pair predicate (|d|^2 < (1.5 r_sum)^2 for JKR,
  simulation.rs:341-356); count conserved, finite state.
"""
import numpy as np
import pytest

from helpers import assert_cells_equal, copy_params
from oracle import OracleSimulation
from rusty_particles_amd import abi, native, scenes

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

FIELDS = ("x", "v", "w", "a", "tau", "rx", "rv", "rw")


def _oracle(p, x, r, m, v=None, meshes=(), fixed=None):
    o = OracleSimulation(p.dt, None, None, params=copy_params(p))
    o.set_state(x, r, m, fixed, v=v)
    for v9, n3 in meshes:
        o.add_mesh_arrays(v9, n3)
    o.set_fast(True)
    o.set_record(True)
    return o


def _engine(p, x, r, m, v=None, meshes=(), solver=None, fixed=None):
    e = native.Engine(copy_params(p))
    e.upload(x, r, m, fixed, v=v)
    for v9, n3 in meshes:
        e.add_mesh(v9, n3)
    if solver:
        e.set_jkr_solver(solver)
    return e


def _first_step_equal(p, x, r, m, v=None, meshes=(), fixed=None):
    """One recorded step on the engine and the oracle: pairs and state bit for
    bit; returns the engine (recording off) and the oracle."""
    e = _engine(p, x, r, m, v, meshes, fixed=fixed)
    e.set_record(True)
    o = _oracle(p, x, r, m, v, meshes, fixed=fixed)
    e.step(1)
    o.step(1)
    oo, oj = o.pairs()
    eo, ej = e.pairs()
    assert np.array_equal(oo, eo) and np.array_equal(oj, ej), "pair lists differ"
    so, se = o.state(), e.download()
    for k in FIELDS:
        assert np.array_equal(so[k], se[k]), k
    e.set_record(False)
    return e, o, oj.size


def _finite(st):
    for k in ("x", "v", "w"):
        assert np.isfinite(st[k]).all(), k


def test_hopper_bed_1m():
    p, x, r, m, meshes, fixed = scenes.hopper_bed_workload(1_000_000)
    n = x.shape[0]
    print(f"case fullsize hopper-bed n={n}", flush=True)
    e, o, npairs = _first_step_equal(p, x, r, m, meshes=meshes, fixed=fixed)
    st = e.stats()
    assert npairs > 1_500_000 and st["contacts"] > 1_500_000 and st["wall_contacts"] > 5_000, st
    assert_cells_equal(e, o)
    e.step(300)
    assert e.count() == n
    s = e.download()
    _finite(s)
    assert np.array_equal(s["x"][fixed == 1], x[fixed == 1])  # fixed feet stay put
    st = e.stats()
    # the columns stand: their contacts persist
    assert st["contacts"] > 0.95 * npairs and st["wall_contacts"] > 5_000, st


def _momentum(st, m):
    return (m[:, None] * st["v"]).sum(0) + (m[:, None] * st["rv"]).sum(0)


def test_jkr_crystal_2m_momentum():
    p, x, r, m, v = scenes.jkr_dense_workload(2_000_000)
    n = x.shape[0]
    print(f"case fullsize jkr-dense n={n}", flush=True)
    e, o, npairs = _first_step_equal(p, x, r, m, v)
    assert npairs > 10 * n
    grid = e.cells(rebuild=False)["grid"]
    assert max(int(s) for s in grid["shift"]) == 3, "pencil numbering expected at this size"
    _, _, fscale = o.forces()  # per-particle sum of contact-force magnitudes
    del o
    e.set_jkr_solver("shortcut")
    p0 = _momentum(e.download(), m)
    steps = 40
    e.step(steps)
    st = e.download()
    assert e.count() == n
    _finite(st)
    g = np.array(p.gravity[:])
    dp = _momentum(st, m) - p0 - m.sum() * g * (steps * p.dt)
    # rounding of each particle's force sum, of F/m and of the half kicks:
    # a few ulps of the contact forces per particle and step
    bound = 64.0 * np.finfo(float).eps * steps * p.dt * float(fscale.sum()) + 64.0 * np.finfo(float).eps * float(
        np.abs(m[:, None] * st["v"]).sum())
    typical = float(fscale.sum()) / max(npairs, 1) * p.dt  # one contact's impulse
    assert typical > 1e3 * bound
    assert np.abs(dp).max() <= bound, (dp, bound)


def test_ensemble_256_instances():
    p, counts, mats, x, r, m, meshes = scenes.batch_ensemble_workload()
    rng = np.random.default_rng(31)
    v = rng.normal(0.0, 0.2, x.shape)
    n = x.shape[0]
    print(f"case fullsize batch instances={len(counts)} n={n}", flush=True)
    e = native.Engine(copy_params(p))
    e.upload_batch(counts, mats, x, r, m, v=v)
    for v9, n3 in meshes:
        e.add_mesh(v9, n3)
    e.step(1)
    got = e.download()
    offs = np.concatenate([[0], np.cumsum(counts)])
    pick = sorted({int(np.argmin(counts)), int(np.argmax(counts)), *rng.choice(len(counts), 2, replace=False).tolist()})
    for k in pick:
        q = copy_params(p)
        q.particle = mats[k]
        sl = slice(int(offs[k]), int(offs[k + 1]))
        o = _oracle(q, x[sl], r[sl], m[sl], v=v[sl], meshes=meshes)
        o.step(1)
        so = o.state()
        for f in FIELDS:
            assert np.array_equal(so[f], got[f][sl]), (k, f)
    assert e.stats()["contacts"] > 0
    e.step(30)
    assert list(e.instance_counts()) == list(counts)
    _finite(e.download())


def test_drum_16m_pairs_sampled():
    from scipy.spatial import cKDTree

    p, x, r, m, meshes = scenes.drum_workload(16_000_000)
    n = x.shape[0]
    print(f"case fullsize drum n={n}", flush=True)
    e = _engine(p, x, r, m, meshes=meshes, solver="shortcut")
    e.set_mesh_motion(0, scenes.DRUM_ORIGIN, scenes.DRUM_OMEGA)
    e.step(20)
    e.set_record(True)
    e.step(1)
    offs, js = e.pairs()
    st = e.download()
    assert e.count() == n
    _finite(st)
    xs = st["x"]
    rad = st["radius"]
    tree = cKDTree(xs)
    rng = np.random.default_rng(17)
    sample = rng.choice(n, 3000, replace=False)
    rmax = float(rad.max())
    nonempty = 0
    for i in sample:
        cand = np.array(tree.query_ball_point(xs[i], 1.5 * (rad[i] + rmax) * (1.0 + 1e-9)), dtype=np.int64)
        cand = cand[cand != i]
        d = xs[i] - xs[cand]
        dsq = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        rsum = rad[i] + rad[cand]
        cut = rsum + rsum * 0.5
        want = np.sort(cand[dsq < cut * cut])
        have = np.sort(js[offs[i]:offs[i + 1]])
        assert np.array_equal(want, have), i
        nonempty += want.size > 0
    assert nonempty > 1000
