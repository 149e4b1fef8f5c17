"""Shared scene builders and comparison helpers for the parity tests."""
from __future__ import annotations

import numpy as np

from rusty_particles_amd import abi, scenes


def copy_params(p: abi.RpParams) -> abi.RpParams:
    return abi.RpParams.from_buffer_copy(p)


def lattice(counts, spacing, origin, jitter, seed):
    return scenes.jittered_lattice(counts, spacing, origin, jitter, seed)


def dense_scene(counts=(10, 10, 10), spacing=0.0098, r=(0.0049, 0.0051), jitter=2e-4, seed=1, vel=0.05,
                omega=5.0, rho=2500.0, centre=(0.0, 0.0, 0.0), fixed_frac=0.0):
    """Jittered lattice whose spacing is below 2r: ~6 overlapping neighbours each,
    random linear and angular velocities so every force term is exercised."""
    rng = np.random.default_rng(seed)
    ext = (np.array(counts) - 1) * spacing
    origin = np.asarray(centre) - 0.5 * ext
    x = scenes.jittered_lattice(counts, spacing, origin, jitter, seed)
    n = x.shape[0]
    rad = rng.uniform(r[0], r[1], n)
    m = scenes.sphere_mass(rad, rho)
    v = rng.normal(0.0, vel, (n, 3))
    w = rng.normal(0.0, omega, (n, 3))
    fixed = (rng.uniform(size=n) < fixed_frac).astype(np.uint8)
    return dict(x=x, radius=rad, mass=m, v=v, w=w, fixed=fixed)


def params(dt, bmin, bmax, normal=abi.RP_NORMAL_HERTZ, tangential=abi.RP_TANGENTIAL_COULOMB, E=1e7, nu=0.3,
           mu=0.5, e=0.5, gamma=0.0, wall=None, periodic=(0, 0, 0)):
    p = abi.default_params(dt, bmin, bmax)
    p.normal_model = normal
    p.tangential_model = tangential
    p.particle = abi.RpMaterial(E, nu, 2500.0, mu, e, gamma)
    if wall is not None:
        p.wall = abi.RpMaterial(*wall)
    for k in range(3):
        p.periodic[k] = periodic[k]
    return p


def run_oracle(p, sc, meshes=(), steps=1, bits=64, record=True):
    from oracle import OracleSimulation
    o = OracleSimulation(p.dt, None, None, bits=bits, params=copy_params(p))
    o.set_state(sc["x"], sc["radius"], sc["mass"], sc.get("fixed"), v=sc.get("v"), w=sc.get("w"),
                a=sc.get("a"), tau=sc.get("tau"))
    for v9, n3 in meshes:
        o.add_mesh_arrays(v9, n3)
    o.set_record(record)
    o.step(steps)
    return o


def run_engine(p, sc, meshes=(), steps=1, record=True, chunks=None, solver=None):
    from rusty_particles_amd import native
    e = native.Engine(copy_params(p))
    if solver:
        e.set_jkr_solver(solver)
    e.upload(sc["x"], sc["radius"], sc["mass"], sc.get("fixed"), v=sc.get("v"), w=sc.get("w"), a=sc.get("a"),
             tau=sc.get("tau"))
    for v9, n3 in meshes:
        e.add_mesh(v9, n3)
    e.set_record(record)
    if chunks:
        for c in chunks:
            e.step(c)
    elif steps:
        e.step(steps)
    return e


def assert_pairs_equal(o, e):
    oo, oj = o.pairs()
    eo, ej = e.pairs()
    assert oo.shape == eo.shape, (oo.shape, eo.shape)
    assert np.array_equal(oo, eo), "pair offsets differ"
    assert np.array_equal(oj, ej), "pair lists differ"
    return oj.size


def assert_forces_close(o, e, p, rtol=1e-10):
    """a and tau after one step: |da| <= rtol*(sum|F_c|/m + |g|), |dtau| <= rtol*r*sum|F_c|.
    sum|F_c| is the oracle's per-particle sum of contact-force magnitudes, the
    scale that summation-order rounding is relative to (the net force of a
    near-static packing can be ~0)."""
    st_o = o.state()
    st_e = e.download()
    _, _, fscale = o.forces()
    g = np.linalg.norm(np.array(p.gravity[:]))
    m = st_e["mass"]
    r = st_e["radius"]
    da = np.abs(st_o["a"] - st_e["a"]).max(1)
    lim_a = rtol * (fscale / m + g) + 1e-300
    dt_ = np.abs(st_o["tau"] - st_e["tau"]).max(1)
    lim_t = rtol * r * fscale + 1e-300
    bad_a = np.nonzero(~(da <= lim_a))[0]
    bad_t = np.nonzero(~(dt_ <= lim_t))[0]
    assert bad_a.size == 0, f"a differs at {bad_a[:5]}: {da[bad_a[:5]]} > {lim_a[bad_a[:5]]}"
    assert bad_t.size == 0, f"tau differs at {bad_t[:5]}: {dt_[bad_t[:5]]} > {lim_t[bad_t[:5]]}"
    return float((da / np.maximum(fscale / m + g, 1e-300)).max())


def assert_state_close(o, e, xtol, vtol, wtol):
    st_o = o.state()
    st_e = e.download()
    assert st_o["x"].shape == st_e["x"].shape
    dx = np.abs(st_o["x"] - st_e["x"]).max()
    dv = np.abs(st_o["v"] - st_e["v"]).max()
    dw = np.abs(st_o["w"] - st_e["w"]).max()
    assert dx <= xtol, f"max |dx| {dx} > {xtol}"
    assert dv <= vtol, f"max |dv| {dv} > {vtol}"
    assert dw <= wtol, f"max |dw| {dw} > {wtol}"
    return dx, dv, dw


def jkr_dimers(d_values, r0=50e-6, E=1e8, nu=0.3, gamma=0.05, rho=1500.0, seed=5, rspread=0.2, vel=0.0):
    """Isolated particle pairs, pair k at scaled overlap d_values[k]
    (delta = d t with s = (R*^2 8 pi gamma / E*)^(1/3), t = s^2 / R*), on a
    lattice 12 r0 apart so no pair sees another (JKR cutoff 1.5 r_sum)."""
    rng = np.random.default_rng(seed)
    d_values = np.asarray(d_values, float)
    npair = d_values.size
    side = int(np.ceil(npair ** (1 / 3)))
    idx = np.arange(npair)
    centres = np.stack([idx % side, (idx // side) % side, idx // (side * side)], 1) * (12 * r0)
    centres -= centres.mean(0)
    ra = r0 * rng.uniform(1 - rspread, 1 + rspread, npair)
    rb = r0 * rng.uniform(1 - rspread, 1 + rspread, npair)
    estar = 1.0 / (2 * (1 - nu * nu) / E)
    rstar = ra * rb / (ra + rb)
    s = (rstar * rstar * 8 * np.pi * gamma / estar) ** (1 / 3)
    t = s * s / rstar
    delta = d_values * t
    u = rng.normal(size=(npair, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    sep = (ra + rb - delta)[:, None]
    x = np.empty((2 * npair, 3))
    x[0::2] = centres - 0.5 * sep * u
    x[1::2] = centres + 0.5 * sep * u
    rad = np.empty(2 * npair)
    rad[0::2], rad[1::2] = ra, rb
    m = scenes.sphere_mass(rad, rho)
    v = rng.normal(0.0, vel, (2 * npair, 3))
    return dict(x=x, radius=rad, mass=m, v=v, w=np.zeros((2 * npair, 3)), fixed=np.zeros(2 * npair, np.uint8),
                rstar=rstar, t=t, s=s, d=d_values)


def fine_keys(x, grid):
    """The engine's fine-cell key of positions x from its grid geometry
    (rp_engine_cells): floor((x - origin) * (1 / cell)) per axis, wrapped on
    periodic axes, clamped to the grid on the others, then the cell number
    sum_k ((c >> shift_k) * hi_k + (c & (2^shift_k - 1)) * lo_k)."""
    key = np.zeros(x.shape[0], dtype=np.int64)
    for k in range(3):
        q = np.floor((x[:, k] - grid["origin"][k]) * grid["inv_cell"][k])
        q = np.where(np.isnan(q), 0.0, np.clip(q, -1e9, 1e9)).astype(np.int64)
        n = int(grid["cells"][k])
        c = np.mod(q, n) if grid["periodic"][k] else np.clip(q, 0, n - 1)
        sh = int(grid["shift"][k])
        key += (c >> sh) * int(grid["hi"][k]) + (c & ((1 << sh) - 1)) * int(grid["lo"][k])
    return key


def assert_cells_equal(e, o):
    """Cell ids and the sort permutation (rp_engine_cells after a forced
    rebuild) against the oracle and the grid formulas, bit for bit:
    reference-grid cell index of every particle = the oracle's
    (grid.rs:30-40); fine keys = the host formula on the engine's own
    positions; slot order = a stable sort of the API indices by key."""
    c = e.cells(rebuild=True)
    ref = o.cell_ids()
    assert np.array_equal(c["ref_cell"], ref), "reference-grid cell ids differ"
    x = e.download()["x"]
    key = fine_keys(x, c["grid"])
    assert np.array_equal(c["fine_key"].astype(np.int64), key), "fine-cell keys differ"
    perm = np.lexsort((np.arange(key.size), key))
    assert np.array_equal(c["slot_api"], perm), "sort permutation differs"
    return c
