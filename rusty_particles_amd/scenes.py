"""Seeded synthetic workloads and procedural STL meshes.

Coordinates are in the reference's frame: gravity along -Y
(src/simulation.rs:267).  A workload described z-up maps (x, y, z) -> (x, z, -y).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import struct

import numpy as np

from . import abi

_SCENE_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "librp_scene.so")
_scene = None


def _lib():
    global _scene
    if _scene is None:
        if not os.path.exists(_SCENE_LIB):
            raise ImportError(f"{_SCENE_LIB} missing: run `python -m rusty_particles_amd.build scene`")
        L = C.CDLL(_SCENE_LIB)
        dp = C.POINTER(C.c_double)
        L.rps_rsa.argtypes = [C.c_longlong, dp, dp, C.c_double, C.c_double, C.c_ulonglong, C.c_longlong, dp, dp]
        L.rps_rsa.restype = C.c_longlong
        L.rps_rsa2.argtypes = [C.c_longlong, dp, dp, C.c_double, C.c_double, dp, C.POINTER(C.c_int),
                               C.c_ulonglong, C.c_longlong, dp, dp]
        L.rps_rsa2.restype = C.c_longlong
        L.rps_rsa_cone.argtypes = [C.c_longlong, dp, dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_double, C.c_ulonglong, C.c_longlong, dp, dp]
        L.rps_rsa_cone.restype = C.c_longlong
        L.rps_set_progress.argtypes = [C.c_int]
        L.rps_set_progress.restype = None
        _scene = L
    return _scene


def rsa(n: int, lo, hi, rmin: float, rmax: float, seed: int, max_attempts: int = 100000):
    """Random sequential addition of n non-overlapping spheres, radii U[rmin, rmax]."""
    x = np.zeros((n, 3))
    r = np.zeros(n)
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    hi = np.ascontiguousarray(hi, dtype=np.float64)
    dp = C.POINTER(C.c_double)
    placed = _lib().rps_rsa(n, lo.ctypes.data_as(dp), hi.ctypes.data_as(dp), rmin, rmax, seed, max_attempts,
                            x.ctypes.data_as(dp), r.ctypes.data_as(dp))
    if placed < n:
        raise RuntimeError(f"RSA placed only {placed} of {n} spheres")
    return x, r


def rsa_radii(radii, lo, hi, periodic, seed: int, max_attempts: int = 200000):
    """Random sequential addition of spheres with the given radii (placed in
    the order given), periodic axes wrapped."""
    radii = np.ascontiguousarray(radii, dtype=np.float64)
    n = radii.size
    x = np.zeros((n, 3))
    r = np.zeros(n)
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    hi = np.ascontiguousarray(hi, dtype=np.float64)
    per = (C.c_int * 3)(*[1 if q else 0 for q in periodic])
    dp = C.POINTER(C.c_double)
    _lib().rps_set_progress(1 if n >= 500_000 else 0)
    placed = _lib().rps_rsa2(n, lo.ctypes.data_as(dp), hi.ctypes.data_as(dp), 0.0, float(radii.max()),
                             radii.ctypes.data_as(dp), per, seed, max_attempts, x.ctypes.data_as(dp),
                             r.ctypes.data_as(dp))
    if placed < n:
        raise RuntimeError(f"RSA placed only {placed} of {n} spheres")
    return x, r


def rsa_cone(n: int, rmin: float, rmax: float, y0: float, y1: float, r0: float, half_angle_deg: float,
             clear: float, seed: int, max_attempts: int = 200000):
    """Random sequential addition of n non-overlapping spheres (radii U[rmin,
    rmax]) inside a cone around +Y (radius r0 at y0, half-angle
    half_angle_deg), centres between y0 and y1 and at least clear * r from the
    cone's surface."""
    x = np.zeros((n, 3))
    r = np.zeros(n)
    rtop = r0 + (y1 - y0) * math.tan(math.radians(half_angle_deg))
    lo = np.array([-rtop, y0, -rtop])
    hi = np.array([rtop, y1, rtop])
    dp = C.POINTER(C.c_double)
    placed = _lib().rps_rsa_cone(n, lo.ctypes.data_as(dp), hi.ctypes.data_as(dp), rmin, rmax, y0, y1, r0,
                                 half_angle_deg, clear, seed, max_attempts, x.ctypes.data_as(dp),
                                 r.ctypes.data_as(dp))
    if placed < n:
        raise RuntimeError(f"RSA placed only {placed} of {n} spheres in the cone")
    return x, r


def jittered_lattice(counts, spacing: float, origin, jitter: float, seed: int, n: int | None = None):
    """Cubic lattice (x fastest), uniform jitter in [-jitter, jitter] per axis."""
    nx, ny, nz = counts
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    g = g.transpose(2, 1, 0, 3).reshape(-1, 3).astype(np.float64)
    pts = np.asarray(origin, dtype=np.float64) + g * spacing
    rng = np.random.default_rng(seed)
    pts = pts + rng.uniform(-jitter, jitter, size=pts.shape)
    if n is not None:
        pts = pts[:n]
    return pts


# ---- meshes -------------------------------------------------------------------

def _tri_normals(v9: np.ndarray) -> np.ndarray:
    a = v9[:, 0:3].astype(np.float64)
    b = v9[:, 3:6].astype(np.float64)
    c = v9[:, 6:9].astype(np.float64)
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return n.astype(np.float32)


def box_mesh(lo, hi):
    """12 triangles closing the box [lo, hi], normals pointing inward."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    c = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], dtype=np.float64)
    centre = c.mean(0)
    quads = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (3, 2, 6, 7), (0, 3, 7, 4), (1, 2, 6, 5)]
    tris = []
    for a, b, cc, d in quads:
        for t in ((a, b, cc), (a, cc, d)):
            p = c[list(t)]
            nrm = np.cross(p[1] - p[0], p[2] - p[0])
            if np.dot(nrm, centre - p[0]) < 0:  # make the normal point inward
                p = p[[0, 2, 1]]
            tris.append(p.reshape(9))
    v9 = np.array(tris, dtype=np.float32)
    return v9, _tri_normals(v9)


def hopper_mesh(r_out: float, half_angle_deg: float, r_top: float, y0: float = 0.0, nseg: int = 64,
                nring: int = 32):
    """Conical hopper around +Y: radius r_out at y0 growing to r_top; 2*nseg*nring triangles."""
    h = (r_top - r_out) / math.tan(math.radians(half_angle_deg))
    ys = y0 + h * np.arange(nring + 1) / nring
    rs = r_out + (ys - y0) * math.tan(math.radians(half_angle_deg))
    th = 2 * np.pi * np.arange(nseg + 1) / nseg
    tris = []
    for i in range(nring):
        for j in range(nseg):
            p00 = (rs[i] * np.cos(th[j]), ys[i], rs[i] * np.sin(th[j]))
            p01 = (rs[i] * np.cos(th[j + 1]), ys[i], rs[i] * np.sin(th[j + 1]))
            p10 = (rs[i + 1] * np.cos(th[j]), ys[i + 1], rs[i + 1] * np.sin(th[j]))
            p11 = (rs[i + 1] * np.cos(th[j + 1]), ys[i + 1], rs[i + 1] * np.sin(th[j + 1]))
            tris.append(p00 + p10 + p11)
            tris.append(p00 + p11 + p01)
    v9 = np.array(tris, dtype=np.float32)
    n3 = _tri_normals(v9)
    # inward (towards the axis)
    cen = (v9[:, 0:3] + v9[:, 3:6] + v9[:, 6:9]) / 3.0
    outward = np.stack([cen[:, 0], np.zeros(len(cen)), cen[:, 2]], 1)
    flip = (n3 * outward).sum(1) > 0
    n3[flip] *= -1
    return v9, n3


def drum_mesh(radius: float = 0.25, length: float = 0.2, centre=(0.0, 0.0, 0.0), nseg: int = 128,
              nlen: int = 49, ncap: int = 8):
    """Closed cylindrical drum around an axis parallel to Z (horizontal: gravity
    is -Y), normals pointing inward.  The side has nseg x nlen quads (2 triangles
    each); each end cap has ncap rings, a fan of nseg triangles at the centre and
    2*nseg per outer ring.  Triangles: 2*nseg*(nlen + 2*ncap - 1); the defaults
    give the 16,384 of a 0.25 m x 0.2 m drum.  Side and caps share their rim
    vertices, so every edge belongs to exactly two triangles."""
    cx, cy, cz = centre
    th = 2 * np.pi * np.arange(nseg + 1) / nseg
    th[-1] = 0.0  # the seam closes on the same vertex
    zs = cz - 0.5 * length + length * np.arange(nlen + 1) / nlen
    rim = np.stack([cx + radius * np.cos(th), cy + radius * np.sin(th)], 1)
    tris = []
    for i in range(nlen):
        for j in range(nseg):
            p00 = (*rim[j], zs[i])
            p01 = (*rim[j + 1], zs[i])
            p10 = (*rim[j], zs[i + 1])
            p11 = (*rim[j + 1], zs[i + 1])
            tris.append(p00 + p10 + p11)
            tris.append(p00 + p11 + p01)
    rr = radius * np.arange(ncap + 1) / ncap
    for z in (zs[0], zs[-1]):
        ring = lambda k, j: (cx + rr[k] * np.cos(th[j]), cy + rr[k] * np.sin(th[j]), z)
        for j in range(nseg):
            tris.append((cx, cy, z) + ring(1, j) + ring(1, j + 1))
        for k in range(1, ncap):
            for j in range(nseg):
                tris.append(ring(k, j) + ring(k + 1, j) + ring(k + 1, j + 1))
                tris.append(ring(k, j) + ring(k + 1, j + 1) + ring(k, j + 1))
    v9 = np.array(tris, dtype=np.float32)
    n3 = _tri_normals(v9)
    # inward: towards the axis on the side, towards the middle plane on the caps
    cen = (v9[:, 0:3].astype(np.float64) + v9[:, 3:6] + v9[:, 6:9]) / 3.0
    cap = np.abs(np.abs(cen[:, 2] - cz) - 0.5 * length) < 1e-6 * max(length, radius)
    out = np.where(cap[:, None], np.stack([np.zeros(len(cen)), np.zeros(len(cen)), cen[:, 2] - cz], 1),
                   np.stack([cen[:, 0] - cx, cen[:, 1] - cy, np.zeros(len(cen))], 1))
    flip = (n3 * out).sum(1) > 0
    n3[flip] *= -1
    return v9, n3


def write_stl_binary(path, v9, n3, header=b"rusty_particles_amd"):
    v9 = np.asarray(v9, dtype="<f4").reshape(-1, 9)
    n3 = np.asarray(n3, dtype="<f4").reshape(-1, 3)
    with open(path, "wb") as f:
        f.write(header[:80].ljust(80, b"\0"))
        f.write(struct.pack("<I", v9.shape[0]))
        for t in range(v9.shape[0]):
            f.write(n3[t].tobytes())
            f.write(v9[t].tobytes())
            f.write(b"\0\0")


def write_stl_ascii(path, v9, n3, name="mesh"):
    v9 = np.asarray(v9, dtype=np.float32).reshape(-1, 9)
    n3 = np.asarray(n3, dtype=np.float32).reshape(-1, 3)
    fmt = lambda v: np.format_float_scientific(np.float32(v), unique=True)  # shortest round-trip form
    with open(path, "w") as f:
        f.write(f"solid {name}\n")
        for t in range(v9.shape[0]):
            f.write(f"  facet normal {fmt(n3[t,0])} {fmt(n3[t,1])} {fmt(n3[t,2])}\n    outer loop\n")
            for k in range(3):
                f.write(f"      vertex {fmt(v9[t,3*k])} {fmt(v9[t,3*k+1])} {fmt(v9[t,3*k+2])}\n")
            f.write("    endloop\n  endfacet\n")
        f.write(f"endsolid {name}\n")


# ---- workloads -------------------------------------------------------------------

def sphere_mass(r, rho):
    return rho * 4.0 / 3.0 * np.pi * np.asarray(r) ** 3


def hopper_workload(n: int = 1_000_000, seed: int = 0xD3E50002, phi: float = 0.25):
    """1M polydisperse spheres (r ~ U[0.5, 1.5] mm, rho 2500) above a conical hopper
    (half-angle 30 deg, outlet radius 8 mm, 4096 triangles), Hertz normal +
    Mindlin tangential, E=5e7, nu=0.25, e=0.6, mu=0.4, dt=2e-6.

    Seeded sequential rejection fills a 0.1 x 0.1 m column above the hopper to
    solid fraction `phi` (the column height grows with n; 0.1 x 0.1 x 0.2 m
    cannot hold 1M such spheres).  Returns (params, x, r, m, meshes)."""
    rmin, rmax = 0.5e-3, 1.5e-3
    mean_r3 = (rmax ** 4 - rmin ** 4) / (4 * (rmax - rmin))
    vol = n * 4.0 / 3.0 * np.pi * mean_r3
    height = vol / phi / (0.1 * 0.1)
    hv9, hn3 = hopper_mesh(0.008, 30.0, 0.075, y0=0.0, nseg=64, nring=32)
    y_top = float(hv9[:, 1::3].max())
    lo = (-0.05, y_top + 0.005, -0.05)
    hi = (0.05, y_top + 0.005 + height, 0.05)
    x, r = rsa(n, lo, hi, rmin, rmax, seed)
    m = sphere_mass(r, 2500.0)
    p = abi.default_params(2e-6, (-0.1, -0.05, -0.1), (0.1, hi[1] + 0.01, 0.1))
    p.normal_model = abi.RP_NORMAL_HERTZ
    p.tangential_model = abi.RP_TANGENTIAL_MINDLIN
    p.particle = abi.RpMaterial(5e7, 0.25, 2500.0, 0.4, 0.6, 0.0)
    return p, x, r, m, [(hv9, hn3)]


HOPPER_OUTLET, HOPPER_HALF_ANGLE, HOPPER_TOP = 0.008, 30.0, 0.075


def hopper_test_workload(n: int = 20_000, seed: int = 0xD3E50012, clear: float = 0.97, inflate: float = 1.04):
    """The hopper configuration's spheres, materials and 4,096-triangle cone
    (hopper_workload) with a bed of n spheres inside the cone instead of the
    column above it: random sequential addition in the cone from its outlet
    up, centres 0.97 r from the wall (the spheres near the wall press on
    it), then radii grown by 4 % so that near neighbours overlap; spheres
    near the outlet fall through it.  Sized for the oracle (20k spheres, a
    bed ~7 cm deep).  Returns (params, x, r, m, meshes)."""
    rmin, rmax = 0.5e-3 / inflate, 1.5e-3 / inflate
    hv9, hn3 = hopper_mesh(HOPPER_OUTLET, HOPPER_HALF_ANGLE, HOPPER_TOP, y0=0.0, nseg=64, nring=32)
    # bed height for solid fraction ~0.2 in the cone
    mean_vol = 4.0 / 3.0 * np.pi * (rmax ** 4 - rmin ** 4) / (4 * (rmax - rmin))
    t = math.tan(math.radians(HOPPER_HALF_ANGLE))
    need = n * mean_vol / 0.2
    h = 0.005
    while np.pi * h / 3.0 * (HOPPER_OUTLET ** 2 + HOPPER_OUTLET * (HOPPER_OUTLET + h * t)
                             + (HOPPER_OUTLET + h * t) ** 2) < need:
        h += 0.001
    x, r = rsa_cone(n, rmin, rmax, 0.0, h, HOPPER_OUTLET, HOPPER_HALF_ANGLE, clear, seed)
    r = r * inflate
    m = sphere_mass(r, 2500.0)
    p = abi.default_params(2e-6, (-0.1, -0.05, -0.1), (0.1, 0.14, 0.1))
    p.normal_model = abi.RP_NORMAL_HERTZ
    p.tangential_model = abi.RP_TANGENTIAL_MINDLIN
    p.particle = abi.RpMaterial(5e7, 0.25, 2500.0, 0.4, 0.6, 0.0)
    return p, x, r, m, [(hv9, hn3)]


def _hertz_kn(e_star: float, r_star):
    """Hertz stiffness 4/3 E* sqrt(R*) (F = kn delta^1.5, physics.rs:86-129)."""
    return 4.0 / 3.0 * e_star * np.sqrt(r_star)


def _e_star(e1: float, nu1: float, e2: float, nu2: float) -> float:
    """Effective modulus 1 / ((1 - nu1^2)/E1 + (1 - nu2^2)/E2) (material.rs:25-37)."""
    return 1.0 / ((1.0 - nu1 * nu1) / e1 + (1.0 - nu2 * nu2) / e2)


def hopper_bed_workload(n: int = 1_000_000, seed: int = 0xD3E50022, top: float = 0.3):
    """The hopper configuration's n spheres (r ~ U[0.5, 1.5] mm, rho 2500,
    Hertz + Mindlin, E=5e7, nu=0.25, e=0.6, mu=0.4, dt=2e-6 s, gravity -Y)
    resting in a cone of the same half-angle (30 deg), outlet (8 mm) and
    triangle count (4,096) whose top radius is `top` (the 75 mm cone of
    hopper_workload holds ~36k such spheres): a contact-bearing window of the
    same workload.

    The bed is a set of vertical columns on a square grid of pitch 2.2 r_max
    (one radius per column), each standing on the cone facet below it: the
    lowest sphere touches the facet's plane (the plane through the apex and
    the facet's two generators), pressed into it by the Hertz overlap that
    carries twice the column's weight (the facet's normal has a vertical
    component of 1/2), and each sphere above rests on the one below with the
    static Hertz overlap of the weight above it.  The lowest sphere of every
    column is fixed (the reference's `fixed` flag, particle.rs:10): on the
    slanted facet a free one slides from under its column, which then falls
    apart (frictionless spheres in the parity mode), while on a fixed foot the
    column stays in equilibrium.  Columns whose lowest contact would lie in
    the outlet are left out.  The columns fill the cone from the bottom up to
    a common height that gives n spheres.  Every sphere but the top of its
    column has two contacts, and every foot touches the cone.  Returns
    (params, x, r, m, meshes, fixed)."""
    rng = np.random.default_rng(seed)
    rmin, rmax = 0.5e-3, 1.5e-3
    e_mod, nu, rho, grav = 5e7, 0.25, 2500.0, 9.81
    wall = abi.default_params(1.0, (0, 0, 0), (1, 1, 1)).wall
    nseg = 64
    hv9, hn3 = hopper_mesh(HOPPER_OUTLET, HOPPER_HALF_ANGLE, top, y0=0.0, nseg=nseg, nring=32)
    y_top = float(hv9[:, 1::3].max())
    half = math.radians(HOPPER_HALF_ANGLE)
    apex_y = -HOPPER_OUTLET / math.tan(half)
    pitch = 2.2 * rmax
    g = np.arange(-top, top + 1e-12, pitch) + 0.5 * pitch
    X, Z = np.meshgrid(g, g, indexing="ij")
    cx, cz = X.ravel(), Z.ravel()
    rc = rng.uniform(rmin, rmax, cx.size)
    mc = sphere_mass(rc, rho)
    # inward unit normals of the facet planes of segments j - 1, j, j + 1
    th = np.mod(np.arctan2(cz, cx), 2.0 * math.pi)
    seg = np.floor(th / (2.0 * math.pi / nseg)).astype(np.int64)

    def plane(jj):
        t0, t1 = 2.0 * math.pi * jj / nseg, 2.0 * math.pi * (jj + 1) / nseg
        s, c = math.sin(half), math.cos(half)
        g0 = np.stack([s * np.cos(t0), np.full(t0.shape, c), s * np.sin(t0)], 1)
        g1 = np.stack([s * np.cos(t1), np.full(t1.shape, c), s * np.sin(t1)], 1)
        nn = -np.cross(g0, g1)
        return nn / np.linalg.norm(nn, axis=1, keepdims=True)

    e_pp = _e_star(e_mod, nu, e_mod, nu)
    e_pw = _e_star(e_mod, nu, wall.youngs_modulus, wall.poissons_ratio)
    kn_pp = _hertz_kn(e_pp, 0.5 * rc)
    kn_pw = _hertz_kn(e_pw, rc)

    def base(count):
        """Lowest sphere's centre height and its foot's height on the cone."""
        w = np.maximum(count, 1) * mc * grav
        press = (2.0 * w / kn_pw) ** (2.0 / 3.0)
        yb = np.full(cx.size, -np.inf)
        foot = np.zeros(cx.size)
        for dj in (-1, 0, 1):
            nn = plane(seg + dj)
            y = apex_y + (rc - press - nn[:, 0] * cx - nn[:, 2] * cz) / nn[:, 1]
            better = y > yb
            yb = np.where(better, y, yb)
            foot = np.where(better, y - (rc - press) * nn[:, 1], foot)
        return yb, foot

    yb, foot = base(np.ones(cx.size))
    ok = (foot > 0.0) & (np.hypot(cx, cz) + rc < top)
    cx, cz, rc, mc, seg, kn_pp, kn_pw = cx[ok], cz[ok], rc[ok], mc[ok], seg[ok], kn_pp[ok], kn_pw[ok]
    yb = yb[ok]

    def counts(h):
        return np.maximum(np.floor((h - rc - yb) / (2.0 * rc)).astype(np.int64) + 1, 0)

    lo, hi = 0.0, y_top - 2.0 * rmax
    if counts(hi).sum() < n:
        raise RuntimeError(f"a bed of {n} spheres does not fit a cone of top radius {top}")
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if counts(mid).sum() < n else (lo, mid)
    per = counts(hi)
    # drop the highest spheres beyond n (column tops, highest first)
    extra = int(per.sum()) - n
    if extra > 0:
        tops = yb + 2.0 * rc * (per - 1)
        for q in np.argsort(-tops)[:extra]:
            per[q] -= 1
    keep = per > 0
    cx, cz, rc, mc, seg, kn_pp, kn_pw, per = (cx[keep], cz[keep], rc[keep], mc[keep], seg[keep], kn_pp[keep],
                                              kn_pw[keep], per[keep])
    yb, _ = base(per)
    cid = np.repeat(np.arange(per.size), per)
    start = np.concatenate([[0], np.cumsum(per)[:-1]])
    k = np.arange(cid.size) - start[cid]                 # 0 = lowest sphere of its column
    above = per[cid] - 1 - k                             # spheres resting on sphere k
    # contact k (between sphere k and k + 1) carries the weight of the spheres above sphere k
    delta = (np.maximum(above, 0) * mc[cid] * grav / kn_pp[cid]) ** (2.0 / 3.0)
    step = 2.0 * rc[cid] - delta
    cs = np.cumsum(step)
    y = yb[cid] + (cs - step) - np.repeat((cs - step)[start], per)
    x = np.stack([cx[cid], y, cz[cid]], 1)
    r = rc[cid]
    m = mc[cid]
    fixed = (k == 0).astype(np.uint8)
    p = abi.default_params(2e-6, (-top - 0.01, -0.05, -top - 0.01), (top + 0.01, y_top + 0.01, top + 0.01))
    p.normal_model = abi.RP_NORMAL_HERTZ
    p.tangential_model = abi.RP_TANGENTIAL_MINDLIN
    p.particle = abi.RpMaterial(e_mod, nu, rho, 0.4, 0.6, 0.0)
    return p, x, r, m, [(hv9, hn3)], fixed


def box_settle_workload(n: int = 10_000, seed: int = 0xD3E50001):
    """10,000 spheres r=5 mm, rho 2500, jittered lattice (spacing 11 mm, +-0.4 mm)
    in the box [-0.15, 0.15]^3 closed by a 12-triangle STL box, Hertz + Coulomb,
    E=1e7, nu=0.3, e=0.5, mu=0.5, dt=1e-5."""
    counts = (22, 22, 21)
    spacing = 0.011
    ext = np.array(counts) - 1
    origin = -0.5 * spacing * ext
    x = jittered_lattice(counts, spacing, origin, 0.0004, seed, n)
    r = np.full(x.shape[0], 0.005)
    m = sphere_mass(r, 2500.0)
    p = abi.default_params(1e-5, (-0.15, -0.15, -0.15), (0.15, 0.15, 0.15))
    p.normal_model = abi.RP_NORMAL_HERTZ
    p.tangential_model = abi.RP_TANGENTIAL_COULOMB
    p.particle = abi.RpMaterial(1e7, 0.3, 2500.0, 0.5, 0.5, 0.0)
    v9, n3 = box_mesh((-0.15, -0.15, -0.15), (0.15, 0.15, 0.15))
    return p, x, r, m, [(v9, n3)]


def jkr_periodic_workload(n: int = 2_000_000, seed: int = 0xD3E50003, phi: float = 0.35):
    """2,000,000 spheres, radius lognormal (median 50 um, sigma 0.3), rho 1500,
    E=1e8, nu=0.3, surface energy 0.05 J/m^2, Coulomb mu=0.3, JKR normal;
    random loose packing at solid fraction `phi` in a periodic cube,
    Gaussian velocities (std 0.05 m/s per component), dt=1e-8 s.  Radii are
    placed largest first (random sequential addition)."""
    rng = np.random.default_rng(seed)
    radii = 50e-6 * np.exp(0.3 * rng.standard_normal(n))
    vol = 4.0 / 3.0 * np.pi * float(np.sum(radii ** 3))
    L = (vol / phi) ** (1.0 / 3.0)
    order = np.argsort(-radii, kind="stable")
    x, r = rsa_radii(radii[order], (0.0, 0.0, 0.0), (L, L, L), (1, 1, 1), seed)
    perm = rng.permutation(n)  # API order independent of the placement order
    x, r = x[perm], r[perm]
    m = sphere_mass(r, 1500.0)
    v = rng.normal(0.0, 0.05, (n, 3))
    p = abi.default_params(1e-8, (0.0, 0.0, 0.0), (L, L, L))
    for k in range(3):
        p.periodic[k] = 1
    p.normal_model = abi.RP_NORMAL_JKR
    p.tangential_model = abi.RP_TANGENTIAL_COULOMB
    p.particle = abi.RpMaterial(1e8, 0.3, 1500.0, 0.3, 0.5, 0.05)
    return p, x, r, m, v


def jkr_dense_workload(n: int = 2_000_000, seed: int = 0xD3E50007, spread: float = 0.02, gap: float = -1e-6):
    """A compressed adhesive crystal: the JKR material of the 2M-particle
    packing (E=1e8, nu=0.3, gamma 0.05, rho 1500, Coulomb mu=0.3) with radii
    50 um x U[1 - spread, 1 + spread] on a periodic FCC lattice whose
    nearest-neighbour distance is 100 um + `gap` (default: 1 um closer), so
    each sphere has 12 adhesive contacts, most of them with a JKR root
    (overlaps of -1 um ... +3 um, i.e. d = delta / t of about -4 ... 13), and
    6 second neighbours inside the 1.5 r_sum cut-off without one; jitter
    +-0.2 um, Gaussian velocities (std 0.05 m/s per component), dt=1e-8 s."""
    rng = np.random.default_rng(seed)
    m_cells = max(1, int(round((n / 4.0) ** (1.0 / 3.0))))
    a = 100e-6 + gap
    side = a * np.sqrt(2.0)  # cubic FCC cell
    base = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    g = np.arange(m_cells)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(float)
    x = ((cells[:, None, :] + base[None, :, :]).reshape(-1, 3) + 0.25) * side
    x += rng.uniform(-0.2e-6, 0.2e-6, x.shape)
    L = m_cells * side
    nn = x.shape[0]
    perm = rng.permutation(nn)
    x = x[perm]
    r = 50e-6 * rng.uniform(1.0 - spread, 1.0 + spread, nn)
    m = sphere_mass(r, 1500.0)
    v = rng.normal(0.0, 0.05, (nn, 3))
    p = abi.default_params(1e-8, (0.0, 0.0, 0.0), (L, L, L))
    for k in range(3):
        p.periodic[k] = 1
    p.normal_model = abi.RP_NORMAL_JKR
    p.tangential_model = abi.RP_TANGENTIAL_COULOMB
    p.particle = abi.RpMaterial(1e8, 0.3, 1500.0, 0.3, 0.5, 0.05)
    return p, x, r, m, v


# The capacity run's drum turns at 30 rpm about its axis (Z through the origin).
DRUM_ORIGIN = (0.0, 0.0, 0.0)
DRUM_OMEGA = (0.0, 0.0, math.pi)


def drum_workload(n: int = 16_000_000, seed: int = 0xD3E50005, fill: float = 0.30, radius: float = 0.25):
    """Capacity run: n spheres (r ~ U[0.4, 0.8] mm, rho 2500) in the lower part
    of a closed drum (radius 0.25 m, axis along Z, 16,384-triangle STL), JKR
    normal (surface energy 0.01 J/m^2) + Coulomb (mu 0.5), E=5e7, nu=0.3,
    e=0.5, dt=2e-6 s, gravity -Y; the drum's surface turns at 30 rpm
    (DRUM_OMEGA about DRUM_ORIGIN, applied with set_mesh_motion on mesh 0).

    The bed starts as a jittered cubic lattice (spacing 2.02 r_max, jitter
    +-4 um, so no two spheres touch), filled layer by layer from the bottom
    of the drum, at least r_max + half a spacing from the wall.  The lattice
    holds solid fraction ~0.24, so a 0.2 m drum cannot take 16M spheres at
    a 30 % fill; the drum's length is set so that the bed's lattice volume is
    `fill` of the drum's (1.15 m at 16M).  Returns (params, x, r, m, meshes)."""
    rng = np.random.default_rng(seed)
    rmin, rmax = 0.4e-3, 0.8e-3
    s = 2.02 * rmax
    length = n * s ** 3 / (fill * np.pi * radius * radius)
    margin = rmax + 0.5 * s
    xs = np.arange(-radius + margin, radius - margin + 1e-12, s)
    zs = np.arange(-0.5 * length + margin, 0.5 * length - margin + 1e-12, s)
    pieces, count = [], 0
    y = -radius + margin
    while count < n:
        if y > radius - margin:
            raise RuntimeError(f"drum holds only {count} of {n} lattice sites")
        xk = xs[xs * xs + y * y <= (radius - margin) ** 2]
        if xk.size:
            X, Z = np.meshgrid(xk, zs, indexing="xy")
            layer = np.stack([X.ravel(), np.full(X.size, y), Z.ravel()], 1)
            take = min(layer.shape[0], n - count)
            pieces.append(layer[:take])
            count += take
        y += s
    x = np.concatenate(pieces)
    x += rng.uniform(-4e-6, 4e-6, x.shape)
    r = rng.uniform(rmin, rmax, n)
    m = sphere_mass(r, 2500.0)
    v9, n3 = drum_mesh(radius, length)
    pad = 0.005
    p = abi.default_params(2e-6, (-radius - pad, -radius - pad, -0.5 * length - pad),
                           (radius + pad, radius + pad, 0.5 * length + pad))
    p.normal_model = abi.RP_NORMAL_JKR
    p.tangential_model = abi.RP_TANGENTIAL_COULOMB
    p.particle = abi.RpMaterial(5e7, 0.3, 2500.0, 0.5, 0.5, 0.01)
    return p, x, r, m, [(v9, n3)]


def batch_ensemble_workload(ninst: int = 256, seed: int = 0xD3E50004, nmin: int = 2000, nmax: int = 50000,
                            spacing: float = 0.0065):
    """Ensemble of `ninst` independent simulations for a material sweep:
    N_k ~ U{nmin..nmax} spheres, r ~ U[1, 3] mm, rho 2500, per-instance
    friction mu ~ U[0.1, 0.8], restitution e ~ U[0.2, 0.9], surface energy
    gamma ~ U[0, 0.02] J/m^2 (E = 1e7, nu = 0.3), simplified JKR + Coulomb,
    dt = 5e-6 s.  Each instance fills a jittered lattice (spacing 6.5 mm,
    jitter +-0.2 mm, so no initial overlap) from the floor of a 0.2 m square
    column upward -- a 0.2 m cube holds only ~25k such sites, so the column is
    0.45 m tall --
    closed by a 12-triangle box STL.  Returns (params, counts, materials, x,
    r, m, meshes)."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(nmin, nmax + 1, ninst)
    side = 0.2
    per_axis = int(side / spacing) - 1            # 31 sites across, 0.2 m
    layers = int(np.ceil(nmax / per_axis ** 2))   # tallest instance
    origin = np.array([0.5 * (side - (per_axis - 1) * spacing), 0.005, 0.5 * (side - (per_axis - 1) * spacing)])
    xs, rs, mats = [], [], []
    for k in range(ninst):
        n = int(counts[k])
        # sites layer by layer (y, the vertical axis, slowest)
        g = np.stack(np.meshgrid(np.arange(per_axis), np.arange(per_axis), np.arange(layers), indexing="ij"), -1)
        sites = g.transpose(2, 1, 0, 3).reshape(-1, 3)[:n]  # (x, z, y) with y slowest
        jit = np.random.default_rng(seed + 17 * k).uniform(-0.0002, 0.0002, (n, 3))
        x = origin + np.stack([sites[:, 0], sites[:, 2], sites[:, 1]], 1) * spacing + jit
        r = np.random.default_rng(seed + 17 * k + 1).uniform(0.001, 0.003, n)
        xs.append(x)
        rs.append(r)
        mats.append(abi.RpMaterial(1e7, 0.3, 2500.0, float(rng.uniform(0.1, 0.8)), float(rng.uniform(0.2, 0.9)),
                                   float(rng.uniform(0.0, 0.02))))
    x = np.concatenate(xs)
    r = np.concatenate(rs)
    m = sphere_mass(r, 2500.0)
    top = max(0.45, float(x[:, 1].max()) + 0.01)
    p = abi.default_params(5e-6, (-0.01, -0.01, -0.01), (side + 0.01, top + 0.01, side + 0.01))
    p.normal_model = abi.RP_NORMAL_SJKR
    p.tangential_model = abi.RP_TANGENTIAL_COULOMB
    p.particle = abi.RpMaterial(1e7, 0.3, 2500.0, 0.5, 0.5, 0.0)
    v9, n3 = box_mesh((0.0, 0.0, 0.0), (side, top, side))
    return p, [int(c) for c in counts], mats, x, r, m, [(v9, n3)]


def jkr_ensemble_workload(ninst: int = 64, seed: int = 0xD3E50006, nmin: int = 5000, nmax: int = 40000,
                          phi_max: float = 0.35):
    """Ensemble of `ninst` periodic JKR packings for a surface-energy sweep:
    N_k ~ U{nmin..nmax} spheres (r lognormal, median 50 um, sigma 0.3,
    rho 1500), gamma_k ~ U[0.02, 0.08] J/m^2 (E = 1e8, nu = 0.3, mu = 0.3),
    JKR + Coulomb, dt = 1e-8 s, velocities ~N(0, 0.05) m/s.  All instances
    share one periodic cube, sized so that the largest N_k fills it to
    `phi_max` (smaller instances are looser); each is placed by random
    sequential addition, largest spheres first.  Returns (params, counts,
    materials, x, r, m, v)."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(nmin, nmax + 1, ninst)
    radii = [50e-6 * np.exp(0.3 * rng.standard_normal(int(c))) for c in counts]
    vmax = max(4.0 / 3.0 * np.pi * float(np.sum(r ** 3)) for r in radii)
    L = (vmax / phi_max) ** (1.0 / 3.0)
    xs, rs, vs, mats = [], [], [], []
    for k, rad in enumerate(radii):
        order = np.argsort(-rad, kind="stable")
        x, r = rsa_radii(rad[order], (0.0, 0.0, 0.0), (L, L, L), (1, 1, 1), seed + 31 * k)
        perm = np.random.default_rng(seed + 31 * k + 1).permutation(r.size)
        xs.append(x[perm])
        rs.append(r[perm])
        vs.append(np.random.default_rng(seed + 31 * k + 2).normal(0.0, 0.05, (r.size, 3)))
        mats.append(abi.RpMaterial(1e8, 0.3, 1500.0, 0.3, 0.5, float(rng.uniform(0.02, 0.08))))
    x, r, v = np.concatenate(xs), np.concatenate(rs), np.concatenate(vs)
    m = sphere_mass(r, 1500.0)
    p = abi.default_params(1e-8, (0.0, 0.0, 0.0), (L, L, L))
    for k in range(3):
        p.periodic[k] = 1
    p.normal_model = abi.RP_NORMAL_JKR
    p.tangential_model = abi.RP_TANGENTIAL_COULOMB
    p.particle = abi.RpMaterial(1e8, 0.3, 1500.0, 0.3, 0.5, 0.05)
    return p, [int(c) for c in counts], mats, x, r, m, v


def drum_chain_workload(n: int = 16_000_000, seed: int = 0xD3E50025, radius: float = 0.25, fill: float = 0.30,
                        press: float = 2e-6):
    """The capacity run with its bed in contact, for the physics-fix mode's
    Mindlin history: n spheres (r ~ U[0.4, 0.8] mm, rho 2500, JKR gamma 0.01,
    E=5e7, nu=0.3, e=0.5, mu=0.5, dt=2e-6 s) in the 16,384-triangle drum of
    radius `radius` turning at 30 rpm (DRUM_OMEGA, mesh 0), arranged as in
    drum_test_workload: chains along the drum's axis, one radius per chain,
    neighbours touching with overlaps U[0.05, 3] um (JKR contacts with a
    root), chains 2.2 r_max apart, the outer ones pressed `press` into the
    plane of the side facet they face (2 um: a Hertz wall contact that
    accelerates a 0.4 mm sphere by ~7e3 m/s^2; 15 um would launch it off the
    wall at ~1 m/s).  The chains fill the lowest `fill` of the drum's cross-section; the
    drum's length is set so that they hold n spheres.  Returns (params, x, r,
    m, meshes)."""
    rng = np.random.default_rng(seed)
    rmin, rmax = 0.4e-3, 0.8e-3
    pitch = 2.2 * rmax
    g = np.arange(-radius, radius + 1e-12, pitch)
    X, Y = np.meshgrid(g, g, indexing="ij")
    cx, cy = X.ravel(), Y.ravel()
    order = np.lexsort((cx, cy))  # lowest chains first
    cx, cy = cx[order], cy[order]
    rc = rng.uniform(rmin, rmax, cx.size)
    rho = np.hypot(cx, cy)
    keep = rho <= radius - rc + press
    cx, cy, rc, rho = cx[keep], cy[keep], rc[keep], rho[keep]
    # the chains below the height that bounds `fill` of the disc's area
    seg = np.cumsum(np.full(cx.size, pitch * pitch))
    nch = max(1, int(np.searchsorted(seg, fill * math.pi * radius * radius)))
    cx, cy, rc, rho = cx[:nch], cy[:nch], rc[:nch], rho[:nch]
    # outer chains pressed `press` into the plane of the drum_mesh facet
    # they face (facet k spans the angles 2 pi k / 128 .. 2 pi (k + 1) / 128;
    # its plane lies R cos(pi / 128) from the axis, up to 75 um inside the
    # circle at R = 0.25 m)
    nseg = 128
    th = np.mod(np.arctan2(cy, cx), 2.0 * math.pi)
    phi = (np.floor(th / (2.0 * math.pi / nseg)) + 0.5) * (2.0 * math.pi / nseg)
    to_plane = (radius * math.cos(math.pi / nseg) - rc + press) / np.cos(th - phi)
    side = (rho > 1e-9) & (rho > radius - rc - 0.5 * pitch)
    sc = np.where(side, to_plane / np.maximum(rho, 1e-300), 1.0)
    px, py = cx * sc, cy * sc
    # length: chains of 2 r_c - 1.5 um per sphere hold n spheres
    length = float(n / np.sum(1.0 / (2.0 * rc - 1.525e-6))) + 4e-4
    for _ in range(20):
        per = np.floor((length - 2e-4 - 2.0 * rc) / (2.0 * rc - 0.05e-6)).astype(np.int64) + 1
        per = np.maximum(per, 0)
        if per.sum() >= n:
            break
        length *= 1.01
    # positions: z_0 = -L/2 + r_c + 1e-4, then steps of 2 r_c - U[0.05, 3] um
    # while the sphere stays 1e-4 m inside the end caps
    cid = np.repeat(np.arange(nch), per)
    step = 2.0 * rc[cid] - rng.uniform(0.05e-6, 3e-6, cid.size)
    start = np.concatenate([[0], np.cumsum(per)[:-1]])
    cs = np.cumsum(step)
    z0 = -0.5 * length + rc[cid] + 1e-4
    z = z0 + cs - step - np.repeat(cs[start] - step[start], per)
    ok = z + rc[cid] < 0.5 * length - 1e-4
    sel = np.nonzero(ok)[0][:n]
    if sel.size < n:
        raise RuntimeError(f"the drum holds only {sel.size} of {n} chained spheres")
    cid = cid[sel]
    x = np.stack([px[cid], py[cid], z[sel]], 1)
    r = rc[cid]
    m = sphere_mass(r, 2500.0)
    v9, n3 = drum_mesh(radius, length)
    pad = 0.005
    p = abi.default_params(2e-6, (-radius - pad, -radius - pad, -0.5 * length - pad),
                           (radius + pad, radius + pad, 0.5 * length + pad))
    p.normal_model = abi.RP_NORMAL_JKR
    p.tangential_model = abi.RP_TANGENTIAL_COULOMB
    p.particle = abi.RpMaterial(5e7, 0.3, 2500.0, 0.5, 0.5, 0.01)
    return p, x, r, m, [(v9, n3)]


def drum_test_workload(n: int = 20_000, seed: int = 0xD3E50015, radius: float = 0.03, length: float = 0.03,
                       press: float = 15e-6):
    """The capacity run's spheres, materials and drum (drum_workload: r ~
    U[0.4, 0.8] mm, JKR gamma 0.01 + Coulomb, E=5e7, 16,384-triangle drum
    turning at 30 rpm) at a size the oracle follows: a 6 cm drum whose lower
    part holds n spheres in chains along the drum's axis.  The spheres of one
    chain share a radius (drawn per chain) and touch their neighbours with
    overlaps U[0.05, 3] um -- JKR contacts with a root, the range a settled
    bed has and the surrogate serves; chains are 2.2 r_max apart (inside each
    other's JKR cut-off, without a root); the outer chains press 15 um into
    the drum's side.  Returns (params, x, r, m, meshes); mesh 0 turns with
    DRUM_OMEGA about DRUM_ORIGIN."""
    rng = np.random.default_rng(seed)
    rmin, rmax = 0.4e-3, 0.8e-3
    pitch = 2.2 * rmax
    g = np.arange(-radius, radius + 1e-12, pitch)
    X, Y = np.meshgrid(g, g, indexing="ij")
    cx, cy = X.ravel(), Y.ravel()
    order = np.lexsort((cx, cy))  # lowest chains first
    xs, rs = [], []
    count = 0
    for q in order:
        rc = rng.uniform(rmin, rmax)
        rho = math.hypot(cx[q], cy[q])
        if rho > radius - rc + press:
            continue
        if rho > 1e-9 and rho > radius - rc - 0.5 * pitch:
            # press the chain onto the side: `press` into the circle, more
            # than the 9 um the 128-sided polygon's facets lie inside it
            s_ = (radius - rc + press) / rho
            px, py = cx[q] * s_, cy[q] * s_
        else:
            px, py = cx[q], cy[q]
        z = -0.5 * length + rc + 1e-4
        while z + rc < 0.5 * length - 1e-4 and count < n:
            xs.append((px, py, z))
            rs.append(rc)
            count += 1
            z += 2.0 * rc - rng.uniform(0.05e-6, 3e-6)
        if count >= n:
            break
    if count < n:
        raise RuntimeError(f"the test drum holds only {count} of {n} spheres")
    x = np.array(xs)
    r = np.array(rs)
    m = sphere_mass(r, 2500.0)
    v9, n3 = drum_mesh(radius, length)
    pad = 0.002
    p = abi.default_params(2e-6, (-radius - pad, -radius - pad, -0.5 * length - pad),
                           (radius + pad, radius + pad, 0.5 * length + pad))
    p.normal_model = abi.RP_NORMAL_JKR
    p.tangential_model = abi.RP_TANGENTIAL_COULOMB
    p.particle = abi.RpMaterial(5e7, 0.3, 2500.0, 0.5, 0.5, 0.01)
    return p, x, r, m, [(v9, n3)]
