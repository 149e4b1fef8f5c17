"""The oracle's pair lists against a brute-force restatement of the
reference's grid semantics (grid.rs:12-87, simulation.rs:323-357): for each
particle, the 27 neighbour cells in z, y, x order (periodic wrap once, range
check), cell members in index order, kept when dist^2 < cutoff^2.  Particles
outside the reference grid are invisible but still gather."""
import math

import numpy as np
import pytest

from rusty_particles_amd import abi

from helpers import params, run_oracle


def brute_pairs(x, r, bmin, bmax, periodic, cs, adhesive):
    bmin, bmax = np.asarray(bmin, float), np.asarray(bmax, float)
    size = bmax - bmin
    dims = [int(math.ceil(size[k] / cs)) + 1 for k in range(3)]

    def coords(p):
        return [int(math.floor((p[k] - bmin[k]) / cs)) for k in range(3)]
    cells = {}
    for j in range(len(x)):
        c = coords(x[j])
        if all(0 <= c[k] < dims[k] for k in range(3)):
            cells.setdefault(tuple(c), []).append(j)
    offs, js = [0], []
    for i in range(len(x)):
        ci = coords(x[i])
        for oz in (-1, 0, 1):
            for oy in (-1, 0, 1):
                for ox in (-1, 0, 1):
                    nb = [ci[0] + ox, ci[1] + oy, ci[2] + oz]
                    for k in range(3):
                        if periodic[k]:
                            if nb[k] < 0:
                                nb[k] += dims[k]
                            elif nb[k] >= dims[k]:
                                nb[k] -= dims[k]
                    if not all(0 <= nb[k] < dims[k] for k in range(3)):
                        continue
                    for j in cells.get(tuple(nb), []):
                        if j == i:
                            continue
                        d = x[i] - x[j]
                        for k in range(3):
                            if periodic[k]:
                                if d[k] > size[k] * 0.5:
                                    d[k] -= size[k]
                                elif d[k] < -size[k] * 0.5:
                                    d[k] += size[k]
                        rs = r[i] + r[j]
                        cut = rs + (rs * 0.5 if adhesive else 0.0)
                        if (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2] < cut * cut:
                            js.append(j)
        offs.append(len(js))
    return np.array(offs), np.array(js, dtype=np.int64)


@pytest.mark.parametrize("case", ["one_cell", "multi_cell", "periodic", "adhesive"])
def test_pairs_match_brute_force(case):
    rng = np.random.default_rng(42)
    n = 400
    if case == "multi_cell":
        bmin, bmax, per = (-2.0, -2.0, -2.0), (2.0, 2.0, 2.0), (0, 0, 0)
        x = rng.uniform(-0.6, 0.4, (n, 3)) * np.array([1.0, 1.0, 1.0])  # straddles the cell planes at -1, 0
        x[:20, 0] = rng.uniform(-2.3, -2.0, 20)  # invisible: below bounds_min.x
        r = rng.uniform(0.04, 0.08, n)
    elif case == "periodic":
        bmin, bmax, per = (0.0, 0.0, 0.0), (0.9, 0.9, 0.9), (1, 1, 0)
        x = rng.uniform(0.0, 0.9, (n, 3))
        r = rng.uniform(0.03, 0.06, n)
    else:
        bmin, bmax, per = (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), (0, 0, 0)
        x = rng.uniform(-0.5, 0.5, (n, 3))
        r = rng.uniform(0.03, 0.06, n)
    normal = abi.RP_NORMAL_JKR if case == "adhesive" else abi.RP_NORMAL_HERTZ
    p = params(1e-9, bmin, bmax, normal=normal, gamma=0.01 if case == "adhesive" else 0.0, periodic=per)
    p.gravity[1] = 0.0
    sc = dict(x=x, radius=r, mass=np.ones(n))
    o = run_oracle(p, sc, (), steps=1)
    # the gather of step 1 sees positions after the first drift; with v = a = 0
    # the drift is exactly zero, so the initial positions are the ones tested
    offs, js = o.pairs()
    bo, bj = brute_pairs(x, r, bmin, bmax, per, 1.0, case == "adhesive")
    assert js.size > 50
    assert np.array_equal(offs, bo) and np.array_equal(js, bj)
