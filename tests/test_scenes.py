"""Procedural STL generators: the drum of the 16M-sphere capacity run.

Checks fixed by geometry, not by the generator's own formulas:
* every edge belongs to exactly two triangles (a closed surface);
* the normals point inward: by the divergence theorem the sum of
  (centroid . n) * area over a closed surface with inward normals is -3 V,
  and V of the polygonal drum is (nseg / 2) R^2 sin(2 pi / nseg) L;
* the triangle count of the capacity run (16,384) and the binary and ASCII
  files read back by the library's reader (mesh.rs:125-143 semantics).
"""
import math

import numpy as np
import pytest

from rusty_particles_amd import native, scenes


def _edges_closed(v9):
    keys = {}
    verts = v9.reshape(-1, 3, 3)
    ids = np.empty(verts.shape[:2], np.int64)
    for t in range(verts.shape[0]):
        for k in range(3):
            ids[t, k] = keys.setdefault(tuple(verts[t, k].tolist()), len(keys))
    e = np.sort(np.concatenate([ids[:, [0, 1]], ids[:, [1, 2]], ids[:, [2, 0]]]), axis=1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    return counts


@pytest.mark.parametrize("nseg,nlen,ncap", [(8, 3, 2), (128, 49, 8)])
def test_drum_is_closed_with_inward_normals(nseg, nlen, ncap):
    R, L = 0.25, 0.2
    v9, n3 = scenes.drum_mesh(R, L, (0.01, -0.02, 0.03), nseg=nseg, nlen=nlen, ncap=ncap)
    assert v9.shape == (2 * nseg * (nlen + 2 * ncap - 1), 9) and n3.shape == (v9.shape[0], 3)
    counts = _edges_closed(v9)
    assert np.all(counts == 2)
    a = v9[:, 0:3].astype(np.float64)
    b = v9[:, 3:6].astype(np.float64)
    c = v9[:, 6:9].astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)
    cen = (a + b + c) / 3.0
    flux = float(((cen * n3.astype(np.float64)).sum(1) * area).sum())
    vol = 0.5 * nseg * R * R * math.sin(2 * math.pi / nseg) * L
    assert abs(-flux / 3.0 - vol) < 1e-5 * vol
    # no normal points away from the inside: centroid + small step along n stays inside
    q = cen + 1e-4 * n3
    rho = np.hypot(q[:, 0] - 0.01, q[:, 1] + 0.02)
    assert np.all(rho < R + 1e-6) and np.all(np.abs(q[:, 2] - 0.03) < 0.5 * L + 1e-6)


def test_drum_default_is_the_capacity_run_mesh():
    v9, _ = scenes.drum_mesh()
    assert v9.shape[0] == 16_384


@pytest.mark.parametrize("fmt", ["binary", "ascii"])
def test_drum_stl_round_trip(tmp_path, fmt):
    v9, n3 = scenes.drum_mesh(nseg=16, nlen=4, ncap=3)
    path = str(tmp_path / f"d.{fmt}.stl")
    (scenes.write_stl_binary if fmt == "binary" else scenes.write_stl_ascii)(path, v9, n3)
    a9, a3 = native.read_stl(path)
    assert np.array_equal(a9, v9) and np.array_equal(a3, n3)
