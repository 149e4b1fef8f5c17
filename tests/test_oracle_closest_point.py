"""The oracle's Triangle::closest_point (mesh.rs:45-117) against hand-worked
known answers, one per branch (tests/closest_point_cases.py), in both
precisions.  A wrong clamp bound ([0, 1] for [0, a]), a swapped parameter
(region 6b clamps -e onto the e0 parameter) or a dropped division (region 0
divides by det) each move at least one of these points by O(1).

The true closest points the table gives for the branches where the
reference's differs are themselves checked against brute-force sampling of
the triangle, so the table pins both semantics."""
import numpy as np
import pytest

from closest_point_cases import CASES, tri12
from oracle import sim as osim


def _brute_closest(tri, p, k=1201):
    v0, v1, v2 = (np.array(v) for v in tri)
    s, t = np.meshgrid(np.linspace(0, 1, k), np.linspace(0, 1, k))
    m = s + t <= 1
    pts = v0 + s[m][:, None] * (v1 - v0) + t[m][:, None] * (v2 - v0)
    d = np.linalg.norm(pts - np.asarray(p), axis=1)
    return pts[d.argmin()], d.min()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_reference_branch_known_answer(case):
    name, tri, p, want, _ = case
    for bits, tol in ((64, 1e-13), (32, 2e-6)):
        cp, _, _, _ = osim.triangle(bits, tri12(tri), p, 0.0)
        assert np.allclose(cp, want, rtol=0.0, atol=tol * max(1.0, np.abs(want).max())), (name, bits, cp, want)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_table_true_closest_points(case):
    """The table's true closest point (given where the reference's differs;
    otherwise the reference's point is the true one) is the brute-force
    minimiser of the distance over the triangle."""
    name, tri, p, want, true = case
    q = np.asarray(true if true is not None else want)
    b, dmin = _brute_closest(tri, p)
    assert abs(np.linalg.norm(q - np.asarray(p)) - dmin) < 2e-3, (name, q, b)
    assert np.linalg.norm(q - b) < 1e-2, (name, q, b)
    if true is not None:
        # the reference's point is another point (off the triangle in 4a, 5
        # and 6b, where it can even lie nearer to p than the triangle does)
        assert np.linalg.norm(np.asarray(want) - q) > 0.1


def test_every_branch_is_covered():
    assert sorted(c[0] for c in CASES) == ["0", "1", "2a", "2b", "3", "4a", "4b", "5", "6a", "6b"]
