"""Wall contacts in every branch of the reference's closest_point, HIP engine
against the oracle (mesh.rs:45-117, simulation.rs:465-523).

One sphere and one triangle per case of tests/closest_point_cases.py: the
sphere's radius is 1 % above its distance to the reference's closest point, so
the reference sees a wall contact there -- also where that point lies off the
triangle (regions 4a, 5 and 6b on these edges longer than 1 m), which the
engine's triangle binning must cover.  A second sphere 1 % short of that
distance has no contact on either side.  Forces within 1e-10 (the bar of
tests/helpers.assert_forces_close), one wall contact counted.
"""
import numpy as np
import pytest

from closest_point_cases import CASES
from helpers import assert_forces_close, params, run_engine, run_oracle
from rusty_particles_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _mesh(tri):
    v9 = np.array([np.concatenate(tri)], dtype=np.float32)
    n3 = np.array([[0.0, 0.0, 1.0]], dtype=np.float32)
    return v9, n3


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("factor", [1.01, 0.99], ids=["hit", "miss"])
def test_wall_contact_in_every_branch(case, factor):
    name, tri, p, want, _ = case
    dist = float(np.linalg.norm(np.subtract(p, want)))
    r = factor * dist
    print(f"case region={name} factor={factor} r={r:.4g}", flush=True)
    pr = params(1e-6, (-40.0, -40.0, -40.0), (40.0, 40.0, 40.0), normal=abi.RP_NORMAL_HERTZ,
                tangential=abi.RP_TANGENTIAL_COULOMB, E=1e7)
    rr = np.array([r])
    sc = dict(x=np.array([p], float), radius=rr, mass=scenes.sphere_mass(rr, 2500.0),
              v=np.array([[0.01, -0.02, 0.03]]), w=np.array([[0.5, 0.0, -0.5]]), fixed=np.zeros(1, np.uint8))
    meshes = [_mesh(tri)]
    o = run_oracle(pr, sc, meshes, steps=1)
    e = run_engine(pr, sc, meshes, steps=1)
    assert_forces_close(o, e, pr)
    st = e.stats()
    assert st["wall_contacts"] == (1 if factor > 1.0 else 0)
    a = e.download()["a"][0]
    if factor < 1.0:
        assert np.array_equal(a, [0.0, -9.81, 0.0])
    else:
        assert not np.array_equal(a, [0.0, -9.81, 0.0])
