"""The full-size fixtures (tests/golden/fullsize_*.npz, the oracle's
reduction of the 48x24x24 HEX8 block and of the 200x200 ANDES4 plate)
checked against what the mechanics fixes, so that they pin the GPU tests
that read them: symmetric reduced matrices, rigid translations in the null
space of the unconstrained parts' reduced stiffness, the reduced gravity
load of a rigid translation equal to the part's mass (rho V, rho sum t_e A_e)
and positive eigenvalues; and the fixed probe vector of B^T r."""
import os
import sys

import numpy as np
import pytest

from fedempy_amd import parts

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name):
    return np.load(os.path.join(HERE, "golden", "fullsize_%s.npz" % name), allow_pickle=False)


@pytest.mark.parametrize("name,ngen,mtol", [("hex8_48x24x24", 20, 2e-7), ("plate200", 0, 1e-9)])
def test_fixture_mechanics(name, ngen, mtol):
    """External nodes carry six reduced DOFs (a solid node's rotations are
    zero rows).  HEX8 masses are integrated with the reference's
    single-precision Gauss weights (hexa.f), hence 2e-7 on the block's mass."""
    ndn = 6
    f = load(name)
    SK, SM, G = f["SK"], f["SM"], f["G"]
    n2 = SK.shape[0] - ngen
    for A in (SK, SM):
        assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    for d in range(3):  # rigid translation along axis d of the boundary DOFs
        r = np.zeros(SK.shape[0])
        r[d:n2:ndn] = 1.0
        assert np.abs(SK @ r).max() <= 1e-8 * np.abs(SK).max()
    if name == "hex8_48x24x24":
        mass = 7850.0 * 0.48 * 0.24 * 0.24
    else:
        p = parts.config2(seed=0)
        x = p["xyz"][(p["mmnpc"] - 1).reshape(-1, 4)]
        area = 0.5 * np.abs(np.cross(x[:, 2] - x[:, 0], x[:, 3] - x[:, 1])[:, 2])
        mass = np.sum(7850.0 * p["eprop"][:, 4:8].mean(axis=1) * area)
    for d in range(3):
        assert abs(G[d:n2:ndn, d].sum() - mass) <= mtol * mass
    if ngen:
        assert np.all(f["lam"] > 0) and np.all(np.diff(f["lam"]) > 0)


def test_probe_matches_fixture_script():
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_fullsize import probe
    from test_fullsize_gpu import probe as used
    assert np.array_equal(probe(1000), used(1000))
