"""Oracle element restatement (oracle/elements.c) against the reference's own
Femlib routines: golden vectors produced by oracle/_ref (the reference
sources compiled by oracle/Makefile, see tests/golden/make_golden.py), the
live _ref build when present, and invariants of the mathematics."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import reducer as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "femlib_ref")


@pytest.fixture(scope="module")
def golden(golden_dir):
    with open(os.path.join(golden_dir, "femlib_ref_vectors.json")) as f:
        return json.load(f)


def test_hexa31_matches_reference_golden(golden):
    for g in golden["hexa31"]:
        K, _ = orc.hex8_matrices(g["X"], g["Y"], g["Z"], g["E"], g["nu"], 1.0)
        Kr = np.array(g["EK"]).reshape(24, 24, order="F")
        assert np.max(np.abs(K - Kr)) <= 1e-13 * np.max(np.abs(Kr))


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built here")
def test_hexa31_matches_live_reference_build():
    rng = np.random.default_rng(7)
    X = np.array([0, 1, 1, 0, 0, 1, 1, 0.]) * 0.02 + rng.uniform(-1e-3, 1e-3, 8)
    Y = np.array([0, 0, 1, 1, 0, 0, 1, 1.]) * 0.02 + rng.uniform(-1e-3, 1e-3, 8)
    Z = np.array([0, 0, 0, 0, 1, 1, 1, 1.]) * 0.02 + rng.uniform(-1e-3, 1e-3, 8)
    inp = "hexa31\n" + " ".join("%.17e" % v for v in [2e11, 0.3, *X, *Y, *Z]) + "\nquit\n"
    out = subprocess.run([REF], input=inp, capture_output=True, text=True, check=True).stdout.split()
    Kr = np.array(out[1:577], float).reshape(24, 24, order="F")
    K, _ = orc.hex8_matrices(X, Y, Z, 2e11, 0.3, 1.0)
    assert np.max(np.abs(K - Kr)) <= 1e-13 * np.max(np.abs(Kr))


def rigid_modes(X, Y, Z):
    R = np.zeros((24, 6))
    for a in range(8):
        x = np.array([X[a], Y[a], Z[a]])
        R[3 * a:3 * a + 3, :3] = np.eye(3)
        R[3 * a:3 * a + 3, 3:] = np.array([[0, x[2], -x[1]], [-x[2], 0, x[0]], [x[1], -x[0], 0]])
    return R


def test_hex8_invariants():
    rng = np.random.default_rng(3)
    X = np.array([0, 1, 1, 0, 0, 1, 1, 0.]) * 0.03 + rng.uniform(-2e-3, 2e-3, 8)
    Y = np.array([0, 0, 1, 1, 0, 0, 1, 1.]) * 0.02 + rng.uniform(-2e-3, 2e-3, 8)
    Z = np.array([0, 0, 0, 0, 1, 1, 1, 1.]) * 0.01 + rng.uniform(-1e-3, 1e-3, 8)
    K, M = orc.hex8_matrices(X, Y, Z, 2.1e11, 0.3, 7850.0)
    assert np.allclose(K, K.T, rtol=0, atol=1e-12 * np.abs(K).max())
    # six rigid-body modes in the null space, the rest positive
    R = rigid_modes(X, Y, Z)
    assert np.abs(K @ R).max() <= 1e-9 * np.abs(K).max() * np.abs(R).max()
    w = np.linalg.eigvalsh(K)
    assert np.sum(w < 1e-8 * w.max()) == 6
    # consistent mass: symmetric, positive definite
    assert np.allclose(M, M.T) and np.linalg.eigvalsh(M).min() > 0


def test_hex8_brick_mass_closed_form():
    """Rectangular brick a x b x c: M = rho V / 216 * [8,4,2,1 pattern] per
    direction.  The reference's 3x3x3 rule uses single-precision abscissae and
    weights, so the closed form holds to ~1e-7, not to rounding."""
    a, b, c, rho = 0.03, 0.02, 0.01, 7850.0
    X = np.array([0, 1, 1, 0, 0, 1, 1, 0.]) * a
    Y = np.array([0, 0, 1, 1, 0, 0, 1, 1.]) * b
    Z = np.array([0, 0, 0, 0, 1, 1, 1, 1.]) * c
    _, M = orc.hex8_matrices(X, Y, Z, 2.1e11, 0.3, rho)
    nat = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]])
    m = np.zeros((8, 8))
    for i in range(8):
        for j in range(8):
            m[i, j] = np.prod([(2.0 if nat[i][k] == nat[j][k] else 1.0) for k in range(3)])
    m *= rho * a * b * c / 216.0
    Mx = M[0::3, 0::3]
    assert np.allclose(Mx, m, rtol=2e-7, atol=0)
    assert abs(Mx.sum() - rho * a * b * c) <= 2e-7 * rho * a * b * c
    assert np.abs(M[0::3, 1::3]).max() == 0.0
