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


def _andes_local(g):
    """Local frame, local coordinates and constitutive matrix of a golden
    ANDES4 case, as ADD_ANDES4 builds them (inadd.f90:1470-1510)."""
    import ctypes
    L = orc.lib()
    xyz = np.array(g["xyz"])
    tr = np.array(g["trel"])
    Tel = tr[:9].reshape(3, 3, order="F").T
    loc = Tel @ (xyz - xyz[0]).T
    E, nu, t = g["E"], g["nu"], g["t"]
    c11 = E / (1 - nu * nu)
    C3 = np.array([[c11, nu * c11, 0], [nu * c11, c11, 0], [0, 0, 0.5 * E / (1 + nu)]])
    Cm = np.zeros((6, 6))
    Cm[:3, :3] = C3 * t
    Cm[3:, 3:] = C3 * t * t * t / 12.0
    return L, xyz, loc, Cm


def test_andes4_matches_reference_golden(golden):
    """Frame (trans3P), local ANDES stiffness, projector P and FQS35 mass of
    the restatement against the reference's own Femlib build."""
    import ctypes
    for g in golden["andes4"]:
        L, xyz, loc, Cm = _andes_local(g)
        tr = np.zeros(12)
        mid = 0.5 * (xyz[2] + xyz[3])
        L.andes4_trans3p(orc._dp(np.ascontiguousarray(xyz[0])), orc._dp(np.ascontiguousarray(xyz[1])),
                         orc._dp(mid), orc._dp(tr))
        assert np.array_equal(tr, np.array(g["trel"]))
        K = np.zeros(576)
        X, Y, Z = (np.ascontiguousarray(loc[i]) for i in range(3))
        L.andes4_local_stiff(orc._dp(X), orc._dp(Y), orc._dp(Z), orc._dp(Cm.ravel(order="F").copy()),
                             ctypes.c_double(1.5), ctypes.c_double(0.9), orc._dp(K))
        Kr = np.array(g["Kloc"])
        assert np.linalg.norm(K - Kr) <= 1e-14 * np.linalg.norm(Kr)
        P = np.zeros(576)
        L.pmat_stiff(4, orc._dp(np.ascontiguousarray((xyz - xyz[0]).ravel())), orc._dp(P))
        assert np.max(np.abs(P - np.array(g["P"]))) <= 1e-15
    for g in golden["fqs35"]:
        _, M = orc.andes4_matrices(np.array(g["xyz"]), 2.1e11, 0.3, g["th"], g["rho"])
        assert np.array_equal(np.diag(M), np.array(g["EM"]))


def shell_rigid_modes(xyz):
    R = np.zeros((24, 6))
    for a in range(4):
        x = xyz[a]
        R[6 * a:6 * a + 3, :3] = np.eye(3)
        R[6 * a:6 * a + 3, 3:] = np.array([[0, x[2], -x[1]], [-x[2], 0, x[0]], [x[1], -x[0], 0]])
        R[6 * a + 3:6 * a + 6, 3:] = np.eye(3)
    return R


def test_andes4_invariants():
    """Projected stiffness: symmetric, the six rigid-body motions of the
    (warped, rotated) element in its null space, no other zero modes beyond
    the drilling-free ones; lumped mass: translational masses sum to
    rho t A."""
    rng = np.random.default_rng(11)
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0.]]) * 0.05
    xyz = base + rng.uniform(-4e-3, 4e-3, (4, 3)) * np.array([1, 1, 0.3])
    Q = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    xyz = xyz @ Q.T + np.array([0.3, -0.1, 2.0])
    th = np.full(4, 0.004)
    K, M = orc.andes4_matrices(xyz, 2.1e11, 0.3, th, 7850.0)
    assert np.allclose(K, K.T, rtol=0, atol=1e-10 * np.abs(K).max())
    R = shell_rigid_modes(xyz)
    assert np.abs(K @ R).max() <= 1e-8 * np.abs(K).max() * np.abs(R).max()
    w = np.linalg.eigvalsh(K)
    assert np.sum(w < 1e-10 * w.max()) >= 6 and w.min() > -1e-9 * w.max()
    # planar element in the xy plane: translational lumped mass = rho t A
    flat = base + np.array([[0, 0, 0], [0.01, 0.002, 0], [0.0, 0.01, 0], [-0.004, 0, 0.]])
    _, Mf = orc.andes4_matrices(flat, 2.1e11, 0.3, th, 7850.0)
    x, y = flat[:, 0], flat[:, 1]
    A = 0.5 * abs((x[0] - x[2]) * (y[1] - y[3]) - (x[1] - x[3]) * (y[0] - y[2]))
    m = np.diag(Mf)
    for d in range(3):
        assert abs(m[d::6].sum() - 7850.0 * 0.004 * A) <= 1e-12 * 7850.0 * 0.004 * A


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built here")
def test_andes4_matches_live_reference_build():
    rng = np.random.default_rng(5)
    xyz = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0.]]) * 0.03 + rng.uniform(-2e-3, 2e-3, (4, 3))
    inp = "andes4\n" + " ".join("%.17e" % v for v in [2.1e11, 0.3, 0.003, *xyz[:, 0], *xyz[:, 1], *xyz[:, 2]])
    out = subprocess.run([REF], input=inp + "\nquit\n", capture_output=True, text=True, check=True).stdout.split()
    g = {"xyz": xyz.tolist(), "E": 2.1e11, "nu": 0.3, "t": 0.003, "trel": [float(v) for v in out[1:13]]}
    Kr = np.array(out[13:13 + 576], float).reshape(24, 24, order="F")
    Pr = np.array(out[13 + 576:13 + 1152], float).reshape(24, 24, order="F")
    # full ADD_ANDES4 chain: rotate the reference's local K, project with its P
    tr = np.array(g["trel"])[:9].reshape(3, 3, order="F")
    T = np.kron(np.eye(8), tr)
    EKr = Pr.T @ (T @ Kr @ T.T) @ Pr
    K, _ = orc.andes4_matrices(xyz, 2.1e11, 0.3, np.full(4, 0.003), 7850.0)
    assert np.linalg.norm(K - EKr) <= 1e-12 * np.linalg.norm(EKr)


def test_beam_closed_forms():
    """Beam restatement against the mathematics: a cantilever's tip
    deflections P L^3/(3 E I) (1 + 3 phi / (1 + ... )) in Timoshenko form,
    the axial and torsional stiffness, six rigid-body modes of an arbitrarily
    oriented element and the consistent-mass totals (rho A L per direction)."""
    E, G, A, I1, I2, J, L = 2.1e11, 8.1e10, 3e-3, 2e-6, 5e-6, 4e-6, 0.7
    K1 = 0.8
    pr = np.zeros(16)
    pr[0], pr[1], pr[2], pr[3] = E, 0.3, 7800.0, G
    pr[4:8] = A, I1, I2, J
    pr[8:11] = 0.0, 1.0, 0.0  # local y = global y for a beam along x
    pr[11] = pr[12] = K1
    K, M, em = orc.beam_matrices(np.array([[0, 0, 0], [L, 0, 0.]]), pr)
    assert np.allclose(K, K.T, rtol=0, atol=1e-9 * np.abs(K).max())
    f = K[6:, 6:]
    u = np.linalg.solve(f, np.eye(6))  # node 1 clamped: flexibility at node 2
    phi1 = 12 * E * I1 / (K1 * A * G * L * L)
    phi2 = 12 * E * I2 / (K1 * A * G * L * L)
    assert np.isclose(u[0, 0], L / (E * A), rtol=1e-12)
    assert np.isclose(u[3, 3], L / (G * J), rtol=1e-12)
    assert np.isclose(u[1, 1], L ** 3 / (3 * E * I1) * (1 + phi1 / 4), rtol=1e-12)  # y: I1
    assert np.isclose(u[2, 2], L ** 3 / (3 * E * I2) * (1 + phi2 / 4), rtol=1e-12)  # z: I2
    # rotated, skew element: rigid-body modes in the null space
    rng = np.random.default_rng(2)
    x = rng.normal(size=(2, 3))
    pr[8:11] = rng.normal(size=3)
    K, M, em = orc.beam_matrices(x, pr)
    R = np.zeros((12, 6))
    for a in range(2):
        R[6 * a:6 * a + 3, :3] = np.eye(3)
        p = x[a]
        R[6 * a:6 * a + 3, 3:] = np.array([[0, p[2], -p[1]], [-p[2], 0, p[0]], [p[1], -p[0], 0]])
        R[6 * a + 3:6 * a + 6, 3:] = np.eye(3)
    assert np.abs(K @ R).max() <= 1e-8 * np.abs(K).max() * np.abs(R).max()
    Lx = np.linalg.norm(x[1] - x[0])
    m = 7800.0 * A * Lx
    assert np.isclose(em, m, rtol=1e-14)
    for d in range(3):
        assert np.isclose(R[:, d] @ M @ R[:, d], m, rtol=1e-12)
