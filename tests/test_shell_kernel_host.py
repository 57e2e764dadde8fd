"""The device ANDES4 element code (csrc/shell_andes4.hpp), run on the host
lane by lane in the kernel's phase order by tests/harness/shell_host.cpp,
against the reference's own Femlib build (golden vectors of
Andes4shell_stiffmat, trans3P, PMATSTIFF and FQS35) and against the oracle
on warped, rotated elements.  No GPU."""
import os
import subprocess

import numpy as np
import pytest

from oracle import reducer as orc

HERE = os.path.dirname(os.path.abspath(__file__))
PER = 1 + 576 * 2 + 360 + 24


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("shell") / "shell_host")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "harness", "shell_host.cpp")],
                   check=True)

    def run(elements):
        """elements: list of (E, nu, rho, th[4], xyz[4,3]) -> list of dicts"""
        inp = "\n".join(" ".join("%.17e" % v for v in [E, nu, rho, *th, *xyz[:, 0], *xyz[:, 1], *xyz[:, 2]])
                        for E, nu, rho, th, xyz in elements) + "\n"
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split()
        res = []
        for k in range(len(elements)):
            o = out[k * PER:(k + 1) * PER]
            v = np.array(o[1:], float)
            res.append({"bad": int(o[0]), "K": v[:576].reshape(24, 24), "P": v[576:1152].reshape(24, 24),
                        "blocks": v[1152:1512].reshape(10, 6, 6), "em": v[1512:]})
        return res

    return run


def test_matches_reference_femlib_build(harness, golden_dir):
    """Unprojected global stiffness = T K_loc T^T with the reference's local
    ANDES stiffness and frame; projector and lumped mass as the reference
    computes them."""
    import json
    with open(os.path.join(golden_dir, "femlib_ref_vectors.json")) as f:
        g = json.load(f)
    els = [(c["E"], c["nu"], 7850.0, [c["t"]] * 4, np.array(c["xyz"])) for c in g["andes4"]]
    for c, r in zip(g["andes4"], harness(els)):
        assert r["bad"] == 0
        tr = np.array(c["trel"])[:9].reshape(3, 3, order="F")
        T = np.kron(np.eye(8), tr)
        Kr = np.array(c["Kloc"]).reshape(24, 24, order="F")
        Kg = T @ Kr @ T.T
        assert np.linalg.norm(r["K"] - Kg) <= 1e-13 * np.linalg.norm(Kg)
        Pr = np.array(c["P"]).reshape(24, 24, order="F")
        assert np.abs(r["P"] - Pr).max() <= 1e-14
    els = [(2.1e11, 0.3, c["rho"], c["th"], np.array(c["xyz"])) for c in g["fqs35"]]
    for c, r in zip(g["fqs35"], harness(els)):
        em = np.array(c["EM"])
        assert np.abs(r["em"] - em).max() <= 1e-13 * np.abs(em).max()
        assert np.all(np.abs(r["em"] - em) <= 4e-15 * np.abs(em))


def test_matches_oracle_on_warped_elements(harness):
    rng = np.random.default_rng(3)
    els = []
    for k in range(24):
        size = rng.uniform(0.002, 0.1)
        base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0.]]) * size
        xyz = base + rng.uniform(-0.2, 0.2, (4, 3)) * size * np.array([1, 1, 0.3])
        Q = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        xyz = xyz @ Q.T + rng.normal(size=3)
        els.append((2.1e11, rng.uniform(0.2, 0.35), 7850.0, rng.uniform(0.002, 0.008, 4), xyz))
    for (E, nu, rho, th, xyz), r in zip(els, harness(els)):
        assert r["bad"] == 0
        K, M = orc.andes4_matrices(xyz, E, nu, th, rho)
        t = 0
        for a in range(4):
            for b in range(a + 1):
                assert np.abs(r["blocks"][t] - K[6 * a:6 * a + 6, 6 * b:6 * b + 6]).max() <= 1e-13 * np.abs(K).max()
                t += 1
        assert np.allclose(r["em"], np.diag(M), rtol=1e-13, atol=0)


def test_degenerate_element_is_flagged(harness):
    xyz = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0.]]) * 0.01  # collinear corners
    r, = harness([(2.1e11, 0.3, 7850.0, [0.004] * 4, xyz)])
    assert r["bad"] == 1
