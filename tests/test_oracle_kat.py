"""The oracle against the reference's own known-answer fixtures.

FEparts/*.sol hold the reference test's expected superelement solution:
after reduction, solve K_red v = G g with g = (0, 9.81, 0) and the first six
reduced DOFs fixed; every nonzero entry must match within 1e-9 absolute
(solveSupel.f90:120-184, vpmReducerTests/CMakeLists.txt:50-57)."""
import os

import numpy as np
import pytest

from oracle import nastran
from oracle import reducer as orc


def kat(name, golden_dir):
    part = nastran.read(os.path.join(golden_dir, name + ".nas"))
    with open(os.path.join(golden_dir, name + ".sol")) as f:
        S = np.array([float(x) for x in f.read().split()])
    o = orc.reduce(part)
    A = o["SK"][6:, 6:]
    R = o["G"][6:] @ np.array([0.0, 9.81, 0.0])
    v = np.linalg.solve(A, R)
    return v, S


def test_chexa08_known_answer(golden_dir):
    v, S = kat("CHEXA08_", golden_dir)
    for i, s in enumerate(S):
        if abs(s) > 0:
            assert abs(v[i] - s) <= 1e-9, (i, v[i], s)


def test_cquad04_known_answer(golden_dir):
    v, S = kat("CQUAD04_", golden_dir)
    for i, s in enumerate(S):
        if abs(s) > 0:
            assert abs(v[i] - s) <= 1e-9, (i, v[i], s)


def test_cquad04_kat_fails_with_planted_error(golden_dir, monkeypatch):
    """A dropped bending stiffness term (planted in the oracle) must break
    the shell KAT."""
    real = orc.andes4_matrices

    def bad(xyz, E, nu, th, rho):
        return real(xyz, E, nu, np.asarray(th) * 0.99, rho)

    monkeypatch.setattr(orc, "andes4_matrices", bad)
    v, S = kat("CQUAD04_", golden_dir)
    assert abs(v[1] - S[1]) > 1e-9


def test_cbar02_known_answer(golden_dir):
    v, S = kat("CBAR02_", golden_dir)
    for i, s in enumerate(S):
        if abs(s) > 0:
            assert abs(v[i] - s) <= 1e-9, (i, v[i], s)


def test_cbar02_kat_fails_with_swapped_section(golden_dir, monkeypatch):
    """Swapping the two bending inertias (planted in the oracle) must break
    the beam KAT."""
    real = orc.beam_matrices

    def bad(xyz, pr):
        q = np.array(pr, float).copy()
        q[5], q[6] = pr[6], pr[5]
        return real(xyz, q)

    monkeypatch.setattr(orc, "beam_matrices", bad)
    v, S = kat("CBAR02_", golden_dir)
    assert abs(v[1] - S[1]) > 1e-9


def test_kat_fails_with_planted_error(golden_dir, monkeypatch):
    """A dropped mass term (planted in the oracle) must break the KAT."""
    real = orc.hex8_matrices

    def bad(*a):
        K, M = real(*a)
        return K, 0.9 * M

    monkeypatch.setattr(orc, "hex8_matrices", bad)
    v, S = kat("CHEXA08_", golden_dir)
    assert abs(v[1] - S[1]) > 1e-9


# Node displacements of the reference's direct FE solve (solvePartDis):
# src/vpmReducer/vpmReducerTests/test_FEsolve.C:124-133, 1e-4 absolute.
# External (ASET) nodes are clamped in the direct solve; load case = first
# FORCE set.  (01_1D/02_1D2D/L-beam use CBEAM/PBEAM(L) cards and
# 04_3D_VolumeElements uses 20-node CHEXA, outside the element set here.)
FE_SOLVE = [("03_2D_ShellElements", 29358, (0.0, 0.0123, 0.0)),
            ("CylinderBeam", 6, (0.0, 0.3450, 0.0)),
            ("CylinderBeam-generic", 6, (0.0, 0.3450, 0.0)),
            ("CylinderBeam-shearY", 6, (0.3434, 0.3450, 0.0)),
            ("CylinderBeam-noshear", 6, (0.0, 0.3434, 0.0))]


def fe_solve(name, node, golden_dir):
    part = nastran.read(os.path.join(golden_dir, name + ".nas"))
    return orc.direct_solve_displacements(part, [node])[0]


@pytest.mark.parametrize("name,node,want", FE_SOLVE)
def test_fe_solve_known_answers(name, node, want, golden_dir):
    d = fe_solve(name, node, golden_dir)
    assert np.max(np.abs(d - np.array(want))) <= 1e-4, d


def test_fe_solve_kat_fails_with_planted_shear_error(golden_dir, monkeypatch):
    """Shear factors taken as K instead of 1/K (planted in the oracle) must
    break the Timoshenko case."""
    real = orc.beam_matrices

    def bad(xyz, pr):
        q = np.array(pr, float).copy()
        q[11] = 1.0 / q[11] if q[11] > 0 else 0.0
        q[12] = 1.0 / q[12] if q[12] > 0 else 0.0
        return real(xyz, q)

    monkeypatch.setattr(orc, "beam_matrices", bad)
    d = fe_solve("CylinderBeam-shearY", 6, golden_dir)
    assert abs(d[1] - 0.3450) > 1e-4
