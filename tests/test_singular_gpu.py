"""Singular and indefinite K_ii on the device (singularityHandler /
tolFactorize, solExtensionModule.f90:716-760, cmstrs.f90:635-652).

Hinged blocks: three true zero pivots (the second block's rigid rotations
about the hinge).  With the default handler they are replaced and the
reduction goes on; the boundary stiffness does not see the mechanisms, so
SK equals K_bb - K_bi K_ii^+ K_ib with the pseudo-inverse.  Handler 0 makes
them an error that names the equations.  Indefinite block: as many negative
pivots as K_ii has negative eigenvalues (Sylvester), a warning only, and
every output equal to the oracle's."""
import os

import numpy as np
import pytest

from fedempy_amd import gpu, parts
from fedempy_amd.reducer import FedemReducer
from oracle import reducer as orc
from singular_parts import hinged_blocks, indefinite_block

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def pinv_schur(p):
    s = orc.sam_numbering(p)
    K, _, _ = orc.assemble(p, s)
    K = K.toarray()
    i1, i2 = s["meqn1"] - 1, s["meqn2"] - 1
    Kii, Kib, Kbb = K[np.ix_(i1, i1)], K[np.ix_(i1, i2)], K[np.ix_(i2, i2)]
    S = Kbb - Kib.T @ np.linalg.pinv(Kii, rcond=1e-10) @ Kib
    pos = s["dofPosIn2"] - 1
    return S[np.ix_(pos, pos)], s


def test_hinged_true_zero_pivots_are_suppressed():
    p = hinged_blocks()
    print("hinged blocks, default handler", flush=True)
    R = gpu.GpuReducer()
    R.set_part(p)
    R.analyze()
    R.reduce()
    (neg, ntrue, nred), sing = R.pivots()
    assert (neg, ntrue, nred) == (0, 3, 0)
    na = 27
    assert len(sing) == 3 and all(node >= na for node, _, _, kind in sing) and all(k == 1 for *_, k in sing)
    SK, _, _ = R.reduced()
    S, _ = pinv_schur(p)
    assert rel(SK[:len(S), :len(S)], S) <= 1e-9


def test_hinged_handler0_is_an_error():
    p = hinged_blocks()
    print("hinged blocks, handler 0", flush=True)
    R = gpu.GpuReducer()
    R.set_part(p)
    R.set_option("singularityHandler", 0)
    R.analyze()
    with pytest.raises(gpu.ReducerError, match="singular K_ii"):
        R.reduce()
    (neg, ntrue, nred), sing = R.pivots()
    assert ntrue + nred == 3 and len(sing) == 3


def test_indefinite_block_negative_pivots_warn_only():
    p = indefinite_block()
    print("indefinite block", flush=True)
    R = gpu.GpuReducer()
    R.set_part(p)
    R.analyze(vgi=True)
    R.reduce()
    (neg, ntrue, nred), _ = R.pivots()
    s = orc.sam_numbering(p)
    K, _, _ = orc.assemble(p, s)
    i1 = s["meqn1"] - 1
    assert neg == int(np.sum(np.linalg.eigvalsh(K.toarray()[np.ix_(i1, i1)]) < 0)) == 4
    assert ntrue == nred == 0
    SK, SM, G = R.reduced()
    o = orc.reduce(p, want_vgi=True)
    for a, b in ((SK, o["SK"]), (SM, o["SM"]), (G, o["G"]), (R.B(), o["B"]), (R.vgi(), o["vgi"])):
        assert rel(a, b) <= 1e-10


def test_indefinite_block_mixed_precision():
    """mixed precision on the indefinite block: conjugate gradients need a
    positive definite K_ii, so the refinement turns stationary at the first
    non-positive curvature (or r^T z) and still meets the mixed path's
    1e-8 against the oracle"""
    p = indefinite_block()
    print("indefinite block, mixed precision", flush=True)
    R = gpu.GpuReducer()
    try:
        R.set_part(p)
        R.set_option("mixedPrecision", 1)
        R.analyze(vgi=True)
        R.reduce()
        SK, SM, G = R.reduced()
        B, vgi = R.B(), R.vgi()
    finally:
        R.close()
    o = orc.reduce(p, want_vgi=True)
    for a, b in ((SK, o["SK"]), (SM, o["SM"]), (G, o["G"]), (B, o["B"]), (vgi, o["vgi"])):
        assert rel(a, b) <= 1e-8


def test_dropin_reports_singular_nodes(tmp_path):
    p = hinged_blocks()
    nas = tmp_path / "hinge.nas"
    parts.write_nastran(p, str(nas))
    print("drop-in hinged blocks, handler 0 then 1", flush=True)
    red = FedemReducer(solver_options=["-linkfile=%s" % nas, "-singularityHandler=0"])
    assert red.run() != 0
    red.release()
    res = (tmp_path / "hinge.res").read_text()
    assert "Singular or ill-conditioned stiffness matrix" in res
    assert res.count(" local DOF") == 3
    red = FedemReducer(solver_options=["-linkfile=%s" % nas])
    assert red.run() == 0
    red.release()
    res = (tmp_path / "hinge.res").read_text()
    assert "circumvented" in res and res.count(" local DOF") == 3
