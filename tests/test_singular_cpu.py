"""The test parts of the singular-pivot handling, checked on the CPU: the
hinged blocks have exactly three zero-energy internal modes (the rigid
rotations of the second block about the hinge) that the boundary stiffness
does not see; the indefinite block's K_ii is nonsingular with negative
eigenvalues, and the oracle still reduces it (LU instead of Cholesky)."""
import numpy as np

from oracle import reducer as orc
from singular_parts import hinged_blocks, indefinite_block


def _kii(p):
    s = orc.sam_numbering(p)
    K, M, _ = orc.assemble(p, s)
    i1, i2 = s["meqn1"] - 1, s["meqn2"] - 1
    return K.toarray(), i1, i2


def test_hinged_blocks_have_three_internal_mechanisms():
    K, i1, i2 = _kii(hinged_blocks())
    w, V = np.linalg.eigh(K[np.ix_(i1, i1)])
    nz = np.sum(np.abs(w) <= 1e-9 * np.abs(w).max())
    assert nz == 3
    # the mechanisms do not couple to the boundary: K_bi v = 0
    N = V[:, :3]
    assert np.abs(K[np.ix_(i2, i1)] @ N).max() <= 1e-6 * np.abs(K).max()


def test_indefinite_block_reduces_with_lu():
    p = indefinite_block()
    K, i1, i2 = _kii(p)
    w = np.linalg.eigvalsh(K[np.ix_(i1, i1)])
    assert w.min() < 0 < w.max() and np.abs(w).min() > 1e-8 * np.abs(w).max()
    o = orc.reduce(p)
    Kii, Kib, Kbb = K[np.ix_(i1, i1)], K[np.ix_(i1, i2)], K[np.ix_(i2, i2)]
    S = Kbb - Kib.T @ np.linalg.solve(Kii, Kib)
    pos = o["sam"]["dofPosIn2"] - 1
    assert np.allclose(o["SK"], S[np.ix_(pos, pos)], rtol=0, atol=1e-10 * np.abs(S).max())
