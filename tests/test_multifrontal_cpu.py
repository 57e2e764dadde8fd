"""The benchmark's CPU baseline (oracle/multifrontal.py: multifrontal
Cholesky over the library's host ordering, dense front work on BLAS) gives
the oracle's reduction: SK, SM, G at 1e-10 and the eigenvalues at 1e-9, on
a block with modes and on a shell plate.  The front structure comes from
the library's host analysis (no device)."""
import numpy as np
import pytest

from fedempy_amd import gpu, parts
from oracle import multifrontal as mf
from oracle import reducer as orc


@pytest.mark.parametrize("make,ngen", [
    (lambda: parts.hex_block(8, 6, 5, 0.08, 0.06, 0.05, seed=3, n_ext=6), 4),
    (lambda: parts.config2(seed=1, n=24, n_ext=12), 0),
])
def test_multifrontal_baseline_matches_oracle(make, ngen):
    p = make()
    R = gpu.GpuReducer()
    R.set_part(p)
    R.analyze(ngen=ngen)
    m = mf.reduce(p, R.supernodes(), ngen=ngen)
    o = orc.reduce(p, ngen=ngen)
    n2 = len(orc.sam_numbering(p)["meqn2"])
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(m["SK"][:n2, :n2], o["SK"][:n2, :n2]) <= 1e-10
    assert rel(m["SM"][:n2, :n2], o["SM"][:n2, :n2]) <= 1e-10
    assert rel(m["G"][:n2], o["G"][:n2]) <= 1e-10
    if ngen:
        assert np.max(np.abs(m["lam"] - o["lam"]) / o["lam"]) <= 1e-9


def test_multifrontal_served_request_matches_oracle():
    """the served-mix request on the CPU port (oracle/multifrontal.serve):
    reduced load vectors against the oracle's L (modal rows up to each mode's
    sign), internal displacements against a direct solve of
    K_ii u_i = f_i - K_ib u_b"""
    import scipy.sparse.linalg as spla
    p = parts.hex_block(7, 5, 4, 0.07, 0.05, 0.04, seed=5, n_ext=5)
    rng = np.random.default_rng(2)
    loads = parts.random_loads(rng, p, 3)
    s = orc.sam_numbering(p)
    n1, n2 = s["ndof1"], s["ndof2"]
    ub = rng.normal(0.0, 1e-3, (n2, 3))
    R = gpu.GpuReducer()
    R.set_part(p)
    R.analyze(ngen=3)
    r, red, Ui = mf.serve(p, R.supernodes(), 3, loads, ub)
    F = np.zeros((s["neq"], 3))
    for c, case in enumerate(loads):
        for node, dof, val in case:
            d = s["madof"][node - 1] - 1 + dof - 1
            if dof <= s["nndof"][node - 1] and s["meqn"][d] > 0:
                F[s["meqn"][d] - 1, c] += val
    o = orc.reduce(p, ngen=3, loads=F)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(red[:n2], o["L"][:n2]) <= 1e-10
    for k in range(3):
        a, b = red[n2 + k], o["L"][n2 + k]
        assert min(rel(a, b), rel(-a, b)) <= 1e-8
    # internal displacements in perm order against the direct solve
    K = o["K"]
    i1, i2 = s["meqn1"] - 1, s["meqn2"] - 1
    ube = np.zeros((n2, 3))
    ube[s["dofPosIn2"] - 1] = ub  # reduced order -> MEQN2 order
    U = spla.splu(K[np.ix_(i1, i1)].tocsc()).solve(F[i1] - K[np.ix_(i1, i2)] @ ube)
    p2e = np.asarray(R.supernodes()["perm_to_eq"])
    assert rel(Ui, U[p2e[:n1]]) <= 1e-10  # perm row r is equation p2e[r]
