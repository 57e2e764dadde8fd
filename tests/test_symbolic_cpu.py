"""Host-side parts of the product, through the C ABI, on the CPU: SAM
numbering bit-exact against the oracle's restatement, and the symbolic
(supernode) structure proven by a CPU multifrontal factorisation driven by
exactly the fronts the device uses."""
import numpy as np
import pytest

from fedempy_amd import gpu, parts
from oracle import reducer as orc


def small_parts():
    p1 = parts.hex_block(6, 4, 3, 0.06, 0.04, 0.03, seed=3, n_ext=4)
    p2 = parts.hex_block(5, 5, 2, 0.05, 0.05, 0.02, seed=5, n_ext=6)
    # fixed DOFs: clamp the x=0 face (bitmask of the three translations)
    x0 = np.nonzero(p2["xyz"][:, 0] == 0.0)[0]
    st = p2["status"].copy()
    st[x0[st[x0] == 0]] = -7
    p2["status"] = st
    # no boundary at all (the direct solve clamps the external nodes)
    p3 = dict(p2, status=np.where(st == 1, -63, st).astype(np.int32))
    return [p1, p2, p3]


@pytest.mark.parametrize("k", [0, 1, 2])
def test_sam_arrays_bit_exact(k):
    part = small_parts()[k]
    R = gpu.GpuReducer()
    R.set_part(part)
    z = R.analyze(ngen=2)
    a = R.sam()
    o = orc.sam_numbering(part)
    for key in ("madof", "msc", "meqn", "meqn1", "meqn2", "dofPosIn2"):
        assert np.array_equal(np.asarray(a[key]), np.asarray(o[key])), key
    assert z["ndof1"] == o["ndof1"] and z["ndof2"] == o["ndof2"]
    assert z["ndim"] == o["ndim_ext"] + 2


@pytest.mark.parametrize("k", [0, 1, 2])
def test_fronts_reproduce_dense_schur(k):
    check_fronts(small_parts()[k])


def check_fronts(part):
    R = gpu.GpuReducer()
    R.set_part(part)
    z = R.analyze()
    sn = R.supernodes()
    o = orc.sam_numbering(part)
    K, _, _ = orc.assemble(part, o)
    Kd = K.toarray()
    pe = sn["perm_to_eq"]
    Kp = Kd[np.ix_(pe, pe)]
    n1, n2, neq = z["ndof1"], z["ndof2"], z["neq"]
    nsup = len(sn["ncol"])
    rptr = np.concatenate([[0], np.cumsum(sn["nrow"])])
    colsn = np.full(neq, nsup - 1)
    for q in range(nsup - 1):
        colsn[sn["col0"][q]:sn["col0"][q] + sn["ncol"][q]] = q
    # every stored entry of K falls inside the front of its column
    rows_of = [set(sn["rows"][rptr[q]:rptr[q + 1]]) for q in range(nsup)]
    ii, jj = np.nonzero(np.tril(Kp))
    assert all(i in rows_of[colsn[j]] for i, j in zip(ii, jj))
    children = {q: [] for q in range(nsup)}
    for q in range(nsup - 1):
        assert sn["parent"][q] > q
        children[sn["parent"][q]].append(q)
    CB = {}
    for q in range(nsup):
        rows = sn["rows"][rptr[q]:rptr[q + 1]]
        nc, c0 = sn["ncol"][q], sn["col0"][q]
        cols = np.arange(c0, c0 + nc) if q < nsup - 1 else np.arange(n1, n1 + n2)
        F = np.tril(Kp[np.ix_(rows, cols)], 0) if len(cols) else np.zeros((len(rows), 0))
        Fq = np.zeros((len(rows), len(rows)))
        Fq[:, :len(cols)] = F
        pos = {r: i for i, r in enumerate(rows)}
        for c in children[q]:
            crows, cb = CB.pop(c)
            idx = [pos[r] for r in crows]
            Fq[np.ix_(idx, idx)] += np.tril(cb)
        A = np.tril(Fq) + np.tril(Fq, -1).T
        if q == nsup - 1:
            S = A
            break
        A11, A21 = A[:nc, :nc], A[nc:, :nc]
        CB[q] = (rows[nc:], A[nc:, nc:] - A21 @ np.linalg.solve(A11, A21.T))
    i1, i2 = o["meqn1"] - 1, o["meqn2"] - 1
    Sd = Kd[np.ix_(i2, i2)] - Kd[np.ix_(i2, i1)] @ np.linalg.solve(Kd[np.ix_(i1, i1)], Kd[np.ix_(i1, i2)])
    assert np.linalg.norm(S - Sd) <= 1e-12 * np.linalg.norm(Sd)


@pytest.mark.parametrize("name", ["CHEXA08_", "CQUAD04_", "CBAR02_"])
def test_sam_bit_exact_with_rigid_elements(name, golden_dir):
    """Reference parts with RBE2 elements: the library's SAM arrays (MSC -1
    on dependent DOFs) equal the oracle's, and the fronts reproduce the dense
    Schur complement of the constrained stiffness T^T K T."""
    import os
    from oracle import nastran
    part = nastran.read(os.path.join(golden_dir, name + ".nas"))
    R = gpu.GpuReducer()
    R.set_part(part)
    z = R.analyze()
    a = R.sam()
    o = orc.sam_numbering(part)
    for key in ("madof", "msc", "meqn", "meqn1", "meqn2", "dofPosIn2"):
        assert np.array_equal(np.asarray(a[key]), np.asarray(o[key])), key
    assert z["ndof1"] == o["ndof1"] and z["ndof2"] == o["ndof2"]
    check_fronts(part)
