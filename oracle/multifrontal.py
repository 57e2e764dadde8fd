"""ORACLE — test infrastructure only (see oracle/__init__.py): the CPU
baseline of the benchmark, not a parity reference.

The reduction of oracle.reducer with the reference's SPR algorithm on the
CPU (csSolveSupEl, solExtensionModule.f90:708-775, with the boundary
equations ordered last so that the trailing front is the Schur complement):
a multifrontal Cholesky of K_ii over a given front structure (ordering and
supernodes, e.g. the device library's host analysis), with every dense
front operation in numpy/LAPACK/BLAS, which uses all host cores; B by a
multi-right-hand-side solve on the fronts; the fixed-interface modes by
ARPACK shift-invert with the front solve; mass, stiffness and gravity
reduction as in oracle.reducer.reduce.
"""
import numpy as np
import scipy.linalg as sla
import scipy.sparse.linalg as spla

from . import reducer as orc


class Fronts:
    """Multifrontal Cholesky K_ii = L L^T over the fronts of `sn` (postorder,
    the last one the Schur front): col0/ncol/nrow/parent per front, rows
    (perm indices of every front's rows, concatenated), perm_to_eq.  Kp is
    the full matrix in perm order (CSC).  The inverse of every pivot block
    is formed, so that both triangular solves are GEMMs, as on the device."""

    def __init__(self, Kp, sn, n1):
        self.n1 = n1
        col0, ncol, nrow = sn["col0"], sn["ncol"], sn["nrow"]
        rptr = np.concatenate([[0], np.cumsum(nrow)])
        self.fronts = []
        upd = {}
        kids = {}
        for f, p in enumerate(sn["parent"]):
            if p >= 0:
                kids.setdefault(int(p), []).append(f)
        for f in range(len(nrow)):
            rows = sn["rows"][rptr[f]:rptr[f + 1]]
            nc = int(ncol[f])
            if nc:
                F = np.zeros((len(rows), len(rows)), order="F")
                c0 = int(col0[f])
                F[:, :nc] = Kp[:, c0:c0 + nc][rows].toarray()
            else:  # the Schur front: the boundary block itself
                F = np.asfortranarray(Kp[rows][:, rows].toarray())
            for c in kids.get(f, []):
                crow, U = upd.pop(c)
                rel = np.searchsorted(rows, crow)
                if len(rel) <= 128:
                    F[np.ix_(rel, rel)] += U
                else:  # lower triangle only, column by column
                    for j in range(len(rel)):
                        F[rel[j:], rel[j]] += U[j:, j]
            if nc == 0:  # fronts hold their lower triangles
                self.schur = np.tril(F) + np.tril(F, -1).T
                continue
            L11 = sla.cholesky(F[:nc, :nc], lower=True, check_finite=False)
            # L21 = F21 L11^-T and L11^-1 by triangular solves (BLAS 3)
            L21 = sla.blas.dtrsm(1.0, L11, F[nc:, :nc], side=1, lower=1, trans_a=1) if len(rows) > nc else \
                np.zeros((0, nc))
            Li = sla.blas.dtrsm(1.0, L11, np.eye(nc), lower=1)
            if len(rows) > nc:  # lower triangle of F22 - L21 L21^T
                upd[f] = (rows[nc:], sla.blas.dsyrk(-1.0, L21, beta=1.0, c=F[nc:, nc:], lower=1))
            m = rows[nc:] < n1
            self.fronts.append((int(col0[f]), nc, rows[nc:][m], Li, np.ascontiguousarray(L21[m])))

    def solve(self, R):
        """X = K_ii^-1 R (perm order, n1 rows, any number of columns)"""
        R = np.asarray(R, float)
        vec = R.ndim == 1
        Z = R.reshape(self.n1, -1).copy()
        for c0, nc, urows, Li, L21 in self.fronts:
            y = Li @ Z[c0:c0 + nc]
            Z[c0:c0 + nc] = y
            if len(urows):
                Z[urows] -= L21 @ y
        for c0, nc, urows, Li, L21 in reversed(self.fronts):
            rhs = Z[c0:c0 + nc]
            if len(urows):
                rhs = rhs - L21.T @ Z[urows]
            Z[c0:c0 + nc] = Li.T @ rhs
        return Z.ravel() if vec else Z


def reduce(part, sn, ngen=0, keep=False):
    """SK, SM, G of the part (as oracle.reducer.reduce) through the fronts
    of `sn`; eigenvalues when ngen > 0.  keep: also the fronts, B, the modes
    and the numbering, for load cases on the same factor (serve)."""
    s = orc.sam_numbering(part)
    K, M, _ = orc.assemble(part, s)
    p2e = np.asarray(sn["perm_to_eq"])
    n1, n2 = s["ndof1"], s["ndof2"]
    Kp = K[p2e][:, p2e].tocsc()
    Mp = M[p2e][:, p2e].tocsr()
    fr = Fronts(Kp, sn, n1)
    B = fr.solve(-Kp[:n1, n1:].toarray())
    S22 = fr.schur
    lam = phi = None
    if ngen > 0:
        Mii = Mp[:n1, :n1]
        op = spla.LinearOperator((n1, n1), matvec=fr.solve, dtype=np.float64)
        lam, phi = spla.eigsh(Kp[:n1, :n1], k=ngen, M=Mii, sigma=0.0, which="LM", OPinv=op, tol=1e-12)
        o = np.argsort(lam)
        lam, phi = lam[o], phi[:, o]
    Mii, Mib, Mbb = Mp[:n1, :n1], Mp[:n1, n1:].toarray(), Mp[n1:, n1:].toarray()
    MB = Mii @ B
    M22 = B.T @ MB + Mib.T @ B + B.T @ Mib + Mbb
    g = Mp @ orc.gravity_dirs(part, s)[p2e]
    Ge = B.T @ g[:n1] + g[n1:]
    # boundary perm rows are MEQN2 order; reduced layout through dofPosIn2
    pos = s["dofPosIn2"] - 1
    nd = s["ndim_ext"] + ngen
    SK, SM, G = np.zeros((nd, nd)), np.zeros((nd, nd)), np.zeros((nd, 3))
    SK[:n2, :n2] = S22[np.ix_(pos, pos)]
    SM[:n2, :n2] = (0.5 * (M22 + M22.T))[np.ix_(pos, pos)]
    G[:n2] = Ge[pos]
    if ngen > 0:
        Kii = Kp[:n1, :n1]
        SK[n2:, n2:] = phi.T @ (Kii @ phi)
        SM[n2:, n2:] = phi.T @ (Mii @ phi)
        M12 = phi.T @ MB + (Mib.T @ phi).T
        SM[n2:, :n2] = M12[:, pos]
        SM[:n2, n2:] = M12[:, pos].T
        G[n2:] = phi.T @ g[:n1]
    out = dict(SK=SK, SM=SM, G=G, lam=lam)
    if keep:
        out.update(fronts=fr, B=B, phi=phi, s=s, p2e=p2e, pos=pos)
    return out


def serve(part, sn, ngen, loads, ub):
    """One request of the served mix on the CPU: the reduction, then for
    every load case the reduced load vector [f_b + B^T f_i (dofPosIn2 order);
    Phi^T f_i] and the internal displacements B u_b + K_ii^-1 f_i from the
    boundary displacements ub (reduced order), all on the same fronts.
    loads: per case a list of (node 1-based, dof 1-based, value)."""
    r = reduce(part, sn, ngen, keep=True)
    s, p2e, pos, B, fr, phi = r["s"], r["p2e"], r["pos"], r["B"], r["fronts"], r["phi"]
    n1, n2, nlc = s["ndof1"], s["ndof2"], len(loads)
    F = np.zeros((s["neq"], nlc))
    for c, case in enumerate(loads):
        for node, dof, val in case:
            d = s["madof"][node - 1] - 1 + dof - 1
            if dof <= s["nndof"][node - 1] and s["meqn"][d] > 0:
                F[s["meqn"][d] - 1, c] += val
    Fp = F[p2e]
    fred = Fp[n1:] + B.T @ Fp[:n1]
    red = np.zeros((n2 + ngen, nlc))
    red[:n2] = fred[pos]
    if ngen:
        red[n2:] = phi.T @ Fp[:n1]
    ubp = np.zeros((n2, nlc))
    ubp[pos] = np.asarray(ub)[:n2]
    Ui = B @ ubp + fr.solve(Fp[:n1])
    return r, red, Ui
