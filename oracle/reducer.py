"""ORACLE — test infrastructure only (see oracle/__init__.py).

CPU restatement of the reducer's numbering, assembly and CMS reduction:

  sam_numbering     samReducerModule.f90:227-569 (nodal DOF counts, MADOF,
                    MSC), asmExtensionModule.f90:365-377 (unconnected
                    internal DOFs become fixed), :910-980 csCreateEqPartition,
                    samModule.f90:892-991 invertEqPartition.  MEQN itself comes
                    from SAM's SPRPRP, absent from the reference tree: parity
                    unpinned, the rule used is status-1 equations first then
                    status-2, each in DOF order (csCreateEqPartition's
                    expectation, asmExtensionModule.f90:936-947).
  assemble          lhsINADD element loop + csAddEM (inadd.f90:103-282)
  reduce            CMSTRS (cmstrs.f90:511-1062), GRAVfull/GRAV
                    (cmstrs.f90:286-474), EIGVAL (cmstrs.f90:53-263) with the
                    dense LAPACK path of the reference (DSYGVX, used when
                    tolEigval <= 0) or a shift-invert Lanczos (scipy eigsh)
                    for large parts; the small dense reducer of the
                    reference tests (smallReducer.f90:150-225) is the same
                    algebra: B = -K_ii^-1 K_ie, K = K_ee + K_ie^T B,
                    M = M_ee + M_ie^T B + B^T M_ie + B^T M_ii B.
"""
import ctypes
import os

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import wavgm

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so missing: make -C oracle")
        _lib = ctypes.CDLL(path)
        D = ctypes.POINTER(ctypes.c_double)
        _lib.hex8_stiff.argtypes = [D, D, D, ctypes.c_double, ctypes.c_double, D]
        _lib.hex8_mass.argtypes = [D, D, D, ctypes.c_double, D]
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def hex8_matrices(X, Y, Z, E, nu, rho):
    """(K, M) 24x24 of one HEX8 element (Femlib HEXA31 NIP=3 and HEXA35)."""
    X, Y, Z = (np.ascontiguousarray(v, np.float64) for v in (X, Y, Z))
    K = np.zeros(576)
    M = np.zeros(576)
    rc = lib().hex8_stiff(_dp(X), _dp(Y), _dp(Z), E, nu, _dp(K))
    if rc != 0:
        raise ValueError(f"hex8_stiff error {rc}")
    lib().hex8_mass(_dp(X), _dp(Y), _dp(Z), rho, _dp(M))
    return K.reshape(24, 24, order="F"), M.reshape(24, 24, order="F")


def andes4_matrices(xyz, E, nu, th, rho):
    """(K, M) 24x24 of one ANDES4 shell (ADD_ANDES4, inadd.f90:1407-1552):
    P^T K P of the rotated local ANDES stiffness, FQS35 lumped diagonal
    mass.  th: the four nodal thicknesses (their mean is the element's)."""
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(4, 3)
    th = np.ascontiguousarray(th, np.float64)
    t = (th[0] + th[1] + th[2] + th[3]) / 4.0
    K = np.zeros(576)
    rc = lib().andes4_stiff(_dp(xyz.ravel().copy()), ctypes.c_double(E), ctypes.c_double(nu), ctypes.c_double(t),
                            _dp(K))
    if rc != 0:
        raise ValueError(f"andes4_stiff error {rc}")
    em = np.zeros(24)
    X, Y, Z = (np.ascontiguousarray(xyz[:, c]) for c in range(3))
    lib().fqs35(_dp(X), _dp(Y), _dp(Z), _dp(th), ctypes.c_double(rho), _dp(em))
    return K.reshape(24, 24, order="F"), np.diag(em)


def beam_matrices(xyz, pr):
    """(K, M) 12x12 of one beam (ADD11 -> BEAM31/BEAM35, inadd.f90:829-930).
    pr: property record (E, nu, rho, G, A, I1, I2, J, v(3), K1, K2, NSM)
    with the Nastran PBAR/CBAR meaning; I1 bends in the local x-y plane
    (the plane of v), the shear factors kappa = 1/K (0: no shear
    deformation), the density carries the non-structural mass.  The
    section mapping (I1 -> I_y of BELS31, kappa = 1/K) is the one the
    CBAR02_ known answer selects among the alternatives."""
    x1 = np.ascontiguousarray(xyz[0], np.float64)
    x2 = np.ascontiguousarray(xyz[1], np.float64)
    v = np.ascontiguousarray(pr[8:11], np.float64)
    A = pr[4]
    Iy, Iz = pr[5], pr[6]
    kap = lambda k: 1.0 / k if k > 0 else 0.0
    rho = pr[2] + (pr[13] / A if A > 1e-16 else 0.0)
    K = np.zeros(144)
    M = np.zeros(144)
    c = ctypes.c_double
    rc = lib().beam_stiff(_dp(x1), _dp(x2), _dp(v), c(pr[0]), c(pr[3]), c(A), c(Iy), c(Iz), c(pr[7]),
                          c(kap(pr[11])), c(kap(pr[12])), _dp(K))
    if rc != 0:
        raise ValueError(f"beam_stiff error {rc}")
    em = np.zeros(1)
    lib().beam_mass(_dp(x1), _dp(x2), _dp(v), c(A), c(rho), c(pr[5] + pr[6]), _dp(M), _dp(em))
    return K.reshape(12, 12, order="F"), M.reshape(12, 12, order="F"), float(em[0])


ELEM_NODES = {44: 8, 24: 4, 11: 2}
ELEM_NDN = {44: 3, 24: 6, 11: 6}


def sam_numbering(part):
    status = np.asarray(part["status"])
    melcon = np.asarray(part["melcon"])
    mp = np.asarray(part["mpmnpc"]) - 1
    mm = np.asarray(part["mmnpc"]) - 1
    nnod = len(status)
    maxdn = np.zeros(nnod, int)
    for e in range(len(melcon)):
        for n in mm[mp[e]:mp[e + 1]]:
            maxdn[n] = max(maxdn[n], ELEM_NDN[int(melcon[e])])
    nndof = np.where((status == 1) | (maxdn == 0), 6, maxdn)
    madof = np.concatenate([[1], 1 + np.cumsum(nndof)])
    ndof = int(madof[-1] - 1)
    msc = np.zeros(ndof, int)
    for i in range(nnod):
        for j in range(nndof[i]):
            st = int(status[i])
            if st < 0 and ((-st) >> j) & 1:
                m = 0
            elif st == 1:
                m = 2
            else:
                m = 1
            msc[madof[i] - 1 + j] = m
    # WAVGM (RBE3) elements first, on the nodal status codes, then the rigid
    # elements (samReducerModule.f90:385-459)
    wv = wavgm.equations(part.get("rbe3", []), part["xyz"], madof, nndof, msc)
    dep = {}
    for gn, cm, deps in part.get("rbe2", []):
        for dn in deps:
            for c in cm:
                j = int(c) - 1
                if j < nndof[dn]:
                    if msc[madof[dn] - 1 + j] < 0:
                        raise ValueError("DOF constrained by two constraint elements")
                    dep[(dn, j)] = gn
    indep = {g for g, _, _ in part.get("rbe2", [])} | {int(n) for el in part.get("rbe3", []) for n in el["nodes"]}
    for i in range(nnod):
        for j in range(nndof[i]):
            d = madof[i] - 1 + j
            m = msc[d]
            if m == 1 and j >= maxdn[i] and i not in indep:
                m = 0  # unconnected internal DOF
            if m == -3 or (i, j) in dep:
                m = -1  # dependent DOF of a constraint element: no equation
            msc[d] = m
    for terms in wv.values():
        if any(msc[d] < 0 for d, _ in terms):
            raise ValueError("constraint chains (an independent DOF that is dependent elsewhere) are not supported")
    for (dn, j), gn in dep.items():
        if any(msc[madof[gn] - 1 + k] < 0 for k in range(nndof[gn])):
            raise ValueError("constraint chains (an independent DOF that is dependent elsewhere) are not supported")
    meqn = np.zeros(ndof, int)
    n1 = int(np.sum(msc == 1))
    meqn[msc == 1] = np.arange(1, n1 + 1)
    meqn[msc == 2] = np.arange(n1 + 1, n1 + 1 + int(np.sum(msc == 2)))
    neq = int(np.count_nonzero(meqn))
    status_of_eq = np.zeros(neq + 1, int)
    status_of_eq[meqn[msc > 0]] = msc[msc > 0]
    meqn1 = np.nonzero(status_of_eq == 1)[0]
    meqn2 = np.nonzero(status_of_eq == 2)[0]
    pos2 = np.zeros(neq + 1, int)
    pos2[meqn2] = np.arange(1, len(meqn2) + 1)
    dofpos2 = pos2[meqn[msc == 2]]
    ndim_ext = int(np.sum(nndof[status == 1]))
    return dict(madof=madof, msc=msc, meqn=meqn, meqn1=meqn1, meqn2=meqn2, dofPosIn2=dofpos2,
                nndof=nndof, neq=neq, ndof1=len(meqn1), ndof2=len(meqn2), ndim_ext=ndim_ext, dep=dep, wavgm=wv)


def constraint_T(part, s):
    """T (ndof x neq): u_dof = T u_eq.  Free DOFs map to their equation,
    fixed DOFs to zero, dependent DOFs of RBE2 elements to the rigid-body
    motion of the independent node: u_d = u_i + theta_i x (x_d - x_i) for
    translations, theta_d = theta_i for rotations; dependent DOFs of WAVGM
    (RBE3) elements to their equations (oracle.wavgm)."""
    rows, cols, vals = [], [], []
    madof, meqn = s["madof"], s["meqn"]
    xyz = part["xyz"]
    for d in range(len(s["msc"])):
        if s["msc"][d] > 0:
            rows.append(d); cols.append(meqn[d] - 1); vals.append(1.0)
    for (dn, j), gn in s["dep"].items():
        d = madof[dn] - 1 + j
        eqi = [meqn[madof[gn] - 1 + k] - 1 if k < s["nndof"][gn] else -1 for k in range(6)]
        r = xyz[dn] - xyz[gn]
        if j < 3:
            coef = {j: 1.0}
            # theta x r: component j
            if j == 0: coef.update({4: r[2], 5: -r[1]})
            if j == 1: coef.update({5: r[0], 3: -r[2]})
            if j == 2: coef.update({3: r[1], 4: -r[0]})
        else:
            coef = {j: 1.0}
        for k, c in coef.items():
            if eqi[k] >= 0 and c != 0.0:
                rows.append(d); cols.append(eqi[k]); vals.append(c)
    for d, terms in s["wavgm"].items():
        for di, c in terms:
            rows.append(d); cols.append(meqn[di] - 1); vals.append(c)
    return sp.csr_matrix((vals, (rows, cols)), shape=(len(s["msc"]), s["neq"]))


def element_matrices(part, e):
    t = int(part["melcon"][e])
    mp = part["mpmnpc"]
    nodes = part["mmnpc"][mp[e] - 1:mp[e + 1] - 1] - 1
    x = part["xyz"][nodes]
    pr = part["eprop"][e]
    if t == 44:
        K, M = hex8_matrices(x[:, 0], x[:, 1], x[:, 2], pr[0], pr[1], pr[2])
        return nodes, 3, K, M
    if t == 24:
        K, M = andes4_matrices(x, pr[0], pr[1], pr[4:8], pr[2])
        return nodes, 6, K, M
    if t == 11:
        K, M, _ = beam_matrices(x, pr)
        return nodes, 6, K, M
    raise NotImplementedError(f"element type {t}")


def assemble(part, s):
    """Global K, M (CSR, 0-based equation order) and the element masses:
    element matrices are added in DOF space, then K = T^T K_dof T with the
    constraint transformation T (identity on the free DOFs)."""
    rows, cols, kv, mv = [], [], [], []
    emass = 0.0
    for e in range(len(part["melcon"])):
        nodes, ndn, K, M = element_matrices(part, e)
        dof = np.full(len(nodes) * ndn, -1)
        for a, n in enumerate(nodes):
            for j in range(ndn):
                if j < s["nndof"][n]:
                    dof[a * ndn + j] = s["madof"][n] - 1 + j
        ok = dof >= 0
        ii, jj = np.meshgrid(dof[ok], dof[ok], indexing="ij")
        rows.append(ii.ravel())
        cols.append(jj.ravel())
        kv.append(K[np.ix_(ok, ok)].ravel())
        mv.append(M[np.ix_(ok, ok)].ravel())
        emass += element_mass(part, e, M)
    nd = len(s["msc"])
    r, c = np.concatenate(rows), np.concatenate(cols)
    Kd = sp.csr_matrix((np.concatenate(kv), (r, c)), shape=(nd, nd))
    Md = sp.csr_matrix((np.concatenate(mv), (r, c)), shape=(nd, nd))
    T = constraint_T(part, s)
    return (T.T @ Kd @ T).tocsr(), (T.T @ Md @ T).tocsr(), emass


def element_mass(part, e, M):
    """EMASS of the element (inadd.f90: lumpSolidMass for solids)."""
    t = int(part["melcon"][e])
    if t == 44:
        return float(M[:, 0::3].sum())
    if t == 24:  # dsum(NEDOF, EMV, 6): x-translation masses of the lumped diagonal
        return float(np.diag(M)[0::6].sum())
    if t == 11:  # EM(1,1) + EM(7,1) of the diagonal form: rho A L
        mp = part["mpmnpc"]
        nodes = part["mmnpc"][mp[e] - 1:mp[e + 1] - 1] - 1
        return beam_matrices(part["xyz"][nodes], part["eprop"][e])[2]
    return 0.0


def gravity_dirs(part, s):
    a = np.zeros((s["neq"], 3))
    for i in range(len(part["status"])):
        for j in range(min(3, s["nndof"][i])):
            q = s["meqn"][s["madof"][i] - 1 + j]
            if q > 0:
                a[q - 1, j] = 1.0
    return a


def fix_signs(phi, eq_order=None):
    """Largest-magnitude component of each mode positive (first in equation
    order on exact ties)."""
    phi = phi.copy()
    for i in range(phi.shape[1]):
        k = int(np.argmax(np.abs(phi[:, i])))
        if phi[k, i] < 0:
            phi[:, i] = -phi[:, i]
    return phi


def modes(Kii, Mii, nev, dense_limit=3000, solve=None):
    """Lowest nev eigenpairs of Kii phi = lambda Mii phi, M-orthonormal
    (shift-invert about 0, reusing the factor of Kii when given)."""
    n = Kii.shape[0]
    if n <= dense_limit:
        w, V = sla.eigh(Kii.toarray(), Mii.toarray(), subset_by_index=[0, nev - 1])
    else:
        opinv = None
        if solve is not None:
            opinv = spla.LinearOperator(Kii.shape, matvec=lambda x: solve(x), dtype=np.float64)
        w, V = spla.eigsh(Kii.tocsc(), k=nev, M=Mii.tocsc(), sigma=0.0, which="LM", tol=0, OPinv=opinv)
        o = np.argsort(w)
        w, V = w[o], V[:, o]
        nrm = np.sqrt(np.einsum("ij,ij->j", V, Mii @ V))
        V = V / nrm
    return w, fix_signs(V)


def refine_extended(Kii, R, X, solve, steps):
    """steps of iterative refinement of X ~ Kii^-1 R with the residual
    R - Kii X accumulated in extended precision (numpy longdouble, 64-bit
    mantissa on x86): the solution's error falls to about the rounding of
    its own entries, well below what any FP64 factorisation of an
    ill-conditioned K_ii gives by itself."""
    K = Kii.tocsr()
    data = K.data.astype(np.longdouble)
    rows = np.repeat(np.arange(K.shape[0]), np.diff(K.indptr))
    for _ in range(steps):
        Res = np.empty_like(X)
        for j in range(X.shape[1]):
            prod = data * X[K.indices, j].astype(np.longdouble)
            acc = np.zeros(K.shape[0], np.longdouble)
            np.add.at(acc, rows, prod)
            Res[:, j] = (R[:, j].astype(np.longdouble) - acc).astype(np.float64)
        X = X + solve(Res)
    return X


def reduce(part, ngen=0, neval=0, want_vgi=False, dense_limit=3000, loads=None, refine=0):
    """Reduced SK, SM, G (CMSTRS/GRAV layout), B, modes and vgi of a part.

    loads: full load vectors (neq x nrhs, equation order); out["L"] is then
    the reduced load matrix of cmstrs.f90:598-640,763-782,951-955:
    rows dofPosIn2 of f_e + B^T f_i, then Phi^T f_i, zero padded to ndim.
    refine: extended-precision refinement steps of B (refine_extended)."""
    s = sam_numbering(part)
    K, M, emass = assemble(part, s)
    i1 = s["meqn1"] - 1
    i2 = s["meqn2"] - 1
    Kii, Kib, Kbb = K[i1][:, i1], K[i1][:, i2], K[i2][:, i2].toarray()
    Mii, Mib, Mbb = M[i1][:, i1], M[i1][:, i2], M[i2][:, i2].toarray()
    n1, n2 = len(i1), len(i2)
    if n1 <= dense_limit:
        try:
            lu = sla.cho_factor(Kii.toarray())
            solve = lambda R: sla.cho_solve(lu, R)
        except np.linalg.LinAlgError:  # indefinite K_ii: LU with partial pivoting
            lu = sla.lu_factor(Kii.toarray())
            solve = lambda R: sla.lu_solve(lu, R)
    else:
        # K_ii is SPD: symmetric minimum-degree ordering, no pivoting
        f = spla.splu(Kii.tocsc(), permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0,
                      options={"SymmetricMode": True})
        solve = lambda R: f.solve(np.asarray(R))
    B = -solve(Kib.toarray()) if n1 else np.zeros((0, n2))
    if refine and n1:
        B = refine_extended(Kii, -Kib.toarray(), B, solve, refine)
    S22 = Kbb + Kib.T @ B
    S22 = 0.5 * (S22 + S22.T)
    nev = max(neval, ngen) if ngen > 0 else 0
    lam = phi = None
    if nev > 0:
        lam, phi = modes(Kii, Mii, nev, dense_limit, solve)
    MB = Mii @ B
    M22 = B.T @ MB + Mib.T @ B + B.T @ Mib + Mbb
    M22 = 0.5 * (M22 + M22.T)
    a = gravity_dirs(part, s)
    g = M @ a
    Ge = B.T @ g[i1] + g[i2]
    ndim = s["ndim_ext"] + ngen
    pos = s["dofPosIn2"] - 1
    SK = np.zeros((ndim, ndim))
    SM = np.zeros((ndim, ndim))
    G = np.zeros((ndim, 3))
    SK[:n2, :n2] = S22[np.ix_(pos, pos)]
    SM[:n2, :n2] = M22[np.ix_(pos, pos)]
    G[:n2] = Ge[pos]
    if ngen > 0:
        P = phi[:, :ngen]
        K11 = P.T @ (Kii @ P)
        M11 = P.T @ (Mii @ P)
        M12 = P.T @ MB + (Mib.T @ P).T
        SK[n2:n2 + ngen, n2:n2 + ngen] = 0.5 * (K11 + K11.T)
        SM[n2:n2 + ngen, n2:n2 + ngen] = 0.5 * (M11 + M11.T)
        SM[n2:n2 + ngen, :n2] = M12[:, pos]
        SM[:n2, n2:n2 + ngen] = M12[:, pos].T
        G[n2:n2 + ngen] = P.T @ g[i1]
    out = dict(SK=SK, SM=SM, G=G, B=B, S22=S22, sam=s, K=K, M=M, emass=emass, lam=lam, phi=phi)
    if loads is not None:
        F = np.asarray(loads, float).reshape(s["neq"], -1)
        L = np.zeros((ndim, F.shape[1]))
        L[:n2] = (B.T @ F[i1] + F[i2])[pos]
        if ngen > 0:
            L[n2:n2 + ngen] = phi[:, :ngen].T @ F[i1]
        out["L"] = L
    if want_vgi:
        out["vgi"] = solve(g[i1])
    return out


def solve_static(part, F):
    """Direct static solve of solver.f90:30-383: external nodes clamped
    (fixExternalDOFs, samReducerModule.f90:339), K u = F over all remaining
    equations.  F: (neq x nrhs) in the equation order of that numbering.
    Returns (u, sam numbering)."""
    p = dict(part)
    st = np.asarray(part["status"]).copy()
    st[st == 1] = -63
    p["status"] = st
    s = sam_numbering(p)
    K, _, _ = assemble(p, s)
    u = spla.spsolve(K.tocsc(), np.asarray(F, float))
    return u, s


def direct_solve_displacements(part, node_ids, gvec=None):
    """solver.f90:30-383 on a part read from bulk data: external nodes
    clamped, nodal loads of the first load case (plus M a for a gravity
    vector gvec), K u = F; returns the translations of the listed node ids
    (len x 3)."""
    st = np.asarray(part["status"]).copy()
    st[st == 1] = -63
    q = dict(part, status=st)
    s = sam_numbering(q)
    F = np.zeros(s["neq"])
    loads = part.get("loads") or []
    if loads:
        first = min(l[0] for l in loads)
        for sid, n, d0, f in loads:
            for k in range(3):
                if sid == first and d0 + k < s["nndof"][n]:
                    eq = s["meqn"][s["madof"][n] - 1 + d0 + k]
                    if eq > 0:
                        F[eq - 1] += f[k]
    if gvec is not None:
        _, M, _ = assemble(q, s)
        F = F + M @ gravity_dirs(q, s) @ np.asarray(gvec, float)
    u, _ = solve_static(part, F)
    ids = list(part["node_ids"]) if "node_ids" in part else list(range(1, len(st) + 1))
    out = np.zeros((len(node_ids), 3))
    for r, nid in enumerate(node_ids):
        i = ids.index(nid)
        for k in range(min(3, s["nndof"][i])):
            eq = s["meqn"][s["madof"][i] - 1 + k]
            out[r, k] = u[eq - 1] if eq > 0 else 0.0
    return out
