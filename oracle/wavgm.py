"""ORACLE — test infrastructure only (see oracle/__init__.py).

Weighted-average-motion (WAVGM, Nastran RBE3) constraint equations.

  omega        wavgmConstrEqn (wavgmModule.f90:58-367) with the reducer's
               default switches (reducerInterface.C:121-124: useRotCpl and
               useEccCpl on, useUnWeightedCG and useOldWAVGM off): the
               coefficients coupling one dependent DOF of the reference
               node to the 6*nM DOFs of the independent nodes.
  equations    MultiPointConstraints (rigidModule.f90:89-338) and
               defineConstraintEqn (rigidModule.f90:1575-1674): which
               dependent DOFs get an equation and which independent DOFs
               enter it, for all WAVGM elements of a part in order.

A WAVGM element here is a dict: ref (0-based reference node), refc (int,
e.g. 123456: the dependent components), nodes (0-based independent nodes),
indc (6 ints, 1-based start of the component's nM weights in `weights`,
<0 uniform weights, 0 no coupling) and weights.  With no weights given the
element takes the reference's defaults (rigidModule.f90:195-200): refc
123456, uniform translations, no rotational weights.

Parity is pinned by closed forms of rigid motions in tests/test_wavgm_cpu.py:
with explicit translational weights and no rotational ones the reference
node follows the rigid motion of its independent nodes exactly (this needs
the weighted centre, both rotation-coupling signs and both eccentricity
terms); uniform weights give the reference's 1/nM-scaled rotation
(wavgmModule.f90:132,150,203); rotational weights add their average on top
of the fitted rotation.
"""
import numpy as np

EPS_DIV0 = np.finfo(np.float64).eps  # epsDiv0_p (kindModule.f90:44)
TOL_WAVGM = 1.0e-4                   # tolWAVGM default (reducerInterface.C:99)


def _defaults(el):
    if el.get("weights") is None or len(el.get("weights", [])) == 0:
        return 123456, [-1, -1, -1, 0, 0, 0], np.zeros(0)
    return int(el["refc"]), [int(c) for c in el["indc"]], np.asarray(el["weights"], float)


def _seqsum(v):
    t = 0.0
    for x in v:
        t += x
    return t


class _Element:
    """Coordinates of one element and the centre/tolerance of a weight set
    (computePointCoords, wavgmModule.f90:381-418)."""

    def __init__(self, xyz, ref, nodes, weights, eps_x):
        self.X = np.asarray(xyz, float)[[ref] + list(nodes)]  # (1+nM, 3)
        self.nM = len(nodes)
        self.w = weights
        self.eps_x = eps_x
        self.cache = {}

    def centre(self, ic):
        """(dX (3, 1+nM), tolX (3)) for the weight set starting at ic (1-based;
        ic < 0: uniform weights)"""
        if ic not in self.cache:
            nM = self.nM
            # sequential sums in the reference's order (wavgmModule.f90:387-
            # 399): the centre of a planar set then lies exactly in its
            # plane, which decides terms at the epsDiv0_p cut
            W = self.w[ic - 1:ic - 1 + nM] if ic > 0 else np.ones(nM)
            X0 = np.zeros(3)
            for n in range(nM):
                for a in range(3):
                    X0[a] = X0[a] + self.X[1 + n, a] * W[n] if ic > 0 else X0[a] + self.X[1 + n, a]
            X0 = X0 / (_seqsum(W) if ic > 0 else float(nM))
            dX = (self.X - X0).T
            # the reference takes columns 1..nM (reference node and the first
            # nM-1 independent nodes) for the tolerance (wavgmModule.f90:408-411)
            tol = np.array([max(np.max(-dX[n, :nM]) * self.eps_x, np.max(dX[n, :nM]) * self.eps_x, EPS_DIV0)
                            for n in range(3)])
            self.cache[ic] = (dX, tol)
        return self.cache[ic]


def omega(el, xyz, ldof, eps_x=TOL_WAVGM):
    """Coefficients (6*nM) of dependent DOF ldof (1..6) of a WAVGM element."""
    refc, indc, W = _defaults(el)
    nodes = list(el["nodes"])
    nM = len(nodes)
    E = _Element(xyz, int(el["ref"]), nodes, W, eps_x)
    om = np.zeros(6 * nM)
    if ldof < 1 or ldof > 6:
        return om
    i = (ldof - 1) % 3 + 1
    j = i % 3 + 1
    k = j % 3 + 1

    def add(comp, coef):  # omega(comp :: 6) += coef (length nM)
        om[comp - 1::6] += coef

    def wset(c):
        return W[indc[c - 1] - 1:indc[c - 1] - 1 + nM]

    if indc[ldof - 1] > 0:
        s = _seqsum(wset(ldof))
        if s > EPS_DIV0:
            add(ldof, wset(ldof) / s)
    elif indc[ldof - 1] < 0:
        add(ldof, np.full(nM, 1.0 / nM))

    def fit(c, a, b, num):
        """(coefficient vector, denominator) of the least-squares rotation
        term: Sum w (dX_a^2 + dX_b^2) and w * dX_num, or their uniform forms"""
        dX, tol = E.centre(indc[c - 1])
        if indc[c - 1] > 0:
            w = wset(c)
            return w * dX[num - 1, 1:], _seqsum(w * (dX[a - 1, 1:] ** 2 + dX[b - 1, 1:] ** 2)), dX, tol
        return dX[num - 1, 1:].copy(), _seqsum(dX[a - 1, 1:] ** 2 + dX[b - 1, 1:] ** 2) * nM, dX, tol

    if ldof > 3:
        # rotation about axis i from the translations j, k
        # (wavgmModule.f90:115-166)
        if indc[j - 1] != 0:
            work, s, dX, tol = fit(j, j, k, k)
            if s > tol[j - 1] ** 2 + tol[k - 1] ** 2:
                add(j, -work / s)
        if indc[k - 1] != 0:
            work, s, dX, tol = fit(k, j, k, j)
            if s > tol[j - 1] ** 2 + tol[k - 1] ** 2:
                add(k, work / s)
    else:
        # translation from the rotation about axis k times the eccentricity
        # e_j, then the rotation about axis j times e_k
        # (wavgmModule.f90:168-335)
        tolWM = 0.0  # carried between blocks as in the reference
        if indc[i - 1] != 0:
            work, s, dX, tol = fit(i, i, j, j)
            if abs(dX[j - 1, 0]) > tol[j - 1]:
                tolWM = tol[i - 1] ** 2 + tol[j - 1] ** 2
                if s > tolWM:
                    add(i, dX[j - 1, 0] / s * work)
        if indc[j - 1] != 0:
            work, s, dX, tol = fit(j, i, j, i)
            if abs(dX[j - 1, 0]) > tol[j - 1]:
                if s > tolWM:
                    add(j, -dX[j - 1, 0] / s * work)
        if 3 + k <= 6 and indc[3 + k - 1] > 0:
            dX, tol = E.centre(indc[3 + k - 1])
            if abs(dX[j - 1, 0]) > tol[j - 1]:
                s = _seqsum(wset(3 + k))
                if s > EPS_DIV0:
                    add(3 + k, -dX[j - 1, 0] / s * wset(3 + k))
        if indc[i - 1] != 0:
            work, s, dX, tol = fit(i, k, i, k)
            if abs(dX[k - 1, 0]) > tol[k - 1]:
                tolWM = tol[k - 1] ** 2 + tol[i - 1] ** 2
                if s > tolWM:
                    add(i, dX[k - 1, 0] / s * work)
        if indc[k - 1] != 0:
            work, s, dX, tol = fit(k, k, i, i)
            if abs(dX[k - 1, 0]) > tol[k - 1]:
                tolWM = tol[k - 1] ** 2 + tol[i - 1] ** 2
                if s > tolWM:
                    add(k, -dX[k - 1, 0] / s * work)
        if 3 + j <= 6 and indc[3 + j - 1] > 0:
            dX, tol = E.centre(indc[3 + j - 1])
            if abs(dX[k - 1, 0]) > tol[k - 1]:
                s = _seqsum(wset(3 + j))
                if s > EPS_DIV0:
                    add(3 + j, dX[k - 1, 0] / s * wset(3 + j))
    return om


def _components(refc):
    """dependent components (procDofComp, rigidModule.f90:1694-1730): the
    digits 1..6 of refc in six decimal places, all six for refc = 0"""
    if refc < 0:
        raise ValueError("explicit constraint equations (negative refC) are not supported")
    if refc == 0:
        return set(range(1, 7))
    comps, x = set(), refc
    for div in (100000, 10000, 1000, 100, 10, 1):
        d = x // div
        if 1 <= d <= 6:
            comps.add(d)
        elif d != 0:
            raise ValueError(f"invalid dependent components {refc}")
        x -= d * div
    return comps


def equations(elements, xyz, madof, nndof, msc):
    """Constraint equations of the WAVGM elements, in element order.

    madof (1-based pointers), nndof per node and msc (status per DOF at the
    time the constraints are set up: 1 internal, 2 external, 0 fixed) as in
    initiateSAM.  msc is updated in place: a dependent DOF with an equation
    becomes -3, one whose first independent DOFs are all fixed and that
    couples to no free one becomes 0.  Returns {dependent DOF: [(DOF,
    coefficient), ...]} with 0-based DOF indices."""
    eqs = {}
    for el in elements:
        refc, indc, W = _defaults(el)
        ref = int(el["ref"])
        nodes = list(el["nodes"])
        comps = _components(refc)
        for ld in range(1, 7):
            if ld not in comps:
                continue
            if ld > nndof[ref]:
                continue  # the reference node has no such DOF
            dep = madof[ref] - 1 + ld - 1
            if msc[dep] != 1:
                continue  # external, fixed or already dependent
            om = omega(el, xyz, ld)
            terms = []
            for a, n in enumerate(nodes):
                for kk in range(nndof[n]):
                    c = om[6 * a + kk]
                    if abs(c) > EPS_DIV0:
                        terms.append((madof[n] - 1 + kk, c))
            if not terms:
                continue
            # defineConstraintEqn: free DOFs enter the equation; a fixed one
            # first in line fixes the dependent DOF until a free one follows;
            # others (dependent in an earlier element) are dropped
            first, out = True, []
            for d, c in terms:
                if msc[d] > 0:
                    out.append((d, c))
                    first = False
                    msc[dep] = -3
                elif msc[d] == 0 and first:
                    msc[dep] = 0
            if out:
                eqs[dep] = out
    return eqs
