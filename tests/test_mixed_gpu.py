"""Mixed-precision reduction on the GPU (option "mixedPrecision"): FP32
fronts of the equilibrated stiffness with FP32 MFMA trailing and Schur
updates and FP64 diagonal-block factorisation, every K_ii solve refined
against the FP64 matrix.

Tolerances (mixed path): reduced matrices, B, load vectors and vgi within
1e-8 relative Frobenius error of the FP64 oracle, eigenvalues within 1e-9
relative.  The error of every output is printed, so a run states it."""
import numpy as np
import pytest

from fedempy_amd import gpu, parts
from oracle import reducer as orc

pytestmark = pytest.mark.gpu
TOL = 1e-8


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def align(phi, ref):
    s = np.sign(np.einsum("ij,ij->j", phi, ref))
    s[s == 0] = 1
    return phi * s


def beam():
    return parts.beam_frame(n_el=200, seed=5, n_ext=4)


CASES = [
    ("plate60_lognormal_t", lambda: parts.config2(seed=3, n=60, n_ext=16), 0),
    ("block10x7x6", lambda: parts.hex_block(10, 7, 6, 0.10, 0.07, 0.06, seed=2, n_ext=6), 6),
    ("plate30_modes", lambda: parts.config2(seed=4, n=30, n_ext=8), 5),
    ("beam_frame", beam, 4),
]


@pytest.mark.parametrize("name,make,ngen", CASES)
def test_mixed_reduction_matches_oracle(name, make, ngen):
    part = make()
    print("mixed precision: %s, ngen %d" % (name, ngen), flush=True)
    R = gpu.GpuReducer()
    R.set_part(part)
    R.set_option("mixedPrecision", 1)
    R.analyze(ngen=ngen, vgi=True)
    R.reduce()
    SK, SM, G = R.reduced()
    steps = int(R.times()["refine_steps"])
    o = orc.reduce(part, ngen=ngen, want_vgi=True)
    err = dict(SK=rel(SK, o["SK"]), SM=rel(SM, o["SM"]), G=rel(G, o["G"]), B=rel(R.B(), o["B"]),
               vgi=rel(R.vgi(), o["vgi"]))
    if ngen > 0:
        lam, _ = R.eigen()
        err["lambda"] = float(np.max(np.abs(lam - o["lam"]) / o["lam"]))
        err["phi"] = rel(align(R.phi(), o["phi"]), o["phi"])
    print("  refinement steps %d; errors vs FP64 oracle: %s" % (
        steps, ", ".join("%s %.1e" % kv for kv in err.items())), flush=True)
    assert steps >= 1
    for k, v in err.items():
        assert v <= (1e-9 if k == "lambda" else TOL), (k, v)


def test_mixed_load_cases_and_recovery_on_resident_factor():
    """Load vectors and internal displacements through the refined solves of
    the mixed factor, against the oracle's direct solve."""
    import scipy.sparse.linalg as spla
    rng = np.random.default_rng(8)
    part = parts.config2(seed=6, n=40, n_ext=12)
    R = gpu.GpuReducer()
    R.set_part(part)
    R.set_option("mixedPrecision", 1)
    z = R.analyze(ngen=0)
    R.reduce()
    s = orc.sam_numbering(part)
    n1 = z["ndof1"]
    Fi = rng.normal(0.0, 1e3, (n1, 3))
    X = R.solve_internal(Fi)
    o = orc.reduce(part, ngen=0)
    i1 = s["meqn1"] - 1
    ref = spla.spsolve(o["K"][i1][:, i1].tocsc(), Fi)
    print("  K_ii solve error vs FP64 direct solve: %.1e" % rel(X, ref), flush=True)
    assert rel(X, ref) <= TOL


def test_mixed_refinement_deterministic():
    part = parts.config2(seed=3, n=40, n_ext=8)
    R = gpu.GpuReducer()
    R.set_part(part)
    R.set_option("mixedPrecision", 1)
    R.analyze(ngen=0)
    R.reduce()
    a = R.reduced()
    R.reduce()
    b = R.reduced()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
