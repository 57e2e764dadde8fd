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
