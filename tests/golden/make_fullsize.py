#!/usr/bin/env python3
"""Writes the full-size fixtures of the 48x24x24 HEX8 block (seed 0, 12
external nodes, 20 component modes) and of the 200x200 ANDES4 plate
(seed 0, 64 edge nodes, 384 boundary DOFs): the oracle's reduced SK, SM, G,
the eigenvalues, and B^T r for a fixed seeded r (B itself is 53 MB and
743 MB).  B is refined twice with extended-precision residuals
(oracle.reducer.refine_extended): the thin plate's K_ii is ill-conditioned
enough (about 1e10) that two FP64 factorisations in different orders
disagree by 3e-10 in the reduced mass, the refined B by less than 1e-14.  The oracle needs minutes at these sizes, so its outputs are
committed as data (fullsize_<name>.npz) and the GPU tests compare with
them.  Run from the repository root: python tests/golden/make_fullsize.py"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "fedem-solvers_amd"))
from fedempy_amd import parts  # noqa: E402
from oracle import reducer as orc  # noqa: E402


def probe(n1, seed=7):
    """the fixed vector r of B^T r (ndof1 entries, MEQN1 order)"""
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n1)


def write(name, part, ngen):
    t = time.time()
    o = orc.reduce(part, ngen=ngen, dense_limit=0, refine=2)
    r = probe(o["B"].shape[0])
    out = dict(SK=o["SK"], SM=o["SM"], G=o["G"], BTr=o["B"].T @ r)
    if ngen:
        out["lam"] = o["lam"]
    np.savez_compressed(os.path.join(HERE, "fullsize_%s.npz" % name), **out)
    print(name, "%.0f s" % (time.time() - t), {k: v.shape for k, v in out.items()}, flush=True)


def main():
    which = sys.argv[1:] or ["hex8_48x24x24", "plate200"]
    if "hex8_48x24x24" in which:
        write("hex8_48x24x24", parts.config1(seed=0), 20)
    if "plate200" in which:
        write("plate200", parts.config2(seed=0), 0)


if __name__ == "__main__":
    main()
