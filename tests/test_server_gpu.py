"""The continuous-batching server on the GPU: every request's results are
bitwise the same when it runs alone and when it runs next to others, and
they match the oracle (reduced matrices, reduced load vectors, recovered
internal displacements) within 1e-10 relative Frobenius error."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from fedempy_amd import parts
from fedempy_amd.server import Server
from oracle import reducer as orc

pytestmark = pytest.mark.gpu


def requests(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        part, ngen = parts.random_part(rng)
        nlc = int(rng.integers(1, 4))
        loads = parts.random_loads(rng, part, nlc)
        n2 = 6 * int(np.sum(part["status"] == 1))
        out.append((part, ngen, loads, rng.normal(0.0, 1e-3, (n2, nlc))))
    return out


def test_server_results_same_alone_and_in_mix():
    reqs = requests(7, 6)
    print("server: alone", flush=True)
    s1 = Server(slots=1)
    alone = []
    for part, ngen, loads, ub in reqs:
        rid = s1.submit(part, ngen, loads, ub)
        alone.append(s1.result(rid))
    s1.close()
    print("server: mix of %d on 4 slots" % len(reqs), flush=True)
    s4 = Server(slots=4)
    ids = [s4.submit(part, ngen, loads, ub) for part, ngen, loads, ub in reqs]
    mixed = [s4.result(i) for i in ids]
    s4.close()
    for a, m in zip(alone, mixed):
        for k in ("SK", "SM", "G", "Fred", "Ui"):
            assert np.array_equal(a[k], m[k]), k


def test_server_matches_oracle():
    rng = np.random.default_rng(3)
    part = parts.hex_block(6, 5, 4, 0.06, 0.05, 0.04, seed=5, n_ext=4)
    loads = parts.random_loads(rng, part, 2)
    n2 = 24
    ub = rng.normal(0.0, 1e-3, (n2, 2))
    srv = Server(slots=2)
    r = srv.result(srv.submit(part, 3, loads, ub))
    srv.close()
    s = orc.sam_numbering(part)
    F = np.zeros((s["neq"], 2))
    for c, case in enumerate(loads):
        for node, dof, val in case:
            eq = s["meqn"][s["madof"][node - 1] - 1 + dof - 1]
            if eq > 0:
                F[eq - 1, c] += val
    o = orc.reduce(part, ngen=3, loads=F)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(r["SK"], o["SK"]) <= 1e-10 and rel(r["SM"], o["SM"]) <= 1e-10
    sg = np.sign(np.sum(r["Fred"][n2:] * o["L"][n2:], axis=1, keepdims=True))  # mode signs may differ
    L = o["L"].copy()
    L[n2:] *= np.where(sg == 0, 1.0, sg)
    assert rel(r["Fred"], L) <= 1e-10
    pos = s["dofPosIn2"] - 1
    ubm = np.zeros_like(ub)
    ubm[pos] = ub
    i1 = s["meqn1"] - 1
    Kii = o["K"][i1][:, i1].tocsc()
    Ui = o["B"] @ ubm + spla.spsolve(Kii, F[i1])
    assert rel(r["Ui"], Ui) <= 1e-10


def test_server_heavy_mix_same_as_alone():
    """Sixteen slots busy at once (workgroups of one launch start late under
    the load of the other streams): every request still gives the bits of
    its single-slot run."""
    rng = np.random.default_rng(0)
    reqs = [parts.served_request(rng, median_dof=20000, cap_dof=60000) for _ in range(24)]
    print("server: %d requests alone" % len(reqs), flush=True)
    s1 = Server(slots=1)
    alone = [s1.result(s1.submit(*r)) for r in reqs]
    s1.close()
    print("server: the same on 16 slots", flush=True)
    s16 = Server(slots=16)
    ids = [s16.submit(*r) for r in reqs]
    mixed = [s16.result(i) for i in ids]
    s16.close()
    for a, m in zip(alone, mixed):
        for k in ("SK", "SM", "G", "Fred", "Ui"):
            assert np.array_equal(a[k], m[k]), k
