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


def test_server_mixed_precision_requests():
    """requests served with mixedPrecision set on the server: the Guyan
    (boundary) blocks of SK, SM and G, the boundary rows of the reduced loads
    and the recovered internal displacements within the mixed path's 1e-8
    of the same requests in FP64 (the mode blocks depend on the modes' basis,
    which test_mixed_gpu.py checks against the oracle)"""
    reqs = requests(11, 6)
    out = {}
    for mode in ("fp64", "mixed"):
        print("server: %s" % mode, flush=True)
        srv = Server(slots=4)
        if mode == "mixed":
            srv.set_option("mixedPrecision", 1)
        ids = [srv.submit(part, ngen, loads, ub) for part, ngen, loads, ub in reqs]
        out[mode] = [srv.result(i) for i in ids]
        srv.close()
    rel = lambda a, b: np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
    for a, m in zip(out["fp64"], out["mixed"]):
        n2 = a["ndof2"]
        for k in ("SK", "SM"):
            assert rel(m[k][:n2, :n2], a[k][:n2, :n2]) <= 1e-8, k
        assert rel(m["G"][:n2], a["G"][:n2]) <= 1e-8
        assert rel(m["Fred"][:n2], a["Fred"][:n2]) <= 1e-8
        assert rel(m["Ui"], a["Ui"]) <= 1e-8
    with pytest.raises(ValueError):
        Server(slots=1).set_option("noSuchOption", 1)


def test_server_redoes_unconverged_mixed_requests_in_fp64():
    """a mixed-precision request whose refinement does not converge (forced
    here by a step limit of 1) is redone in FP64 on its slot: flagged as
    FP64 and bitwise the same as the request submitted in FP64"""
    reqs = requests(13, 3)
    out = {}
    for mode in ("fp64", "capped"):
        print("server: %s" % mode, flush=True)
        srv = Server(slots=2)
        if mode == "capped":
            srv.set_option("mixedPrecision", 1)
            srv.set_option("refineMaxSteps", 1)
        ids = [srv.submit(part, ngen, loads, ub) for part, ngen, loads, ub in reqs]
        out[mode] = [srv.result(i) for i in ids]
        srv.close()
    for a, m in zip(out["fp64"], out["capped"]):
        assert not a["mixed"] and not m["mixed"]
        for k in ("SK", "SM", "G", "Fred", "Ui"):
            assert np.array_equal(a[k], m[k]), k


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


def _alone_and_mixed(reqs, slots, budget_gb=200.0):
    s1 = Server(slots=1, budget_gb=budget_gb)
    alone = [s1.result(s1.submit(*r)) for r in reqs]
    s1.close()
    sn = Server(slots=slots, budget_gb=budget_gb)
    ids = [sn.submit(*r) for r in reqs]
    mixed = [sn.result(i) for i in ids]
    held = sn.device_bytes()
    sn.close()
    for a, m in zip(alone, mixed):
        for k in ("SK", "SM", "G", "Fred", "Ui"):
            assert np.array_equal(a[k], m[k]), k
    return alone, held


def test_served_mix_at_the_300k_cap():
    """The served distribution as configured (DOF lognormal, median 20k,
    capped at 300k; 1-8 load cases with recovery): 24 requests, among them
    parts above 100k DOF, bitwise the same alone and on 16 slots."""
    rng = np.random.default_rng(11)
    reqs = [parts.served_request(rng) for _ in range(24)]
    dofs = [6 * len(r[0]["xyz"]) for r in reqs]
    assert max(dofs) > 100000, max(dofs)
    print("served mix at the 300k cap: largest %d DOF" % max(dofs), flush=True)
    _alone_and_mixed(reqs, 16)


def test_batched_256_parts_same_alone_and_oracle_sample():
    """The 256-part batched mix (plates 10-80 squared, HEX8 blocks 6-24
    cubed, beam frames 50-500 elements, 2-16 external nodes, 0-10 modes):
    bitwise the same alone and on 16 slots, and the eight smallest parts
    against the oracle at 1e-10."""
    rng = np.random.default_rng(0)
    reqs = []
    for _ in range(256):
        part, ngen = parts.random_part(rng)
        reqs.append((part, ngen, None, None))
    print("batched mix of 256 parts", flush=True)
    alone, _ = _alone_and_mixed(reqs, 16)
    order = np.argsort([len(r[0]["xyz"]) for r in reqs])[:8]
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    for k in order:
        part, ngen, _, _ = reqs[k]
        o = orc.reduce(part, ngen=ngen)
        n2 = alone[k]["ndof2"]
        assert rel(alone[k]["SK"][:n2, :n2], o["SK"][:n2, :n2]) <= 1e-10
        assert rel(alone[k]["G"][:n2], o["G"][:n2]) <= 1e-10
        if ngen == 0:
            assert rel(alone[k]["SM"], o["SM"]) <= 1e-10


def test_tight_budget_releases_memory():
    """A 1 GB budget with one large request among small ones: the large one
    runs alone, idle slots give their buffers back when a request needs the
    room, results stay bitwise the same, and the slots end within the
    budget."""
    rng = np.random.default_rng(5)
    big = parts.hex_block(30, 16, 16, 0.3, 0.16, 0.16, seed=2, n_ext=8)
    reqs = [(big, 4, parts.random_loads(rng, big, 2), None)]
    for _ in range(8):
        part, ngen = parts.random_part(rng, target_dof=4000)
        reqs.append((part, ngen, parts.random_loads(rng, part, 1), None))
    print("tight budget: one %d-DOF request among 8 small ones" % (3 * len(big["xyz"])), flush=True)
    _, held = _alone_and_mixed(reqs, 4, budget_gb=1.0)
    assert held <= 1.0 * 2**30, held
