"""The C-ABI library without a GPU: it loads, exports every function that
include/fedem_reducer_gpu.h declares, parses options and part files the way
the drop-in entry points need, and writes .fmx files.  No compute calls."""
import ctypes
import os
import re
import struct

import numpy as np
import pytest

from fedempy_amd import LIB_PATH, fmx, parts
from fedempy_amd.reducer import FedemReducer
from oracle import nastran
from oracle import reducer as orc

HEADER = os.path.join(os.path.dirname(__file__), "..", "include", "fedem_reducer_gpu.h")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB_PATH):
        pytest.fail("libfedem_reducer_core.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    L.fgpu_opt_double.restype = ctypes.c_double
    return L


def declared_functions():
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt, re.M)
    return sorted(set(names))


def test_exports_every_declared_function(lib):
    names = declared_functions()
    for must in ("initSolverArgs", "reducePart", "solvePart", "solvePartDis", "freeSingeltons", "fgpu_reduce"):
        assert must in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_only_declared_symbols_are_exported():
    out = os.popen("nm -D --defined-only %s" % LIB_PATH).read().split("\n")
    exported = {l.split()[-1] for l in out if " T " in l}
    assert exported <= set(declared_functions()), exported - set(declared_functions())


def _opt_str(lib, name):
    buf = ctypes.create_string_buffer(512)
    n = lib.fgpu_opt_string(name.encode(), buf, 512)
    return buf.raw[:n].decode()


def test_options_like_fedempy(lib, tmp_path):
    fco = tmp_path / "extra.fco"
    fco.write_text("# comment\n-neval 7\n-tolEigval=1.0e-9\n")
    red = FedemReducer(solver_options=["-linkfile=part.nas", "-ngen=5", "-gvec", "0", "-9.81", "0",
                                       "-fco=%s" % fco, "-lumpedmass"])
    assert lib.fgpu_opt_int(b"ngen") == 5
    assert lib.fgpu_opt_int(b"neval") == 7
    assert lib.fgpu_opt_double(b"tolEigval") == 1.0e-9
    assert lib.fgpu_opt_double(b"tolFactorize") == 1.0e-12  # default
    assert lib.fgpu_opt_bool(b"lumpedmass") == 1
    assert lib.fgpu_opt_isset(b"dispfile") == 0
    assert _opt_str(lib, "linkfile") == "part.nas"
    v = (ctypes.c_double * 3)()
    assert lib.fgpu_opt_doubles(b"gvec", v, 3) == 3 and list(v) == [0.0, -9.81, 0.0]
    # a later part keeps the definitions and takes new values only
    assert red.solver_init(["-ngen=2"]) == 0 or True
    red.release()


def test_initsolverargs_second_call_replaces_values(lib):
    argv = (ctypes.c_char_p * 2)(b"fedem_reducer", b"-ngen=3")
    assert lib.initSolverArgs(2, argv, ctypes.c_bool(True)) == 0
    assert lib.fgpu_opt_int(b"ngen") == 3
    argv = (ctypes.c_char_p * 2)(b"fedem_reducer", b"-neval=4")
    assert lib.initSolverArgs(2, argv, ctypes.c_bool(False)) == 0
    assert lib.fgpu_opt_int(b"neval") == 4 and lib.fgpu_opt_int(b"ngen") == 0
    lib.freeSingeltons()


def test_missing_option_file_is_an_error(lib):
    argv = (ctypes.c_char_p * 2)(b"fedem_reducer", b"-fco=/nonexistent/x.fco")
    assert lib.initSolverArgs(2, argv, ctypes.c_bool(True)) < 0
    lib.freeSingeltons()


def _link_sizes(lib):
    s = (ctypes.c_int * 4)()
    assert lib.fgpu_link_sizes(s) == 0
    return list(s)


def test_nastran_reader_on_reference_parts(lib, golden_dir):
    for name in ("CHEXA08_", "CQUAD04_", "CBAR02_"):
        path = os.path.join(golden_dir, name + ".nas")
        assert lib.fgpu_link_read(path.encode()) == 0
        o = nastran.read(path)
        nnod, nel, nrgd, nlc = _link_sizes(lib)
        assert nnod == len(o["xyz"]) and nel == len(o["melcon"]) and nrgd == len(o.get("rbe2", []))
        ids = (ctypes.c_int * nnod)()
        lib.fgpu_link_node_ids(ids)
        assert list(ids) == list(o["node_ids"])


def test_written_part_roundtrip_and_sam(lib, tmp_path):
    """A part written as bulk data, read by the library, numbered by the
    library's SAM (host side of fgpu_analyze): bit-exact with the oracle."""
    p = parts.hex_block(4, 3, 2, 0.04, 0.03, 0.02, seed=3, n_ext=3)
    st = p["status"].copy()
    st[np.nonzero(p["xyz"][:, 0] == 0.0)[0][:2]] = -5
    p["status"] = st
    f = tmp_path / "blk.nas"
    parts.write_nastran(p, str(f), loads=[(7, 5, (0.0, 10.0, 0.0)), (3, 9, (1.0, 0.0, 2.0))])
    back = nastran.read(str(f))
    assert np.array_equal(back["xyz"], p["xyz"]) and np.array_equal(back["eprop"][:, :3], p["eprop"][:, :3])
    assert lib.fgpu_link_read(str(f).encode()) == 0
    nnod, nel, nrgd, nlc = _link_sizes(lib)
    assert (nnod, nel, nrgd, nlc) == (len(p["xyz"]), len(p["melcon"]), 0, 2)
    n = lib.fgpu_link_loads(None, None, None, None)
    lc, nd, dof = (ctypes.c_int * n)(), (ctypes.c_int * n)(), (ctypes.c_int * n)()
    val = (ctypes.c_double * n)()
    assert lib.fgpu_link_loads(lc, nd, dof, val) == n == 3
    got = sorted(zip(lc, nd, dof, val))
    assert got == [(1, 9, 1, 1.0), (1, 9, 3, 2.0), (2, 5, 2, 10.0)]  # set 3 -> case 1, set 7 -> case 2
    h = ctypes.c_void_p()
    assert lib.fgpu_create(ctypes.byref(h)) == 0
    try:
        assert lib.fgpu_link_attach(h) == 0
        assert lib.fgpu_analyze(h, 2, 0, 0) == 0
        sz = (ctypes.c_int * 8)()
        lib.fgpu_get_sizes(h, sz)
        s = orc.sam_numbering(p)
        nnod, ndof, neq, n1, n2 = sz[0], sz[1], sz[2], sz[3], sz[4]
        arrs = [(ctypes.c_int * k)() for k in (nnod + 1, ndof, ndof, n1, n2, n2)]
        assert lib.fgpu_get_sam(h, *arrs) == 0
        for a, key in zip(arrs, ("madof", "msc", "meqn", "meqn1", "meqn2", "dofPosIn2")):
            assert np.array_equal(np.array(a[:]), np.asarray(s[key])), key
    finally:
        lib.fgpu_destroy(h)


def _thru_part_text():
    """A 3x1x1 HEX8 strip whose x=0 face is clamped by "SPC1 ... THRU" (grids
    1-4, with the non-existent grid 3 inside the range skipped) and whose
    external grids 13, 15, 16 come from "ASET1 ... THRU"."""
    lines = ["BEGIN BULK", "MAT1,1,2.1e11,,0.3,7850.", "PSOLID,1,1"]
    ids = {}
    k = 0
    for i in range(4):
        for j in range(2):
            for m in range(2):
                k += 1
                gid = k if k != 3 else 40  # grid 3 does not exist
                ids[(i, j, m)] = gid
                lines.append("GRID,%d,,%g,%g,%g" % (gid, 0.01 * i, 0.01 * j, 0.01 * m))
    for e in range(3):
        n = [ids[(e, 0, 0)], ids[(e + 1, 0, 0)], ids[(e + 1, 1, 0)], ids[(e, 1, 0)],
             ids[(e, 0, 1)], ids[(e + 1, 0, 1)], ids[(e + 1, 1, 1)], ids[(e, 1, 1)]]
        lines.append("CHEXA,%d,1,%s" % (e + 1, ",".join(map(str, n))))
    lines += ["SPC1,1,123,1,THRU,4", "SPC1,1,123,40", "ASET1,123456,13,THRU,16", "ENDDATA"]
    return "\n".join(lines) + "\n"


def test_nastran_thru_ranges(lib, tmp_path):
    """SPC1 / ASET1 "G1 THRU G2" ranges in the library's reader and in the
    oracle's: the same node status, hence bit-exact SAM numbering."""
    f = tmp_path / "thru.nas"
    f.write_text(_thru_part_text())
    o = nastran.read(str(f))
    ids = list(o["node_ids"])
    st = dict(zip(ids, o["status"]))
    assert all(st[g] == -7 for g in (1, 2, 4, 40)) and st[14] == 1 == st[13] == st[15] == st[16]
    assert sum(1 for v in st.values() if v == 0) == len(ids) - 8
    assert lib.fgpu_link_read(str(f).encode()) == 0
    h = ctypes.c_void_p()
    assert lib.fgpu_create(ctypes.byref(h)) == 0
    try:
        assert lib.fgpu_link_attach(h) == 0
        assert lib.fgpu_analyze(h, 0, 0, 0) == 0
        sz = (ctypes.c_int * 8)()
        lib.fgpu_get_sizes(h, sz)
        s = orc.sam_numbering(o)
        arrs = [(ctypes.c_int * k)() for k in (sz[0] + 1, sz[1], sz[1])]
        assert lib.fgpu_get_sam(h, arrs[0], arrs[1], arrs[2], None, None, None) == 0
        for a, key in zip(arrs, ("madof", "msc", "meqn")):
            assert np.array_equal(np.array(a[:]), np.asarray(s[key])), key
    finally:
        lib.fgpu_destroy(h)
    bad = tmp_path / "bad_thru.nas"
    bad.write_text(_thru_part_text().replace("SPC1,1,123,1,THRU,4", "SPC1,1,123,THRU,4"))
    assert lib.fgpu_link_read(str(bad).encode()) < 0
    with pytest.raises(ValueError):
        nastran.read(str(bad))
    bad.write_text(_thru_part_text().replace("SPC1,1,123,40", "SPC1,1,123,4O"))
    assert lib.fgpu_link_read(str(bad).encode()) < 0
    buf = ctypes.create_string_buffer(256)
    lib.fgpu_link_error(buf, 256)
    assert b"not a grid id" in buf.value


def test_ngen_larger_than_ndof1_is_rejected(lib):
    """inputReducerModule.f90:403-409: NGEN > NDOF1 is an error (host side
    of fgpu_analyze, no device work)."""
    p = parts.hex_block(2, 1, 1, 0.02, 0.01, 0.01, seed=1, n_ext=4)
    s = orc.sam_numbering(p)
    h = ctypes.c_void_p()
    assert lib.fgpu_create(ctypes.byref(h)) == 0
    try:
        arrs = [np.ascontiguousarray(p[k], dt) for k, dt in
                (("xyz", np.float64), ("status", np.int32), ("melcon", np.int32), ("mpmnpc", np.int32),
                 ("mmnpc", np.int32), ("eprop", np.float64))]
        ptr = [a.ctypes.data_as(ctypes.c_void_p) for a in arrs]
        assert lib.fgpu_set_part(h, len(p["xyz"]), ptr[0], ptr[1], len(p["melcon"]), ptr[2], ptr[3], ptr[4],
                                 ptr[5]) == 0
        assert lib.fgpu_analyze(h, s["ndof1"] + 1, 0, 0) == -1
        lib.fgpu_last_error.restype = ctypes.c_char_p
        assert b"NGEN" in lib.fgpu_last_error(h) and b"NDOF1" in lib.fgpu_last_error(h)
        assert lib.fgpu_analyze(h, s["ndof1"], s["ndof1"] + 5, 0) == 0
    finally:
        lib.fgpu_destroy(h)


def test_nastran_reader_rejects_bad_input(lib, tmp_path):
    f = tmp_path / "bad.nas"
    f.write_text("BEGIN BULK\nGRID,1,,0.,0.,0.\nCHEXA,1,1,1,2,3,4,5,6,7,8\nENDDATA\n")
    assert lib.fgpu_link_read(str(f).encode()) < 0
    buf = ctypes.create_string_buffer(256)
    lib.fgpu_link_error(buf, 256)
    assert b"missing property" in buf.raw


MOTIONS = os.path.join(os.path.dirname(__file__), "golden", "motions.fnd")


def test_file_tag_matches_reference_data_file(lib, tmp_path):
    """The reference's motions.fnd (createTestFnd.C:38-40: FFaTag::write of
    "#FEDEM nodal displacements" with checksum 12345678, then ";7.3;") pins
    the tag bytes: the tag writer reproduces its first 45 bytes exactly."""
    ref = open(MOTIONS, "rb").read()
    f = tmp_path / "motions.fnd"
    assert lib.fgpu_write_tag(str(f).encode(), b"#FEDEM nodal displacements", ctypes.c_uint(12345678), b";7.3;") == 0
    assert f.read_bytes() == ref[:45]
    # a planted change of the checksum or of the padding breaks the match
    assert lib.fgpu_write_tag(str(f).encode(), b"#FEDEM nodal displacements", ctypes.c_uint(12345679), b";7.3;") == 0
    assert f.read_bytes() != ref[:45]
    assert lib.fgpu_write_tag(str(f).encode(), b"#FEDEM nodal displacements ", ctypes.c_uint(12345678), b";7.3;") == 0
    assert f.read_bytes() == ref[:45]  # trailing blanks are padding
    assert lib.fgpu_write_tag(str(f).encode(), b"x" * 31, ctypes.c_uint(0), b";1.0;\n") < 0


def test_tag_reader_parses_reference_data_file(lib):
    """readtagdb semantics on motions.fnd, then the file's own records
    (createTestFnd.C:60-84): 5 node numbers, 3 steps of (step, time, 15
    displacements 1..15, raised by 100 per step)."""
    tag = ctypes.create_string_buffer(64)
    ver = ctypes.create_string_buffer(64)
    cs = ctypes.c_uint(0)
    lib.fgpu_read_tag.restype = ctypes.c_int64
    off = lib.fgpu_read_tag(MOTIONS.encode(), tag, 64, ctypes.byref(cs), ver, 64)
    assert off == 48
    assert tag.value == b"#FEDEM nodal displacements" and cs.value == 12345678 and ver.value == b";7.3;5;\n"
    raw = open(MOTIONS, "rb").read()
    t, c, v, o = fmx.read_tag(raw)
    assert (t, c, v, o) == ("#FEDEM nodal displacements", 12345678, ";7.3;5;\n", 48)
    nodes = np.frombuffer(raw, "<i4", 5, o)
    assert list(nodes) == [7, 12, 38, 42, 59]
    p = o + 20
    for step in range(3):
        assert struct.unpack_from("<i", raw, p)[0] == step
        assert abs(struct.unpack_from("<d", raw, p + 4)[0] - 0.1 * step) < 1e-15
        assert np.array_equal(np.frombuffer(raw, "<f8", 15, p + 12), np.arange(1.0, 16.0) + 100.0 * step)
        p += 12 + 120
    assert p == len(raw) == 464
    with pytest.raises(ValueError):
        fmx.read_tag(b"#FEDEM stiffness matrix;1234;1.0;\n" + bytes(16))


def test_fmx_writer_layout(lib, tmp_path):
    a = np.arange(12.0).reshape(3, 4, order="F")
    f = tmp_path / "x_S.fmx"
    assert lib.fgpu_write_fmx(str(f).encode(), b"stiffness matrix", 1234,
                              a.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(a.size)) == 0
    raw = f.read_bytes()
    want = b"#FEDEM stiffness matrix".ljust(30) + struct.pack("<IHI", 0x1234, 0, 1234) + b";1.0;\n"
    assert raw[:46] == want
    b, tag, cs = fmx.read_fmx(f, (3, 4))
    assert tag == "#FEDEM stiffness matrix" and cs == 1234
    assert np.array_equal(b, a)


def test_fmxwriter_fortran_entry_roundtrip(lib, tmp_path):
    """writefmx_/readfmx_ as fedempy.write_fmx calls them (fmxWriter.f90)."""
    prefix = str(tmp_path / "part")
    data = np.linspace(-1.0, 1.0, 10)
    back = np.zeros(10)
    for itype, suffix in ((1, "_K.fmx"), (2, "_M.fmx"), (3, "_G.fmx")):
        it, nv = ctypes.c_int(itype), ctypes.c_int(10)
        pb = prefix.encode() + b"   "  # Fortran blank padding
        assert lib.writefmx_(pb, ctypes.byref(it), data.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nv),
                             ctypes.c_size_t(len(pb))) == 0
        assert os.path.exists(prefix + suffix)
        assert lib.readfmx_(pb, ctypes.byref(it), back.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nv),
                            ctypes.c_size_t(len(pb))) == 0
        assert np.array_equal(back, data)
    it = ctypes.c_int(4)
    nv = ctypes.c_int(10)
    assert lib.writefmx_(b"x", ctypes.byref(it), data.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nv),
                         ctypes.c_size_t(1)) < 0


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-device error path")
def test_reducepart_without_device_fails_cleanly(tmp_path):
    p = parts.hex_block(3, 2, 2, 0.03, 0.02, 0.02, seed=1, n_ext=2)
    f = tmp_path / "blk.nas"
    parts.write_nastran(p, str(f))
    red = FedemReducer(solver_options=["-linkfile=%s" % f, "-ngen=2"])
    assert red.run() < 0
    res = (tmp_path / "blk.res").read_text()
    assert "FE MODEL SUMMARY" in res and "Reduction failed" in res
    red.release()


def test_oracle_load_reduction_reproduces_full_solve():
    """Pins the oracle's reduced load vectors (and stiffness): with ngen = 0
    the condensed system SK u_b = L gives the boundary displacements of the
    full solve K u = F of a clamped part."""
    p = parts.hex_block(5, 3, 3, 0.05, 0.03, 0.03, seed=5, n_ext=3)
    st = p["status"].copy()
    x0 = np.nonzero(p["xyz"][:, 0] == 0.0)[0]
    st[x0[st[x0] == 0]] = -7
    p["status"] = st
    s = orc.sam_numbering(p)
    rng = np.random.default_rng(0)
    F = rng.normal(size=(s["neq"], 2))
    o = orc.reduce(p, loads=F)
    K = o["K"].toarray()
    act = np.nonzero(np.diag(K) > 0)[0]  # rotations of the external nodes carry no stiffness
    F[np.diag(K) == 0] = 0.0
    o = orc.reduce(p, loads=F)
    u = np.zeros_like(F)
    u[act] = np.linalg.solve(K[np.ix_(act, act)], F[act])
    n2 = s["ndof2"]
    pos = s["dofPosIn2"] - 1
    SK = o["SK"][:n2, :n2]
    b = np.nonzero(np.diag(SK) > 0)[0]
    ub = np.linalg.solve(SK[np.ix_(b, b)], o["L"][:n2][b])
    assert np.allclose(ub, u[s["meqn2"] - 1][pos][b], rtol=1e-9, atol=1e-14)
    planted = F.copy()
    planted[s["meqn1"][0] - 1] *= 2.0  # a changed internal load must move u_b
    o2 = orc.reduce(p, loads=planted)
    ub2 = np.linalg.solve(SK[np.ix_(b, b)], o2["L"][:n2][b])
    assert not np.allclose(ub2, ub, rtol=1e-9, atol=0.0)


def test_reduced_eigen_matches_generalized_eigh(lib):
    """fgpu_reduced_eigen (the EIGCMS eigenvalue call) against scipy's
    generalized symmetric solver on seeded SPD pairs, and DSYGVX's INFO
    convention (n + order of the failing minor) for a singular mass."""
    import scipy.linalg as sla
    f = lib.fgpu_reduced_eigen
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    rng = np.random.default_rng(11)
    for n, nval in ((1, 1), (12, 12), (92, 12), (200, 30)):
        X = rng.normal(size=(n, n))
        A = X @ X.T + 1e-3 * np.eye(n)
        A[0, 0] -= 2.0 * A[0, 0]  # one negative direction: indefinite stiffness is allowed
        Y = rng.normal(size=(n, n))
        B = Y @ Y.T + n * np.eye(n)
        A, B = np.asfortranarray(A), np.asfortranarray(B)
        lam = np.zeros(nval)
        assert f(n, A.ctypes.data, B.ctypes.data, nval, lam.ctypes.data) == 0
        ref = sla.eigh(A, B, eigvals_only=True, subset_by_index=[0, nval - 1])
        assert np.allclose(lam, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max()), (n, lam - ref)
    # a zero-padded DOF (fixed DOF of an external node) makes the mass singular
    n = 6
    B = np.eye(n)
    B[3, 3] = 0.0
    lam = np.zeros(2)
    A = np.asfortranarray(np.eye(n))
    assert f(n, A.ctypes.data, np.asfortranarray(B).ctypes.data, 2, lam.ctypes.data) == n + 4
    assert f(n, A.ctypes.data, A.ctypes.data, n + 1, lam.ctypes.data) == -1


def test_written_beam_and_shell_parts_roundtrip(tmp_path):
    """The bulk-data writer keeps every property the element kernels read:
    PBAR with its continuation (shear factors K1, K2), PSHELL thickness,
    CBAR orientation vectors."""
    for p in (parts.beam_frame(n_el=60, seed=3, n_ext=3), parts.config0()):
        f = tmp_path / "part.nas"
        parts.write_nastran(p, str(f))
        back = nastran.read(str(f))
        assert np.array_equal(back["xyz"], p["xyz"])
        assert np.array_equal(back["eprop"], p["eprop"])
