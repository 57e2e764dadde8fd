"""Seeded synthetic FE parts in the reducer's input form.

A part is a dict with SAM-style arrays (samReducerModule.f90:227-569):
  xyz    (nnod, 3) nodal coordinates
  status (nnod,)   1 external, 0 internal, negative = fixed-DOF bitmask
                   -(sum of 2^i over the fixed local DOFs i), as testModels.C:512-525
  melcon (nel,)    element type (11 beam, 24 ANDES4 shell, 44 HEX8)
  mpmnpc (nel+1,)  1-based pointers into mmnpc
  mmnpc           1-based node numbers of the elements
  eprop  (nel, 16) per-element properties (E, nu, rho, G, thickness/section ...)
"""
import numpy as np

NPROP = 16
STEEL = dict(E=2.1e11, nu=0.3, rho=7850.0)
FIXED_ALL = -63  # all six DOFs fixed


def _finish(xyz, status, melcon, conn, props):
    nel = len(melcon)
    sizes = np.array([len(c) for c in conn], np.int32)
    mpmnpc = np.concatenate([[1], 1 + np.cumsum(sizes)]).astype(np.int32)
    mmnpc = (np.concatenate(conn) + 1).astype(np.int32) if nel else np.zeros(0, np.int32)
    return dict(xyz=np.ascontiguousarray(xyz, np.float64), status=np.ascontiguousarray(status, np.int32),
                melcon=np.ascontiguousarray(melcon, np.int32), mpmnpc=mpmnpc, mmnpc=mmnpc,
                eprop=np.ascontiguousarray(props, np.float64))


def hex_block(nx, ny, nz, lx, ly, lz, E=2.1e11, nu=0.3, rho=7850.0, jitter=0.05, seed=0,
              n_ext=12, ext_nodes=None):
    """nx x ny x nz HEX8 block on [0,lx]x[0,ly]x[0,lz].

    Nodal coordinates are jittered by U(-jitter, jitter) times the element size
    per axis, except the component normal to a boundary face for nodes on that
    face (faces stay planar).  n_ext external nodes are drawn without
    replacement from the x=0 and x=lx faces with the same seed.
    """
    rng = np.random.default_rng(seed)
    hx, hy, hz = lx / nx, ly / ny, lz / nz
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).ravel() for a in (i, j, k))  # x fastest
    xyz = np.stack([i * hx, j * hy, k * hz], axis=1).astype(np.float64)
    if jitter > 0:
        d = rng.uniform(-jitter, jitter, size=xyz.shape) * np.array([hx, hy, hz])
        d[(i == 0) | (i == nx), 0] = 0.0
        d[(j == 0) | (j == ny), 1] = 0.0
        d[(k == 0) | (k == nz), 2] = 0.0
        xyz += d
    nid = lambda a, b, c: a + (nx + 1) * (b + (ny + 1) * c)
    ei, ej, ek = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ei, ej, ek = (a.transpose(2, 1, 0).ravel() for a in (ei, ej, ek))
    conn = np.stack([nid(ei, ej, ek), nid(ei + 1, ej, ek), nid(ei + 1, ej + 1, ek), nid(ei, ej + 1, ek),
                     nid(ei, ej, ek + 1), nid(ei + 1, ej, ek + 1), nid(ei + 1, ej + 1, ek + 1),
                     nid(ei, ej + 1, ek + 1)], axis=1)
    status = np.zeros(len(xyz), np.int32)
    if ext_nodes is None:
        cand = np.nonzero((i == 0) | (i == nx))[0]
        ext_nodes = rng.choice(cand, size=min(n_ext, len(cand)), replace=False)
    status[np.asarray(ext_nodes, int)] = 1
    nel = len(conn)
    props = np.zeros((nel, NPROP))
    props[:, 0], props[:, 1], props[:, 2] = E, nu, rho
    return _finish(xyz, status, np.full(nel, 44, np.int32), list(conn), props)


def config1(seed=0):
    """48x24x24 HEX8 steel block, 0.48x0.24x0.24 m, +-5 % jitter, 12 external
    nodes on the x=0 and x=max faces (72 boundary DOFs), 20 component modes."""
    return hex_block(48, 24, 24, 0.48, 0.24, 0.24, jitter=0.05, seed=seed, n_ext=12)


def write_nastran(part, path, loads=None):
    """Writes a part as Nastran bulk data in free-field format with every
    real at full double precision (GRID, CHEXA/CQUAD4/CBAR, MAT1,
    PSOLID/PSHELL/PBAR, ASET1 for the external nodes, SPC1 for fixed DOFs,
    FORCE for nodal loads), so that the part read back is bit-identical.
    Node ids are 1-based node numbers; one material and one property card
    per element.
    loads: iterable of (load set id, node number 1-based, (fx, fy, fz))."""
    xyz, status = part["xyz"], part["status"]
    mel, mp, mm, pr = part["melcon"], part["mpmnpc"], part["mmnpc"], part["eprop"]
    out = ["BEGIN BULK"]
    r = lambda v: repr(float(v))
    for n, (x, y, z) in enumerate(xyz):
        out.append("GRID,%d,,%s,%s,%s" % (n + 1, r(x), r(y), r(z)))
    name = {44: "CHEXA", 24: "CQUAD4", 11: "CBAR"}
    for e in range(len(mel)):
        nodes = [str(int(v)) for v in mm[mp[e] - 1:mp[e + 1] - 1]]
        t = int(mel[e])
        fields = [str(e + 1), str(e + 1)] + nodes
        if t == 11:
            fields += [r(v) for v in pr[e, 8:11]]
        out.append(name[t] + "," + ",".join(fields))
        E, nu, rho = pr[e, 0], pr[e, 1], pr[e, 2]
        G = pr[e, 3] if pr[e, 3] else E / (2 * (1 + nu))
        out.append("MAT1,%d,%s,%s,%s,%s" % (e + 1, r(E), r(G), r(nu), r(rho)))
        if t == 44:
            out.append("PSOLID,%d,%d" % (e + 1, e + 1))
        elif t == 24:
            out.append("PSHELL,%d,%d,%s" % (e + 1, e + 1, r(pr[e, 4])))
        else:
            out.append("PBAR,%d,%d,%s" % (e + 1, e + 1, ",".join(r(v) for v in (*pr[e, 4:8], pr[e, 13]))))
    ext = [i + 1 for i in range(len(status)) if status[i] == 1]
    for k in range(0, len(ext), 6):
        out.append("ASET1,123456," + ",".join(str(v) for v in ext[k:k + 6]))
    for i in range(len(status)):
        if status[i] < 0:
            comp = "".join(str(j + 1) for j in range(6) if (-int(status[i]) >> j) & 1)
            out.append("SPC1,1,%s,%d" % (comp, i + 1))
    for sid, node, f in loads or []:
        out.append("FORCE,%d,%d,0,1.0,%s,%s,%s" % (sid, node, *(r(v) for v in f)))
    out.append("ENDDATA")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")
