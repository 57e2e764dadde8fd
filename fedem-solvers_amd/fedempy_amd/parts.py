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


def plate(nx, ny, lx, ly, t=0.005, E=2.1e11, nu=0.3, rho=7850.0, thickness=None, ext_nodes=(), clamp_x0=False,
          warp=0.0, seed=0):
    """nx x ny ANDES4 shell plate on [0,lx]x[0,ly] in the xy plane.

    Node (i, j) (i along x, j along y) has index i*(ny+1) + j, so its id is
    (ny+1)*i + j + 1 (the testModels.C:397-405 numbering).  thickness: per
    element thickness array (nx*ny, element (i, j) at i*ny + j), else t.
    ext_nodes: node indices made external; clamp_x0 fixes all six DOFs of the
    x=0 edge.  warp: out-of-plane jitter amplitude (times the element size)
    drawn with the seed, for non-planar test elements."""
    i, j = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
    i, j = i.ravel(), j.ravel()
    hx, hy = lx / nx, ly / ny
    xyz = np.stack([i * hx, j * hy, np.zeros_like(i, float)], axis=1).astype(np.float64)
    if warp > 0:
        rng = np.random.default_rng(seed)
        xyz[:, 2] = rng.uniform(-warp, warp, len(xyz)) * min(hx, hy)
    nid = lambda a, b: a * (ny + 1) + b
    ei, ej = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ei, ej = ei.ravel(), ej.ravel()
    conn = np.stack([nid(ei, ej), nid(ei + 1, ej), nid(ei + 1, ej + 1), nid(ei, ej + 1)], axis=1)
    status = np.zeros(len(xyz), np.int32)
    if clamp_x0:
        status[i == 0] = FIXED_ALL
    status[np.asarray(ext_nodes, int)] = 1
    nel = len(conn)
    th = np.full(nel, t) if thickness is None else np.asarray(thickness, float)
    props = np.zeros((nel, NPROP))
    props[:, 0], props[:, 1], props[:, 2] = E, nu, rho
    props[:, 3] = E / (2 * (1 + nu))
    props[:, 4:8] = th[:, None]
    return _finish(xyz, status, np.full(nel, 24, np.int32), list(conn), props)


def config0():
    """Cantilever plate 1.0 x 0.5 m, 20 x 10 ANDES4, t = 5 mm, steel, x = 0
    edge clamped, external nodes at the free corners (1.0, 0) and (1.0, 0.5)
    (12 boundary DOFs), no component modes."""
    nx, ny = 20, 10
    return plate(nx, ny, 1.0, 0.5, 0.005, ext_nodes=[nx * (ny + 1), nx * (ny + 1) + ny], clamp_x0=True)


def config2(seed=0, n=200, n_ext=64):
    """n x n ANDES4 plate, 1 x 1 m, per-element thickness lognormal(ln 5 mm,
    0.3) from the seed, n_ext external nodes drawn from the edge nodes with
    the same seed (6 DOFs each), no fixed DOFs."""
    rng = np.random.default_rng(seed)
    th = np.exp(rng.normal(np.log(0.005), 0.3, n * n))
    i, j = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    i, j = i.ravel(), j.ravel()
    edge = np.nonzero((i == 0) | (i == n) | (j == 0) | (j == n))[0]
    ext = rng.choice(edge, size=n_ext, replace=False)
    return plate(n, n, 1.0, 1.0, thickness=th, ext_nodes=ext)


TUBE = dict(ro=0.020, t=0.002)


def tube_section(ro=0.020, t=0.002):
    """Area, I1 = I2 and J = 2 I of a thin-walled tube."""
    ri = ro - t
    A = np.pi * (ro * ro - ri * ri)
    I = np.pi * (ro ** 4 - ri ** 4) / 4.0
    return A, I, 2.0 * I


def beam_frame(n_el=100, seed=0, n_ext=4, bay=0.5, E=2.1e11, nu=0.3, rho=7850.0, shear_k=0.5):
    """Space-frame tower of 2-node tube beams (r_o = 20 mm, t = 2 mm): square
    bays of side `bay` stacked along z, each with four columns, four
    horizontal members and two face diagonals, every member split into `seg`
    elements; bays and segments are chosen so the element count is close to
    n_el.  n_ext external nodes are drawn with the seed among the member
    joints, always including one base corner so the internal problem is
    constrained."""
    rng = np.random.default_rng(seed)
    per_bay = 10
    seg = max(1, int(round(np.sqrt(n_el / (per_bay * 4.0)))))
    nbay = max(1, int(round(n_el / (per_bay * seg))))
    corners = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float) * bay
    joints = []
    for k in range(nbay + 1):
        for c in corners:
            joints.append([c[0], c[1], k * bay])
    xyz = [list(j) for j in joints]
    conn, vecs = [], []

    def member(a, b):
        pa, pb = np.array(xyz[a]), np.array(xyz[b])
        d = pb - pa
        v = [1.0, 0.0, 0.0] if abs(d[2]) > 0.9 * np.linalg.norm(d) else [0.0, 0.0, 1.0]
        prev = a
        for s_ in range(1, seg + 1):
            if s_ < seg:
                xyz.append(list(pa + d * s_ / seg))
                cur = len(xyz) - 1
            else:
                cur = b
            conn.append([prev, cur])
            vecs.append(v)
            prev = cur

    J = lambda k, c: 4 * k + c
    for k in range(nbay):
        for c in range(4):
            member(J(k, c), J(k + 1, c))                # columns
            member(J(k + 1, c), J(k + 1, (c + 1) % 4))  # ring at the top of the bay
        member(J(k, 0), J(k + 1, 1))                    # face diagonals
        member(J(k, 2), J(k + 1, 3))
    for c in range(4):
        member(J(0, c), J(0, (c + 1) % 4))              # base ring
    xyz = np.array(xyz, float)
    status = np.zeros(len(xyz), np.int32)
    cand = np.arange(1, 4 * (nbay + 1))
    ext = [0] + list(rng.choice(cand, size=min(max(n_ext - 1, 0), len(cand)), replace=False))
    status[ext] = 1
    nel = len(conn)
    A, I, Jt = tube_section()
    props = np.zeros((nel, NPROP))
    props[:, 0], props[:, 1], props[:, 2], props[:, 3] = E, nu, rho, E / (2 * (1 + nu))
    props[:, 4], props[:, 5], props[:, 6], props[:, 7] = A, I, I, Jt
    props[:, 8:11] = np.array(vecs)
    props[:, 11] = props[:, 12] = shear_k
    return _finish(xyz, status, np.full(nel, 11, np.int32), [np.array(c) for c in conn], props)


def _stiff_beam_props(n, E=2.1e11, nu=0.3, rho=7850.0, ro=0.05, t=0.01, v=(1.0, 0.0, 0.0)):
    A, I, Jt = tube_section(ro, t)
    props = np.zeros((n, NPROP))
    props[:, 0], props[:, 1], props[:, 2], props[:, 3] = E, nu, rho, E / (2 * (1 + nu))
    props[:, 4], props[:, 5], props[:, 6], props[:, 7] = A, I, I, Jt
    props[:, 8:11] = np.array(v)
    props[:, 11] = props[:, 12] = 0.5
    return props


def _wavgm(ref, nodes, weights, rng):
    """a WAVGM element dict: weights "default" (no properties: uniform
    translations), "equal" (explicit equal translational weights), "random"
    (random translational weights) or "rotational" (random weights on all six
    components, translations and rotations in separate sets)"""
    el = dict(ref=int(ref), nodes=[int(n) for n in nodes])
    nM = len(nodes)
    if weights == "equal":
        el.update(refc=123456, indc=[1, 1, 1, 0, 0, 0], weights=np.ones(nM))
    elif weights == "random":
        el.update(refc=123456, indc=[1, 1, 1, 0, 0, 0], weights=rng.uniform(0.5, 2.0, nM))
    elif weights == "rotational":
        el.update(refc=123456, indc=[1, 1, 1, 1 + nM, 1 + nM, 1 + nM],
                  weights=np.concatenate([rng.uniform(0.5, 2.0, nM), rng.uniform(0.5, 2.0, nM)]))
    elif weights != "default":
        raise ValueError(weights)
    return el


def cylinder_spiders(nL=8, nC=12, L=1.0, R=0.1, t=0.004, weights="default", seed=0, gap=0.1):
    """Shell cylinder along z of ANDES4 elements (nC around, nL along) with a
    weighted-average-motion spider (WAVGM / RBE3) at each end, as the flexible
    spiders of the reference's cylinder test model (testModels.C:139-260):
    the reference node on the axis at each end takes the weighted motion of
    its ring; a stiff tube beam of length `gap` joins it to an external node
    further out.  Free-free: the two external nodes carry the boundary."""
    rng = np.random.default_rng(seed)
    th = 2.0 * np.pi * np.arange(nC) / nC
    rings = [np.stack([R * np.cos(th), R * np.sin(th), np.full(nC, z)], 1) for z in np.linspace(0.0, L, nL + 1)]
    xyz = list(np.concatenate(rings))
    ring = lambda i: np.arange(nC) + i * nC
    conn = []
    for i in range(nL):
        a, b = ring(i), ring(i + 1)
        for j in range(nC):
            conn.append(np.array([a[j], a[(j + 1) % nC], b[(j + 1) % nC], b[j]]))
    n0 = len(xyz)
    xyz += [[0.0, 0.0, 0.0], [0.0, 0.0, L], [0.0, 0.0, -gap], [0.0, 0.0, L + gap]]
    ref0, ref1, ext0, ext1 = n0, n0 + 1, n0 + 2, n0 + 3
    nsh = len(conn)
    conn += [np.array([ext0, ref0]), np.array([ref1, ext1])]
    props = np.zeros((nsh + 2, NPROP))
    props[:nsh, 0], props[:nsh, 1], props[:nsh, 2] = 2.1e11, 0.3, 7850.0
    props[:nsh, 3] = 2.1e11 / 2.6
    props[:nsh, 4:8] = t
    props[nsh:] = _stiff_beam_props(2)
    status = np.zeros(len(xyz), np.int32)
    status[[ext0, ext1]] = 1
    melcon = np.array([24] * nsh + [11, 11], np.int32)
    part = _finish(np.array(xyz), status, melcon, conn, props)
    part["rbe3"] = [_wavgm(ref0, ring(0), weights, rng), _wavgm(ref1, ring(nL), weights, rng)]
    return part


def block_spider(nx=6, ny=4, nz=4, lx=0.3, ly=0.2, lz=0.2, weights="random", seed=0, n_ext=4, gap=0.05):
    """HEX8 block (hex_block, n_ext external nodes on its x=0 and x=lx faces)
    whose x=lx face is tied by a weighted-average-motion element (WAVGM /
    RBE3) to a reference node off the face centre; a stiff beam joins the
    reference node to one more external node.  The independent nodes carry
    three DOFs (solid nodes), the reference node six."""
    rng = np.random.default_rng(seed + 1)
    p = hex_block(nx, ny, nz, lx, ly, lz, jitter=0.05, seed=seed, n_ext=n_ext)
    xyz = p["xyz"]
    face = np.nonzero(np.isclose(xyz[:, 0], lx))[0]
    nn = len(xyz)
    ref, ext = nn, nn + 1
    xyz = np.vstack([xyz, [[lx + gap, 0.45 * ly, 0.55 * lz], [lx + 2 * gap, 0.45 * ly, 0.55 * lz]]])
    mp, mm = p["mpmnpc"], p["mmnpc"]
    conn = [mm[mp[e] - 1:mp[e + 1] - 1] - 1 for e in range(len(p["melcon"]))] + [np.array([ref, ext])]
    props = np.vstack([p["eprop"], _stiff_beam_props(1, v=(0.0, 0.0, 1.0))])
    status = np.concatenate([p["status"], [0, 1]]).astype(np.int32)
    melcon = np.concatenate([p["melcon"], [11]]).astype(np.int32)
    part = _finish(xyz, status, melcon, conn, props)
    part["rbe3"] = [_wavgm(ref, face, weights, rng)]
    return part


def config1(seed=0):
    """48x24x24 HEX8 steel block, 0.48x0.24x0.24 m, +-5 % jitter, 12 external
    nodes on the x=0 and x=max faces (72 boundary DOFs), 20 component modes."""
    return hex_block(48, 24, 24, 0.48, 0.24, 0.24, jitter=0.05, seed=seed, n_ext=12)


def write_nastran(part, path, loads=None):
    """Writes a part as Nastran bulk data in free-field format with every
    real at full double precision (GRID, CHEXA/CQUAD4/CBAR, MAT1,
    PSOLID/PSHELL/PBAR, ASET1 for the external nodes, SPC1 for fixed DOFs,
    RBE3 for weighted-average-motion elements with explicit weights, FORCE
    for nodal loads), so that the part read back is bit-identical.
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
            # PID MID A I1 I2 J NSM, blank fields up to C1..F2, then the shear
            # factors K1 K2 (fields 17 and 18, one free-field line)
            out.append("PBAR,%d,%d,%s%s%s,%s" % (e + 1, e + 1, ",".join(r(v) for v in (*pr[e, 4:8], pr[e, 13])),
                                                 "," * 10, r(pr[e, 11]), r(pr[e, 12])))
    ext = [i + 1 for i in range(len(status)) if status[i] == 1]
    for k in range(0, len(ext), 6):
        out.append("ASET1,123456," + ",".join(str(v) for v in ext[k:k + 6]))
    for i in range(len(status)):
        if status[i] < 0:
            comp = "".join(str(j + 1) for j in range(6) if (-int(status[i]) >> j) & 1)
            out.append("SPC1,1,%s,%d" % (comp, i + 1))
    for k, el in enumerate(part.get("rbe3") or []):
        # one group per independent node and weight set (WT, C, G); an element
        # without weights is written with weight 1.0 on the translations
        w = np.asarray(el.get("weights", []), float)
        indc = list(el["indc"]) if len(w) else [1, 1, 1, 0, 0, 0]
        if not len(w):
            w = np.ones(len(el["nodes"]))
        fields = [str(len(mel) + k + 1), "", str(int(el["ref"]) + 1), str(int(el.get("refc", 123456)) if len(el.get("weights", [])) else 123456)]
        for ic in sorted({c for c in indc if c > 0}):
            comp = "".join(str(c + 1) for c in range(6) if indc[c] == ic)
            for a, n in enumerate(el["nodes"]):
                fields += [r(w[ic - 1 + a]), comp, str(int(n) + 1)]
        out.append("RBE3," + ",".join(fields[:8]))
        for b in range(8, len(fields), 8):
            out.append("+," + ",".join(fields[b:b + 8]))
    for sid, node, f in loads or []:
        out.append("FORCE,%d,%d,0,1.0,%s,%s,%s" % (sid, node, *(r(v) for v in f)))
    out.append("ENDDATA")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def random_part(rng, target_dof=None, kinds=("plate", "hex", "beam")):
    """One synthetic part of the batched / served mixes: type uniform over
    shell plate (nx, ny in [10, 80]), HEX8 block (n in [6, 24] per axis) and
    beam frame (50 to 500 elements, tube section), 2 to 16 external nodes
    (at least 4 for solid blocks, whose nodes have no rotations),
    0 to 10 component modes.  With target_dof the size is chosen to come
    close to that DOF count, for every type.
    Returns (part, ngen)."""
    kind = kinds[int(rng.integers(len(kinds)))]
    n_ext = int(rng.integers(2, 17))
    ngen = int(rng.integers(0, 11))
    seed = int(rng.integers(1 << 30))
    if kind == "plate":
        if target_dof:
            nx = ny = max(4, int(round(np.sqrt(target_dof / 6.0))) - 1)
        else:
            nx, ny = int(rng.integers(10, 81)), int(rng.integers(10, 81))
        i, j = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
        i, j = i.ravel(), j.ravel()
        edge = np.nonzero((i == 0) | (i == nx) | (j == 0) | (j == ny))[0]
        ext = np.random.default_rng(seed).choice(edge, size=min(n_ext, len(edge)), replace=False)
        th = np.exp(np.random.default_rng(seed + 1).normal(np.log(0.005), 0.3, nx * ny))
        part = plate(nx, ny, 0.02 * nx, 0.02 * ny, thickness=th, ext_nodes=ext)
    elif kind == "hex":
        if target_dof:
            n = max(3, int(round((target_dof / 3.0) ** (1.0 / 3.0))) - 1)
            nx = ny = nz = n
        else:
            nx, ny, nz = (int(rng.integers(6, 25)) for _ in range(3))
        # solid nodes carry translations only: fewer than three (non-collinear)
        # external nodes leave a rigid rotation of the block free
        part = hex_block(nx, ny, nz, 0.01 * nx, 0.01 * ny, 0.01 * nz, jitter=0.05, seed=seed, n_ext=max(n_ext, 4))
    else:
        n_el = int(max(50, target_dof / 6.0)) if target_dof else int(rng.integers(50, 501))
        part = beam_frame(n_el=n_el, seed=seed, n_ext=n_ext)
    return part, ngen


def random_loads(rng, part, nlc, n_nodes=4, scale=1000.0):
    """nlc load cases of nodal forces N(0, scale) on n_nodes random nodes
    each (three force components per node): [(node1, dof1, value), ...]."""
    nnod = len(part["xyz"])
    cases = []
    for _ in range(nlc):
        nodes = rng.choice(nnod, size=min(n_nodes, nnod), replace=False)
        cases.append([(int(n) + 1, d + 1, float(rng.normal(0.0, scale))) for n in nodes for d in range(3)])
    return cases


def served_request(rng, median_dof=20000, cap_dof=300000):
    """One request of the served mix: part type as random_part, DOF count
    lognormal (median 20k, sigma 1, capped at 300k), 1 to 8 load cases of
    N(0, 1 kN) nodal forces and N(0, 1 mm) boundary displacements for the
    recovery.  Returns (part, ngen, loads, ub)."""
    dof = min(cap_dof, float(np.exp(rng.normal(np.log(median_dof), 1.0))))
    part, ngen = random_part(rng, target_dof=max(dof, 600.0))
    nlc = int(rng.integers(1, 9))
    loads = random_loads(rng, part, nlc)
    n2 = sum(6 for s in part["status"] if s == 1)
    ub = rng.normal(0.0, 1e-3, (n2, nlc))
    return part, ngen, loads, ub
