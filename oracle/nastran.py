"""ORACLE — test infrastructure only (see oracle/__init__.py).

Nastran bulk-data subset reader for the reference's FE part fixtures
(src/vpmReducer/vpmReducerTests/FEparts/*.nas, read in the reference by FFl,
which is absent from the reference tree).  Cards: GRID, CQUAD4, CHEXA, CBAR,
MAT1, PSHELL, PSOLID, PBAR, PBARL (TUBE), ASET, ASET1, SPC1 (with "G1 THRU G2"
ranges), RBE2, RBE3, FORCE,
MOMENT, in free (comma), small (8-column) and
large (16-column, '*') field formats.  Returns a part dict (see
fedempy_amd.parts) plus the RBE2 constraint list and the RBE3 elements in the
WAVGM form of oracle.wavgm.  How FFl turns an RBE3 card into WAVGM
properties is not in the reference tree; the reading here: the independent
nodes in order of first appearance, per component the weight of every node
summed over the groups that list it with that component, components with
equal weights sharing one weight set, no group for a component: no coupling.
"""
import re

import numpy as np

NPROP = 16


def _num(s):
    s = s.strip()
    if not s:
        return None
    try:
        return int(s)
    except ValueError:
        pass
    t = s.upper().replace("D", "E")
    m = re.match(r"^([+-]?[0-9.]+)([+-][0-9]+)$", t)  # Nastran implied exponent 1.5-3
    if m and "E" not in t:
        t = m.group(1) + "E" + m.group(2)
    try:
        return float(t)
    except ValueError:
        return s  # a name field (e.g. the PBARL section type)


def _grid_list(vals):
    """(listed, ranged) grid ids of SPC1/ASET1 grid fields: "G1 THRU G2"
    adds the ids strictly between G1 and G2 to `ranged` (they need not
    exist); anything but an integer or THRU raises ValueError."""
    listed, ranged = [], []
    vals = [x for x in vals if x is not None]
    k = 0
    while k < len(vals):
        x = vals[k]
        if isinstance(x, str) and x.upper() == "THRU":
            if not listed or k + 1 >= len(vals) or not isinstance(vals[k + 1], int) or vals[k + 1] < listed[-1]:
                raise ValueError("THRU needs grid ids before and after it")
            ranged.extend(range(listed[-1] + 1, vals[k + 1]))
            listed.append(vals[k + 1])
            k += 2
            continue
        if not isinstance(x, int) or x <= 0:
            raise ValueError("grid field %r is not a grid id" % (x,))
        listed.append(x)
        k += 1
    return listed, ranged


def _cards(text):
    """Yield (name, [fields]) with continuations merged."""
    lines = [ln.rstrip("\n") for ln in text.splitlines()]
    cur = None
    for ln in lines:
        if not ln.strip() or ln.startswith("$"):
            continue
        head = ln[:8].strip()
        if ln.startswith("*") or ln.startswith("+"):
            if cur is None:
                continue
            if "," in ln:
                cur[1].extend(f.strip() for f in ln.split(",")[1:])
            elif cur[2]:  # large field continuation
                body = ln[8:72]
                cur[1].extend(body[i:i + 16] for i in range(0, len(body), 16))
            else:
                body = ln[8:72]
                cur[1].extend(body[i:i + 8] for i in range(0, len(body), 8))
            continue
        if cur is not None:
            yield cur[0], cur[1]
        if "," in ln:
            parts = ln.split(",")
            name = parts[0].strip().upper()
            cur = [name.rstrip("*"), [p.strip() for p in parts[1:]], name.endswith("*")]
        elif head.endswith("*"):
            body = ln[8:72]
            cur = [head[:-1].upper(), [body[i:i + 16] for i in range(0, len(body), 16)], True]
        else:
            body = ln[8:72]
            cur = [head.upper(), [body[i:i + 8] for i in range(0, len(body), 8)], False]
    if cur is not None:
        yield cur[0], cur[1]


def read(path):
    with open(path) as f:
        text = f.read()
    grids, elems, mats, props, aset, rbe2, spc, loads = {}, [], {}, {}, [], [], {}, []
    rbe3 = []
    aset_range = []
    known = {"GRID", "CHEXA", "CQUAD4", "CBAR", "MAT1", "PSHELL", "PSOLID", "PBAR", "PBARL", "ASET", "ASET1",
             "SPC1", "RBE2", "RBE3", "FORCE", "MOMENT"}
    for name, fl in _cards(text):
        if name not in known:
            continue
        v = [_num(x) for x in fl] + [None] * 24
        if name == "GRID":
            grids[int(v[0])] = (v[2] or 0.0, v[3] or 0.0, v[4] or 0.0)
        elif name == "CHEXA":
            elems.append((44, int(v[0]), int(v[1]), [int(x) for x in v[2:10]], None))
        elif name == "CQUAD4":
            elems.append((24, int(v[0]), int(v[1]), [int(x) for x in v[2:6]], None))
        elif name == "CBAR":
            elems.append((11, int(v[0]), int(v[1]), [int(x) for x in v[2:4]],
                          tuple(x or 0.0 for x in (v[4], v[5], v[6]))))
        elif name == "MAT1":
            mats[int(v[0])] = dict(E=v[1], G=v[2], nu=v[3], rho=v[4] or 0.0)
        elif name == "PSHELL":
            props[int(v[0])] = dict(kind="shell", mid=int(v[1]), t=v[2])
        elif name == "PSOLID":
            props[int(v[0])] = dict(kind="solid", mid=int(v[1]))
        elif name == "PBAR":
            vv = v
            props[int(v[0])] = dict(kind="bar", mid=int(v[1]), A=vv[2], I1=vv[3], I2=vv[4], J=vv[5],
                                    nsm=vv[6] or 0.0, K1=vv[16], K2=vv[17])
        elif name == "PBARL":
            # TUBE: DIM1 outer radius, DIM2 inner radius; thin-tube shear area A/2
            if str(fl[3]).strip().upper() != "TUBE":
                raise ValueError("PBARL type %s not supported" % fl[3])
            R, r = v[8], v[9]
            A = np.pi * (R * R - r * r)
            I = np.pi * (R ** 4 - r ** 4) / 4.0
            props[int(v[0])] = dict(kind="bar", mid=int(v[1]), A=A, I1=I, I2=I, J=2 * I, nsm=0.0, K1=0.5, K2=0.5)
        elif name == "ASET":
            vals = [x for x in v if x is not None]
            aset.extend(int(vals[i]) for i in range(0, len(vals), 2))
        elif name == "ASET1":
            listed, ranged = _grid_list(v[1:])
            aset.extend(listed)
            aset_range.extend(ranged)
        elif name == "SPC1":
            mask = sum(1 << (int(c) - 1) for c in str(fl[1]).strip() if c in "123456")
            listed, ranged = _grid_list(v[2:])
            for x in listed + ranged:
                spc[x] = spc.get(x, 0) | mask
        elif name in ("FORCE", "MOMENT"):
            if v[2]:
                raise ValueError("only the basic coordinate system is supported")
            F = v[3] or 0.0
            loads.append((int(v[0]), int(v[1]), 0 if name == "FORCE" else 3,
                          tuple(F * (x or 0.0) for x in v[4:7])))
        elif name == "RBE3":
            groups, state = [], 0
            for x in v[4:]:
                if x is None:
                    continue
                if isinstance(x, str):
                    if x.upper() == "UM":
                        raise ValueError("RBE3 %d: UM dependent DOFs are not supported" % v[0])
                    if x.upper() in ("ALPHA", "TREF"):
                        break
                    raise ValueError("RBE3 %d: unexpected field %r" % (v[0], x))
                if isinstance(x, float):
                    groups.append([x, 0, []])
                    state = 1
                elif state == 1:
                    groups[-1][1] = sum(1 << (int(c) - 1) for c in str(x) if c in "123456")
                    state = 2
                elif state == 2:
                    groups[-1][2].append(x)
                else:
                    raise ValueError("RBE3 %d: unexpected field %r" % (v[0], x))
            rbe3.append((int(v[2]), int(v[3] or 0), groups))
        elif name == "RBE2":
            vals = [x for x in v if x is not None]
            rbe2.append((int(vals[1]), str(int(vals[2])), [int(x) for x in vals[3:]]))
    ids = sorted(grids)
    idx = {g: i for i, g in enumerate(ids)}
    xyz = np.array([grids[g] for g in ids], float)
    status = np.zeros(len(ids), np.int32)
    for g, m in spc.items():
        if g in idx:
            status[idx[g]] = -m
    for g in aset:
        status[idx[g]] = 1
    for g in aset_range:
        if g in idx:
            status[idx[g]] = 1
    melcon, conn, eprop = [], [], []
    for t, eid, pid, nodes, orient in elems:
        p = props[pid]
        m = mats[p["mid"]]
        pr = np.zeros(NPROP)
        pr[0], pr[1], pr[2] = m["E"], m["nu"], m["rho"]
        pr[3] = m["G"] if m["G"] is not None else m["E"] / (2 * (1 + m["nu"]))
        if t == 24:
            pr[4:8] = p["t"]
        elif t == 11:
            pr[4], pr[5], pr[6], pr[7] = p["A"], p["I1"], p["I2"], p["J"]
            pr[8:11] = orient
            pr[11] = p["K1"] if p["K1"] is not None else 0.0
            pr[12] = p["K2"] if p["K2"] is not None else 0.0
            pr[13] = p["nsm"]
        melcon.append(t)
        conn.append([idx[n] for n in nodes])
        eprop.append(pr)
    wv = []
    for ref, refc, groups in rbe3:
        order = []
        for _, _, gl in groups:
            for g in gl:
                if g not in order:
                    order.append(g)
        if not order:
            raise ValueError("RBE3 without independent grids")
        pos = {g: a for a, g in enumerate(order)}
        indc, wts = [0] * 6, []
        for c in range(6):
            w = np.zeros(len(order))
            hit = False
            for wt, comps, gl in groups:
                if (comps >> c) & 1:
                    for g in gl:
                        w[pos[g]] += wt
                        hit = True
            if not hit:
                continue
            for prev in range(c):
                if indc[prev] > 0 and np.array_equal(w, np.asarray(wts[indc[prev] - 1:indc[prev] - 1 + len(order)])):
                    indc[c] = indc[prev]
                    break
            if not indc[c]:
                indc[c] = len(wts) + 1
                wts.extend(w)
        wv.append(dict(ref=idx[ref], refc=refc, nodes=[idx[g] for g in order], indc=indc, weights=np.array(wts)))
    sizes = [len(c) for c in conn]
    part = dict(xyz=xyz, status=status, melcon=np.array(melcon, np.int32),
                mpmnpc=np.concatenate([[1], 1 + np.cumsum(sizes)]).astype(np.int32),
                mmnpc=(np.concatenate([np.asarray(c) for c in conn]) + 1).astype(np.int32),
                eprop=np.array(eprop), node_ids=np.array(ids),
                rbe2=[(idx[gn], cm, [idx[d] for d in dep]) for gn, cm, dep in rbe2], rbe3=wv,
                loads=[(sid, idx[g], d0, f) for sid, g, d0, f in loads])
    return part
