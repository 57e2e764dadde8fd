"""Test parts with a singular or an indefinite K_ii (shared by the CPU and
GPU tests of the singular-pivot handling)."""
import numpy as np

from fedempy_amd import parts


def hinged_blocks():
    """Two 2x2x2 HEX8 blocks joined at a single corner node: the second
    block can rotate rigidly about the hinge, so K_ii has a three-dimensional
    null space (true zero pivots).  External nodes: four nodes of the first
    block's x = 0 face."""
    a = parts.hex_block(2, 2, 2, 0.02, 0.02, 0.02, jitter=0.0, seed=1, ext_nodes=[0, 2, 6, 8])
    b = parts.hex_block(2, 2, 2, 0.02, 0.02, 0.02, jitter=0.0, seed=2, ext_nodes=[])
    na = len(a["xyz"])
    hinge_a = int(np.argmax(a["xyz"].sum(axis=1)))  # corner (0.02, 0.02, 0.02)
    xyz_b = b["xyz"] + 0.02
    keep = np.ones(len(xyz_b), bool)
    keep[0] = False  # b's origin corner is the hinge
    idx_b = np.full(len(xyz_b), -1)
    idx_b[0] = hinge_a
    idx_b[keep] = na + np.arange(keep.sum())
    xyz = np.vstack([a["xyz"], xyz_b[keep]])
    status = np.concatenate([a["status"], np.zeros(keep.sum(), np.int32)])
    conn_a = (a["mmnpc"] - 1).reshape(-1, 8)
    conn_b = idx_b[(b["mmnpc"] - 1).reshape(-1, 8)]
    conn = np.vstack([conn_a, conn_b])
    props = np.vstack([a["eprop"], b["eprop"]])
    return parts._finish(xyz, status, np.full(len(conn), 44, np.int32), list(conn), props)


def indefinite_block():
    """3x2x2 HEX8 block whose element 5 has the Young's modulus negated: K_ii
    stays nonsingular (smallest |eigenvalue| 7e-4 of the largest) but has
    four negative eigenvalues."""
    p = parts.hex_block(3, 2, 2, 0.03, 0.02, 0.02, seed=4, n_ext=4)
    p["eprop"][5, 0] = -2.1e11
    return p
