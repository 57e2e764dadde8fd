"""Continuous-batching reduction server (fgpu_server_* of
libfedem_reducer_core.so): slots with their own HIP streams take parts in
arrival order, admitted by device-memory footprint; every request is a
reduction plus its load cases on the resident factor (reduced load vectors
and recovered internal displacements)."""
import ctypes

import numpy as np

from . import LIB_PATH

_lib = None


def _load():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(LIB_PATH)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        lib.fgpu_server_create.argtypes = [ctypes.POINTER(P), I, D]
        lib.fgpu_server_destroy.argtypes = [P]
        lib.fgpu_server_destroy.restype = None
        lib.fgpu_server_submit.argtypes = [P, I, P, P, I, P, P, P, P, I, I, P, P, P, P, I, P]
        lib.fgpu_server_set_option.argtypes = [P, ctypes.c_char_p, D]
        lib.fgpu_server_wait.argtypes = [P, I]
        lib.fgpu_server_precision.argtypes = [P, I]
        lib.fgpu_server_sizes.argtypes = [P, I, P]
        lib.fgpu_server_result.argtypes = [P, I, P, P, P, P, P, P]
        lib.fgpu_server_error.argtypes = [P, I, ctypes.c_char_p, I]
        lib.fgpu_server_release.argtypes = [P, I]
        lib.fgpu_server_device_bytes.argtypes = [P]
        lib.fgpu_server_device_bytes.restype = ctypes.c_int64
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Server:
    def __init__(self, slots=4, budget_gb=200.0):
        self.lib = _load()
        self.h = ctypes.c_void_p()
        if self.lib.fgpu_server_create(ctypes.byref(self.h), int(slots), float(budget_gb)) != 0:
            raise RuntimeError("fgpu_server_create failed")
        self._keep = {}

    def submit(self, part, ngen=0, loads=None, ub=None):
        """loads: list of load cases, each a list of (node (1-based), dof
        (1..6), value); ub: (ndof2, nlc) boundary displacements in reduced
        order, or None."""
        arrs = [np.ascontiguousarray(part[k], dt) for k, dt in
                (("xyz", np.float64), ("status", np.int32), ("melcon", np.int32), ("mpmnpc", np.int32),
                 ("mmnpc", np.int32), ("eprop", np.float64))]
        loads = loads or []
        nlc = len(loads)
        lptr = np.concatenate([[0], np.cumsum([len(c) for c in loads])]).astype(np.int32)
        flat = [x for c in loads for x in c]
        lnode = np.array([x[0] for x in flat], np.int32)
        ldof = np.array([x[1] for x in flat], np.int32)
        lval = np.array([x[2] for x in flat], np.float64)
        ubf = None if ub is None else np.asfortranarray(ub, np.float64)
        nub = 0 if ub is None else ubf.shape[0]
        ubf = None if ubf is None else ubf.ravel(order="F").copy()
        xyz, st, mel, mp, mm, ep = arrs
        rid = self.lib.fgpu_server_submit(self.h, len(st), _p(xyz), _p(st), len(mel), _p(mel), _p(mp), _p(mm),
                                          _p(ep), int(ngen), nlc, _p(lptr), _p(lnode), _p(ldof), _p(lval), nub,
                                          _p(ubf))
        if rid < 0:
            raise ValueError("fgpu_server_submit failed (%d)" % rid)
        self._keep[rid] = (arrs, lptr, lnode, ldof, lval, ubf)
        return rid

    def set_option(self, name, value):
        """numeric option (fgpu_set_option's names) of the requests submitted from now on"""
        if self.lib.fgpu_server_set_option(self.h, name.encode(), float(value)) != 0:
            raise ValueError("unknown option " + name)

    def wait(self, rid):
        return self.lib.fgpu_server_wait(self.h, rid)

    def result(self, rid):
        st = self.wait(rid)
        if st != 0:
            buf = ctypes.create_string_buffer(512)
            self.lib.fgpu_server_error(self.h, rid, buf, 512)
            raise RuntimeError("request %d failed (%d): %s" % (rid, st, buf.value.decode()))
        s = np.zeros(5, np.int32)
        self.lib.fgpu_server_sizes(self.h, rid, _p(s))
        nd, n1, n2, nlc, nev = (int(v) for v in s)
        SK, SM, G = np.zeros((nd, nd), order="F"), np.zeros((nd, nd), order="F"), np.zeros((nd, 3), order="F")
        Fr, Ui = np.zeros((nd, max(nlc, 1)), order="F"), np.zeros((n1, max(nlc, 1)), order="F")
        t = np.zeros(4)
        self.lib.fgpu_server_result(self.h, rid, _p(SK), _p(SM), _p(G), _p(Fr), _p(Ui), _p(t))
        return dict(SK=SK, SM=SM, G=G, Fred=Fr[:, :nlc], Ui=Ui[:, :nlc], ndof2=n2,
                    queue_ms=t[0], service_ms=t[1], latency_ms=t[2], footprint=t[3],
                    mixed=self.lib.fgpu_server_precision(self.h, rid) == 1)

    def device_bytes(self):
        """device memory the slots' buffers hold now"""
        return int(self.lib.fgpu_server_device_bytes(self.h))

    def release(self, rid):
        self.lib.fgpu_server_release(self.h, rid)
        self._keep.pop(rid, None)

    def close(self):
        if self.h:
            self.lib.fgpu_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
