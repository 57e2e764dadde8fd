"""FedemReducer: drives the drop-in entry points of libfedem_reducer_core.so
the way fedempy does (PythonAPI/src/fedempy/reducer.py:39-119).

Same constructor, methods, argument meaning and return values as the
reference class: ``solver_init(options)`` passes command-line options to
``initSolverArgs`` (three arguments; a program name is prepended when the
first option starts with '-'), ``run(options)`` calls ``reducePart(first,
False)`` and returns its status (0 success, negative error).  Results are the
.fmx/.fsm files named by the options or derived from -linkfile.
"""
from ctypes import POINTER, c_bool, c_char_p, c_double, c_int, cdll

from . import LIB_PATH


class FedemReducer:
    def __init__(self, lib_path=None, solver_options=None):
        self._lib = cdll.LoadLibrary(lib_path or LIB_PATH)
        self._lib.initSolverArgs.restype = c_int
        self._lib.reducePart.restype = c_int
        self._lib.solvePartDis.restype = c_int
        self._lib.solvePartDis.argtypes = [c_int, POINTER(c_int), POINTER(c_double)]
        self._first = True
        self.solver_init(solver_options)

    @staticmethod
    def _argv(options):
        if isinstance(options, list):
            arr = (c_char_p * len(options))(*[o.encode("utf-8") for o in options])
            return c_int(len(options)), arr
        raise TypeError(f"Expected {list}, got {type(options)}.")

    def solver_init(self, options):
        """Registers (first part) or updates the options; 0 on success."""
        if options is None:
            return 0
        options = list(options)
        if options and options[0][:1] == "-":
            options.insert(0, "fedem_reducer")
        argc, argv = self._argv(options)
        return self._lib.initSolverArgs(argc, argv, c_bool(self._first))

    def run(self, options=None):
        """Reduces the part named by -linkfile; 0 on success, < 0 on error."""
        self.solver_init(options)
        status = self._lib.reducePart(c_bool(self._first), c_bool(False))
        self._first = False
        return status

    def solve(self, nodes, options=None):
        """Direct static solve of the part (solvePartDis): returns (status,
        displacements (len(nodes), 3)) of the listed node ids."""
        import numpy as np

        self.solver_init(options)
        self._first = False
        ids = (c_int * len(nodes))(*[int(n) for n in nodes])
        out = (c_double * (3 * len(nodes)))()
        status = self._lib.solvePartDis(len(nodes), ids, out)
        return status, np.array(out[:]).reshape(len(nodes), 3)

    def release(self):
        """freeSingeltons: drops the option registry and the part."""
        self._lib.freeSingeltons()
