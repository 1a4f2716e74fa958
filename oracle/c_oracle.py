"""ctypes wrapper of the C + OpenMP oracle (oracle/c/mgpcg.c).

TEST INFRASTRUCTURE ONLY (tests and bench.py's CPU-baseline leg).
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from . import build_c

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(build_c.LIB):
            build_c.build(verbose=False)
        L = ctypes.CDLL(build_c.LIB)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.oracle_pcg.restype = I
        L.oracle_pcg.argtypes = [I, I, I, P, P, P, P, P, P, I, I, I, I, D, P, ctypes.POINTER(D), ctypes.POINTER(I)]
        L.oracle_vcycle.restype = I
        L.oracle_vcycle.argtypes = [I, I, I, P, P, P, P, P, P, P, I]
        _lib = L
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def pcg(widths, obstacle, rhs0, x0=None, *, outlet=False, remove_mean=True, multigrid=True, max_iters=500,
        tolerance=1e-10, origin=(0.0, 0.0, 0.0)):
    """Returns (x, iterations, relative_residual, converged)."""
    shape = tuple(len(w) for w in widths)
    w = [_c(v, np.float64) for v in widths]
    o = _c(origin, np.float64)
    obst = _c(obstacle, np.int32)
    r = _c(rhs0, np.float64)
    x = np.zeros(shape) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    rel = ctypes.c_double()
    conv = ctypes.c_int()
    it = lib().oracle_pcg(shape[0], shape[1], shape[2], w[0].ctypes.data, w[1].ctypes.data, w[2].ctypes.data,
                          o.ctypes.data, obst.ctypes.data, r.ctypes.data, int(outlet), int(remove_mean),
                          1 if multigrid else 0, int(max_iters), float(tolerance), x.ctypes.data,
                          ctypes.byref(rel), ctypes.byref(conv))
    return x, int(it), float(rel.value), bool(conv.value)


def vcycle(widths, obstacle, r, *, outlet=False, origin=(0.0, 0.0, 0.0)):
    shape = tuple(len(w) for w in widths)
    w = [_c(v, np.float64) for v in widths]
    o = _c(origin, np.float64)
    obst = _c(obstacle, np.int32)
    rr = _c(r, np.float64)
    z = np.zeros(shape)
    lib().oracle_vcycle(shape[0], shape[1], shape[2], w[0].ctypes.data, w[1].ctypes.data, w[2].ctypes.data,
                        o.ctypes.data, obst.ctypes.data, rr.ctypes.data, z.ctypes.data, int(outlet))
    return z
