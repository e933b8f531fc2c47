"""ctypes binding of ``cpu_ref.c`` (test oracle and CPU baseline only)."""

import ctypes
import os

import numpy as np

_LIB = None
_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build",
                     "libcpuref.so")


def build():
    """Compile ``libcpuref.so`` with the oracle Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.dirname(_PATH) + "/.."],
                   check=True)


def library():
    global _LIB
    if _LIB is None:
        if not os.path.exists(_PATH):
            raise FileNotFoundError(
                f"{_PATH} is missing; run `make -C oracle` first")
        lib = ctypes.CDLL(_PATH)
        d = ctypes.POINTER(ctypes.c_double)
        i = ctypes.c_int
        lib.cpuref_propagate.argtypes = [i, i, i, i, i, i, i, i, d, d, d,
                                         ctypes.c_double, i, d, i, d, d, i]
        lib.cpuref_propagate.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def propagate(gens, weights, coef, h, R, residual, rotating, d_step=None,
              s_step=None, threads=0):
    """Reference propagator core in C: returns TE (N, n, n).

    ``gens`` is (G, n, n) or, rotating, (S, G, n, n); ``coef`` (F, N, S, G).
    """
    gens = np.ascontiguousarray(gens, dtype=np.float64)
    weights = np.ascontiguousarray(weights, dtype=np.float64)
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    F, N, S, G = coef.shape
    J = weights.shape[0]
    n = gens.shape[-1]
    nd = 0 if d_step is None else d_step.shape[0]
    ns = 0 if s_step is None else s_step.size
    d = np.ascontiguousarray(d_step if nd else np.zeros((1, 2)),
                             dtype=np.float64)
    s = np.ascontiguousarray(s_step if ns else np.zeros(1), dtype=np.float64)
    te = np.empty((N, n, n), dtype=np.float64)
    flags = (1 if residual else 0) | (2 if rotating else 0)
    rc = library().cpuref_propagate(
        n, G, J, S, F, N, int(R), flags, _ptr(gens), _ptr(weights),
        _ptr(coef), float(h), nd, _ptr(d), ns, _ptr(s), _ptr(te),
        int(threads))
    if rc != 0:
        raise MemoryError("cpu_ref scratch allocation failed")
    return te
