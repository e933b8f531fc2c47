"""ctypes binding of ``libssmx.so`` (the HIP propagator library).

The library is built in-tree by ``superspinsim_amd.build.build_library`` (or
``__graft_entry__.build()``) into ``superspinsim_amd/lib/libssmx.so``.  There
is no fallback: every entry point of the package that simulates raises when
the library or a HIP device is missing.
"""

import ctypes
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.environ.get("SSMX_LIB") or os.path.join(LIB_DIR, "libssmx.so")

SSMX_ROTATING = 1
SSMX_COEF_WAVE = 2
SSMX_READ_FULL = 4
SSMX_READ_INTEGRAL = 8
SSMX_STORE_INPUTS = 16

MAX_N = 64
MAX_N_WIDE = 512
MAX_GEN = 8
MAX_EXP = 6
MAX_NODE = 3
MAX_OBS = 2
MAX_TAYLOR = 40
#: n_inst + n_waves limit of ssmx_propagate_queue_noalloc
NOALLOC_MAX_QUEUE = 0x7fffffff
ADJ_MAX_TAYLOR = 10
ADJ_MAX_THETA = 16
ADJ_MAX_SLOTS = 32
#: n > 64 adjoint: matrix and transposed slots staged per row
ADJ_WIDE_MAX_SLOTS = 16


class PropagateDesc(ctypes.Structure):
    """Mirror of ``ssmx_propagate_desc`` (include/ssmx.h)."""
    _fields_ = [
        ("n", ctypes.c_int32), ("n_inst", ctypes.c_int32),
        ("n_gen", ctypes.c_int32), ("slots", ctypes.c_int32),
        ("n_exp", ctypes.c_int32), ("n_node", ctypes.c_int32),
        ("fine", ctypes.c_int32), ("n_out", ctypes.c_int32),
        ("taylor", ctypes.c_int32), ("substeps", ctypes.c_int32),
        ("n_obs", ctypes.c_int32), ("flags", ctypes.c_int32),
        ("first_out", ctypes.c_int32), ("partner", ctypes.c_int32),
        ("ckpt_every", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("coef_inst_stride", ctypes.c_int64),
        ("h", ctypes.c_double), ("t0", ctypes.c_double),
        ("dt", ctypes.c_double),
    ]


class AdjointDesc(ctypes.Structure):
    """Mirror of ``ssmx_adjoint_desc`` (include/ssmx.h)."""
    _fields_ = [
        ("n", ctypes.c_int32), ("n_inst", ctypes.c_int32),
        ("n_gen", ctypes.c_int32), ("slots", ctypes.c_int32),
        ("slots_t", ctypes.c_int32), ("n_exp", ctypes.c_int32),
        ("n_node", ctypes.c_int32), ("fine", ctypes.c_int32),
        ("n_out", ctypes.c_int32), ("taylor", ctypes.c_int32),
        ("taylor_adj", ctypes.c_int32), ("n_obs", ctypes.c_int32),
        ("flags", ctypes.c_int32), ("first_out", ctypes.c_int32),
        ("partner", ctypes.c_int32), ("ckpt_every", ctypes.c_int32),
        ("n_theta", ctypes.c_int32), ("theta_gen", ctypes.c_int32),
        ("grad_gen", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("coef_inst_stride", ctypes.c_int64),
        ("h", ctypes.c_double), ("t0", ctypes.c_double),
        ("dt", ctypes.c_double),
    ]


class DenseDesc(ctypes.Structure):
    """Mirror of ``ssmx_dense_desc`` (include/ssmx.h)."""
    _fields_ = [
        ("n", ctypes.c_int32), ("n_inst", ctypes.c_int32),
        ("n_gen", ctypes.c_int32), ("n_exp", ctypes.c_int32),
        ("n_node", ctypes.c_int32), ("fine", ctypes.c_int32),
        ("n_out", ctypes.c_int32), ("squarings", ctypes.c_int32),
        ("flags", ctypes.c_int32), ("n_doubles", ctypes.c_int32),
        ("n_singles", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("coef_inst_stride", ctypes.c_int64),
        ("h", ctypes.c_double), ("scale", ctypes.c_double),
    ]


class ChunkDesc(ctypes.Structure):
    """Mirror of ``ssmx_chunk_desc`` (include/ssmx.h)."""
    _fields_ = [(f, ctypes.c_int32) for f in (
        "n", "n_inst", "n_chunks", "chunk_out", "n_out", "pack", "n_obs",
        "reserved")]


class Segment(ctypes.Structure):
    """Mirror of ``ssmx_segment``."""
    _fields_ = [("t_start", ctypes.c_double), ("t_end", ctypes.c_double),
                ("amp", ctypes.c_double), ("omega", ctypes.c_double),
                ("phase", ctypes.c_double)]


class Wave(ctypes.Structure):
    """Mirror of ``ssmx_wave``."""
    _fields_ = [("base", ctypes.c_double * MAX_GEN),
                ("seg_begin", ctypes.c_int32 * MAX_GEN),
                ("seg_count", ctypes.c_int32 * MAX_GEN)]


#: every symbol include/ssmx.h declares
EXPORTS = ("ssmx_propagate", "ssmx_propagate_dense", "ssmx_chunk_combine",
           "ssmx_propagate_queue",
           "ssmx_propagate_queue_noalloc", "ssmx_adjoint",
           "ssmx_adjoint_workspace_bytes", "ssmx_last_error", "ssmx_version",
           "ssmx_device_count")

_lib = None


class SsmxError(RuntimeError):
    """A library call returned a negative code."""


def load():
    """Load ``libssmx.so`` once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} not found: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    # torch first: its bundled HIP runtime (soname libamdhip64.so.7) then
    # satisfies this library's dependency.  Loaded the other way round, the
    # process would map two HIP runtimes (torch asks for "libamdhip64.so",
    # which does not match the soname already loaded), and launches from
    # this library would find no device.
    import torch  # noqa: F401
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    lib.ssmx_propagate.argtypes = [ctypes.POINTER(PropagateDesc)] + [vp] * 13
    lib.ssmx_propagate.restype = ctypes.c_int
    lib.ssmx_propagate_queue.argtypes = (
        [ctypes.POINTER(PropagateDesc)] + [vp] * 13 +
        [ctypes.c_int32, ctypes.c_int32, vp, vp])
    lib.ssmx_propagate_queue.restype = ctypes.c_int
    lib.ssmx_propagate_queue_noalloc.argtypes = (
        [ctypes.POINTER(PropagateDesc)] + [vp] * 12 +
        [ctypes.c_int32, ctypes.c_int32, vp, vp])
    lib.ssmx_propagate_queue_noalloc.restype = ctypes.c_int
    lib.ssmx_propagate_dense.argtypes = [ctypes.POINTER(DenseDesc)] + \
        [vp] * 7
    lib.ssmx_propagate_dense.restype = ctypes.c_int
    lib.ssmx_chunk_combine.argtypes = [ctypes.POINTER(ChunkDesc)] + [vp] * 5
    lib.ssmx_chunk_combine.restype = ctypes.c_int
    lib.ssmx_adjoint.argtypes = [ctypes.POINTER(AdjointDesc)] + [vp] * 17
    lib.ssmx_adjoint.restype = ctypes.c_int
    lib.ssmx_adjoint_workspace_bytes.argtypes = [
        ctypes.POINTER(AdjointDesc), ctypes.POINTER(ctypes.c_uint64)]
    lib.ssmx_adjoint_workspace_bytes.restype = ctypes.c_int
    lib.ssmx_last_error.argtypes = []
    lib.ssmx_last_error.restype = ctypes.c_char_p
    lib.ssmx_version.argtypes = []
    lib.ssmx_version.restype = ctypes.c_int
    lib.ssmx_device_count.argtypes = []
    lib.ssmx_device_count.restype = ctypes.c_int
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = load().ssmx_last_error().decode(errors="replace")
        raise SsmxError(f"ssmx error {rc}: {msg}")
    return rc
