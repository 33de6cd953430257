"""ctypes binding of ``libddhip.so`` (include/ddhip.h). The product path calls the HIP kernels
only through this module; a missing library is an error, never a fallback."""

from __future__ import annotations

import ctypes
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "libddhip.so"

DDH_FP64, DDH_MIXED = 0, 1
PRECISIONS = {"fp64": DDH_FP64, "mixed": DDH_MIXED}
#: The impedance operator's per-device table: fields per grid point, scalars per device.
IMP_TABLE_FIELDS, IMP_ION_TABLE_FIELDS, IMP_SCALARS = 48, 64, 8


class Batch(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("NP", ctypes.c_int32), ("nseg", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("pt", ctypes.c_void_p), ("sid", ctypes.c_void_p),
                ("seg", ctypes.c_void_p), ("sc", ctypes.c_void_p)]


class Structure(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("NP", ctypes.c_int32), ("nl", ctypes.c_int32),
                ("nseg", ctypes.c_int32), ("lid", ctypes.c_void_p), ("sid", ctypes.c_void_p),
                ("flags", ctypes.c_void_p), ("i0", ctypes.c_void_p), ("unit", ctypes.c_void_p),
                ("unit_sum", ctypes.c_void_p)]


class TOpts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_iter", ctypes.c_int32),
                ("bernoulli", ctypes.c_int32), ("zimt_current", ctypes.c_int32),
                ("failure_mode", ctypes.c_int32)]


class Opts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_iter", ctypes.c_int32),
                ("start_iter", ctypes.c_int32), ("strict", ctypes.c_int32),
                ("bernoulli", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("jv_mode", ctypes.c_int32), ("lanes", ctypes.c_int32)]


#: The symbols include/ddhip.h declares.
SYMBOLS = ("ddh_abi_version", "ddh_last_error", "ddh_workspace_bytes", "ddh_solve_batch",
           "ddh_solve_batch_states", "ddh_adjoint_workspace_bytes", "ddh_adjoint",
           "ddh_block_sweep", "ddh_assemble", "ddh_assemble_ions",
           "ddh_impedance_workspace_bytes", "ddh_impedance", "ddh_noalloc_scratch_bytes",
           "ddh_solve_batch_noalloc", "ddh_transient_workspace_bytes", "ddh_transient",
           "ddh_impedance_ions_workspace_bytes", "ddh_impedance_ions", "ddh_build",
           "ddh_graded_layout", "ddh_adjoint_vjp_workspace_bytes", "ddh_adjoint_vjp")

_lib = None


def load():
    """The loaded library (once per process). Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        raise RuntimeError(f"{LIB} is missing: build it first (python -m jaxdd_amd._build, or "
                           "__graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(str(LIB))
    vp, i32 = ctypes.c_void_p, ctypes.c_int32
    lib.ddh_abi_version.restype = ctypes.c_int
    lib.ddh_last_error.restype = ctypes.c_char_p
    lib.ddh_workspace_bytes.restype = ctypes.c_size_t
    lib.ddh_workspace_bytes.argtypes = [i32, i32, i32, i32]
    lib.ddh_solve_batch.restype = ctypes.c_int
    lib.ddh_solve_batch.argtypes = [ctypes.POINTER(Batch), vp, i32, ctypes.POINTER(Opts), vp, vp,
                                    vp, vp, vp, ctypes.c_size_t, vp]
    lib.ddh_solve_batch_states.restype = ctypes.c_int
    lib.ddh_solve_batch_states.argtypes = [ctypes.POINTER(Batch), vp, i32, ctypes.POINTER(Opts),
                                           vp, vp, vp, vp, vp, vp, vp, ctypes.c_size_t, vp]
    lib.ddh_adjoint_workspace_bytes.restype = ctypes.c_size_t
    lib.ddh_adjoint_workspace_bytes.argtypes = [i32, i32, i32]
    lib.ddh_adjoint.restype = ctypes.c_int
    lib.ddh_adjoint.argtypes = [ctypes.POINTER(Batch), vp, i32, vp, vp, vp, i32, vp, vp, vp, i32,
                                vp, vp, vp, ctypes.c_size_t, vp]
    lib.ddh_assemble.restype = ctypes.c_int
    lib.ddh_assemble.argtypes = [ctypes.POINTER(Batch), vp, vp, vp, i32, vp, vp, vp, vp, vp, vp,
                                 vp]
    lib.ddh_assemble_ions.restype = ctypes.c_int
    lib.ddh_assemble_ions.argtypes = [ctypes.POINTER(Batch), vp, vp, vp, i32, vp, vp, vp, vp, vp,
                                      vp]
    lib.ddh_impedance_workspace_bytes.restype = ctypes.c_size_t
    lib.ddh_impedance_workspace_bytes.argtypes = [i32, i32]
    lib.ddh_impedance.restype = ctypes.c_int
    lib.ddh_impedance.argtypes = [ctypes.POINTER(Batch), vp, vp, vp, vp, i32, ctypes.c_double,
                                  i32, vp, vp, vp, ctypes.c_size_t, vp]
    lib.ddh_impedance_ions_workspace_bytes.restype = ctypes.c_size_t
    lib.ddh_impedance_ions_workspace_bytes.argtypes = [i32, i32]
    lib.ddh_impedance_ions.restype = ctypes.c_int
    lib.ddh_impedance_ions.argtypes = [ctypes.POINTER(Batch), vp, vp, vp, vp, vp, i32,
                                       ctypes.c_double, i32, vp, vp, vp, ctypes.c_size_t, vp]
    lib.ddh_noalloc_scratch_bytes.restype = ctypes.c_size_t
    lib.ddh_noalloc_scratch_bytes.argtypes = []
    lib.ddh_solve_batch_noalloc.restype = ctypes.c_int
    lib.ddh_solve_batch_noalloc.argtypes = [ctypes.POINTER(Batch), vp, i32, ctypes.POINTER(Opts),
                                            vp, vp, vp, vp, vp]
    lib.ddh_transient_workspace_bytes.restype = ctypes.c_size_t
    lib.ddh_transient_workspace_bytes.argtypes = [i32, i32]
    lib.ddh_transient.restype = ctypes.c_int
    lib.ddh_transient.argtypes = [ctypes.POINTER(Batch), vp, vp, vp, vp, vp, i32, vp,
                                  ctypes.POINTER(TOpts), i32, vp, vp, vp, vp, vp, vp, vp,
                                  ctypes.c_size_t, vp]
    lib.ddh_adjoint_vjp_workspace_bytes.restype = ctypes.c_size_t
    lib.ddh_adjoint_vjp_workspace_bytes.argtypes = [i32, i32, i32, i32]
    lib.ddh_adjoint_vjp.restype = ctypes.c_int
    lib.ddh_adjoint_vjp.argtypes = [ctypes.POINTER(Batch), vp, i32, vp, vp, vp, i32, vp, vp, vp,
                                    i32, vp, vp, vp, vp, ctypes.c_size_t, vp]
    lib.ddh_graded_layout.restype = ctypes.c_int
    lib.ddh_graded_layout.argtypes = [i32, i32, i32, vp, ctypes.c_double, ctypes.c_double, vp,
                                      vp, vp, vp, vp, vp, vp, vp]
    lib.ddh_build.restype = ctypes.c_int
    lib.ddh_build.argtypes = [ctypes.POINTER(Structure), vp, vp, vp, i32, i32, vp, vp, vp, vp,
                              vp, vp, vp, vp, vp, vp]
    lib.ddh_block_sweep.restype = ctypes.c_int
    lib.ddh_block_sweep.argtypes = [i32, i32, i32, i32, vp, vp, vp, vp, i32, vp, vp]
    if lib.ddh_abi_version() != 5:
        raise RuntimeError(f"{LIB}: ABI version {lib.ddh_abi_version()}, expected 5")
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().ddh_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")
