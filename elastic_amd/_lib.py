"""ctypes binding of libelastic_hip.so (the C-ABI of include/elastic_hip.h).

The library is built in-tree by elastic_amd.build; there is no fallback: if it
is missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libelastic_hip.so")
HEADER = os.path.join(HERE, "..", "include", "elastic_hip.h")

OK, ZERO_DIAG, MAXIT, NEG_JACOBIAN, OOM, HIP_ERROR, BAD_ARG, NOT_READY, BREAKDOWN, BUSY = range(10)
STATUS = {0: "ok", 1: "zero diagonal", 2: "maxit reached", 3: "non-positive Jacobian",
          4: "out of device memory", 5: "HIP error", 6: "bad argument", 7: "not ready",
          8: "Krylov breakdown", 9: "busy"}
PCG, GMRES, BICGSTAB = 1, 2, 4
DIAG, BDIAG3 = 2, 6
FP64, MIXED = 64, 32

I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
I64P = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class ElasticError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__("%s (%d)%s" % (STATUS.get(code, "error"), code, (": " + msg) if msg else ""))


class Stats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("outer", C.c_int32), ("status", C.c_int32),
                ("spmv_count", C.c_int32), ("xkx", C.c_double), ("norm_x", C.c_double),
                ("norm_r", C.c_double), ("norm_m_r", C.c_double), ("conv", C.c_double),
                ("rel_res", C.c_double), ("t_device_ms", C.c_double), ("t_total_ms", C.c_double),
                ("op64_from", C.c_int32), ("reserved", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Loads the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ElasticError(NOT_READY, "%s is missing: run `python -m elastic_amd.build`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.elastic_create.argtypes = [C.POINTER(vp), C.c_int32]
    L.elastic_destroy.argtypes = [vp]
    L.elastic_last_error.argtypes = [vp]
    L.elastic_last_error.restype = C.c_char_p
    L.elastic_reord.argtypes = [C.c_int32, C.c_int32, C.c_int32, I32P, I32P]
    L.elastic_numeq.argtypes = [C.c_int32, C.c_int32, I32P, I32P, C.POINTER(C.c_int32)]
    L.elastic_csrstruct.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, I32P,
                                    I32P, I32P, vp, C.POINTER(C.c_int64)]
    L.elastic_set_structure.argtypes = [vp, C.c_int32, C.c_int64, I32P, I32P]
    L.elastic_set_values.argtypes = [vp, F64P, F64P]
    L.elastic_get_values.argtypes = [vp, F64P, F64P]
    L.elastic_spmv.argtypes = [vp, C.c_int32, F64P, F64P]
    L.elastic_solve.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, F64P, F64P, C.c_double,
                                C.c_int32, C.c_int32, C.POINTER(Stats)]
    L.elastic_upload_rhs.argtypes = [vp, F64P]
    L.elastic_solve_resident.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                         C.c_int32, C.c_int32, C.POINTER(Stats)]
    L.elastic_download_x.argtypes = [vp, F64P]
    L.elastic_time_spmv.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]
    L.elastic_layout_info.argtypes = [vp, I64P]
    L.elastic_set_tile_nnz.argtypes = [vp, C.c_int32]
    L.elastic_set_tile_blocks.argtypes = [vp, C.c_int32]
    L.elastic_set_option.argtypes = [vp, C.c_char_p, C.c_int64]
    L.elastic_get_option.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64)]
    L.elastic_product_timing.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                         C.c_int32]
    L.elastic_autotune.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_double)]
    L.elastic_get_bdiag.argtypes = [vp, F64P]
    L.elastic_get_block_precond.argtypes = [vp, C.c_int32, F64P]
    L.elastic_spmv_csrc_noalloc.argtypes = [C.c_int32, C.c_int64, vp, vp, vp, vp, vp, vp, vp]
    L.elastic_set_mesh.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                   C.c_int32, I32P, F64P, I32P, I32P, F64P, I32P, C.c_int32, I32P,
                                   I32P, F64P]
    L.elastic_assemble.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, vp, vp]
    L.elastic_time_assembly.argtypes = [vp, C.c_int32, C.POINTER(C.c_double)]
    L.elastic_stress.argtypes = [vp, vp, C.c_int32, vp, vp]
    L.elastic_j2_setup.argtypes = [vp, C.c_double, C.c_double, vp]
    L.elastic_j2_increment.argtypes = [vp, C.c_double]
    L.elastic_j2_assemble.argtypes = [vp, C.c_int32, C.POINTER(C.c_double)]
    L.elastic_j2_correct.argtypes = [vp]
    L.elastic_j2_step.argtypes = [vp, C.c_double]
    L.elastic_j2_commit.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    L.elastic_j2_get_u.argtypes = [vp, F64P]
    L.elastic_j2_get_residual.argtypes = [vp, F64P]
    L.elastic_j2_get_state.argtypes = [vp, C.c_int32, vp, vp]
    L.elastic_batch_create.argtypes = [C.POINTER(vp), C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    L.elastic_batch_destroy.argtypes = [vp]
    L.elastic_batch_last_error.argtypes = [vp]
    L.elastic_batch_last_error.restype = C.c_char_p
    L.elastic_batch_submit.argtypes = [vp, vp, C.c_int32, C.c_double, C.c_int32, C.c_int64]
    L.elastic_batch_step.argtypes = [vp]
    L.elastic_batch_poll.argtypes = [vp, I64P, vp, C.c_int32, C.POINTER(C.c_int32)]
    L.elastic_batch_active.argtypes = [vp, C.POINTER(C.c_int32)]
    _lib = L
    return L


def declared_functions() -> list[str]:
    """Names of every entry point declared in include/*.h."""
    inc = os.path.dirname(HEADER)
    names = []
    for f in sorted(os.listdir(inc)):
        if f.endswith(".h"):
            txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(inc, f)).read(), flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(elastic_\w+)\s*\(", txt, re.M)
    return sorted(set(names))


# ------------------------------------------------------------ host set-up
def reord(mesh) -> np.ndarray:
    """RCM node order (Reord.f:1-44): perm[k-1] = old node at position k."""
    perm = np.zeros(mesh.nnode, np.int32)
    _check(None, lib().elastic_reord(mesh.nnode, mesh.numel, mesh.nen,
                                     np.ascontiguousarray(mesh.ix, np.int32), perm))
    return perm


def numeq(mesh, perm):
    """Equation numbers (Numeq.f:139-191): returns (id with eq numbers, neq)."""
    idn = np.ascontiguousarray(mesh.id, np.int32).copy()
    neq = C.c_int32()
    _check(None, lib().elastic_numeq(mesh.nnode, mesh.ndf, perm, idn, C.byref(neq)))
    return idn, neq.value


def csrstruct(mesh, idn, neq):
    """CSRC pattern (Csr.f:13-305): returns (ia, ja), 1-based."""
    ia = np.zeros(neq + 1, np.int32)
    nad = C.c_int64()
    ix = np.ascontiguousarray(mesh.ix, np.int32)
    _check(None, lib().elastic_csrstruct(mesh.nnode, mesh.numel, mesh.nen, mesh.ndf, neq, ix, idn,
                                         ia, None, C.byref(nad)))
    ja = np.zeros(max(nad.value, 1), np.int32)
    _check(None, lib().elastic_csrstruct(mesh.nnode, mesh.numel, mesh.nen, mesh.ndf, neq, ix, idn,
                                         ia, ja.ctypes.data_as(C.c_void_p), C.byref(nad)))
    return ia, ja[: nad.value]


def _check(ctx, rc, allow=()):
    if rc != OK and rc not in allow:
        msg = lib().elastic_last_error(ctx).decode() if ctx else ""
        raise ElasticError(rc, msg)
    return rc
