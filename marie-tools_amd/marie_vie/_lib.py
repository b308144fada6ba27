"""ctypes binding of libvie_hip.so (the C ABI declared in include/vie_op.h).

The library is loaded from this package directory and nowhere else; when it
is missing the import fails with an error that says how to build it.  There
is no CPU fallback: every operator call goes to the HIP library.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvie_hip.so")

VIE_OK = 0
VIE_ERR_ARG = -1
VIE_ERR_HIP = -2
VIE_ERR_OOM = -3
VIE_ERR_STATE = -4

VIE_BASIS_PWC = 0
VIE_BASIS_PWL = 1

VIE_OP_N = 0
VIE_OP_K = 1
VIE_OP_NH = 2
VIE_OP_KH = 3
VIE_OP_VIE = 4
VIE_OP_VIEH = 5

VIE_PREC_FP64 = 0
VIE_PREC_FP32 = 1

VIE_MEM_HOST = 0
VIE_MEM_DEVICE = 1

#: every entry point include/vie_op.h declares
EXPORTS = ("vie_create", "vie_destroy", "vie_last_error", "vie_set_stream", "vie_set_gram", "vie_fft_dims",
           "vie_set_green_tucker", "vie_set_green_dense", "vie_set_material",
           "vie_num_scatterers", "vie_apply", "vie_gmres", "vie_green_quadrature", "vie_autotune",
           "vie_box_dims", "vie_apply_noalloc", "vie_axis_schedule", "vie_cocr", "vie_bcocg")


class VieDesc(ctypes.Structure):
    _fields_ = [("n1", ctypes.c_int32), ("n2", ctypes.c_int32), ("n3", ctypes.c_int32),
                ("basis", ctypes.c_int32), ("res", ctypes.c_double), ("device", ctypes.c_int32)]


class VieError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libvie_hip status {code}: {msg}")
        self.code = code


_lib = None


def load() -> ctypes.CDLL:
    """Load libvie_hip.so from the package directory (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `make -C marie-tools_amd/csrc` "
                          "or __graft_entry__.build()")
    # One HIP runtime per process.  torch ships its own libamdhip64 (soname
    # libamdhip64.so.7, loaded by torch as "libamdhip64.so"); the library needs
    # "libamdhip64.so.7".  Loaded after torch, the library binds to torch's
    # runtime, so torch tensors, streams and the library share devices and
    # pointers.  Loaded first, it would bring /opt/rocm's copy, and torch,
    # imported later, would load a second runtime that finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i32 = ctypes.c_int32
    dbl = ctypes.c_double
    i64 = ctypes.c_int64
    lib.vie_create.argtypes = [ctypes.POINTER(VieDesc), ctypes.POINTER(P)]
    lib.vie_destroy.argtypes = [P]
    lib.vie_last_error.argtypes = [P]
    lib.vie_last_error.restype = ctypes.c_char_p
    lib.vie_set_stream.argtypes = [P, P]
    lib.vie_set_gram.argtypes = [P, dbl]
    lib.vie_fft_dims.argtypes = [P, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)]
    lib.vie_set_green_tucker.argtypes = [P, i32, i32, P, i32, i32, i32, P, P, P]
    lib.vie_set_green_dense.argtypes = [P, i32, P, i32]
    lib.vie_set_material.argtypes = [P, P, P, P, i64]
    lib.vie_num_scatterers.argtypes = [P]
    lib.vie_num_scatterers.restype = i64
    lib.vie_apply.argtypes = [P, i32, i32, P, P, i32, i32]
    lib.vie_gmres.argtypes = [P, i32, i32, P, P, i32, dbl, i32, i32, i32, P, P, P, P]
    lib.vie_green_quadrature.argtypes = [P, i32, dbl, i32, P, P, i32, i32, i32, P, i32]
    lib.vie_autotune.argtypes = [P, i32, i32, i32, P]
    lib.vie_box_dims.argtypes = [P, P, P, P]
    lib.vie_apply_noalloc.argtypes = [P, i32, i32, P, P, i32]
    lib.vie_axis_schedule.argtypes = [P, i32, P]
    lib.vie_cocr.argtypes = [P, i32, i32, P, P, i32, dbl, i32, i32, P, P, P, P]
    lib.vie_bcocg.argtypes = [P, i32, i32, P, P, i32, dbl, i32, i32, P, P, P, P]
    for name in EXPORTS:
        if name not in ("vie_last_error", "vie_num_scatterers"):
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def check(rc: int, handle=None) -> None:
    if rc != VIE_OK:
        msg = load().vie_last_error(handle)
        raise VieError(rc, msg.decode() if msg else "")
