"""ctypes bindings of the native libraries (``include/f2jk.h``, ``include/f2ints.h``,
``include/f2probe.h``).

Nothing here falls back to Python: if a library is missing the loader raises,
so a GPU path can never silently run on the host.
"""

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))

c_int64_p = ctypes.POINTER(ctypes.c_int64)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_voidp_p = ctypes.POINTER(ctypes.c_void_p)

# enum values of include/f2jk.h
F2JK_F64, F2JK_F32 = 0, 1
PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2}
F2JK_WANT_J, F2JK_WANT_K, F2JK_TIMING = 1, 2, 16
F2JK_NOALLOC_SCRATCH_BYTES = 64 << 20
PHASES = ("prep", "j_gamma", "j_acc", "kx", "ks", "k_reduce", "total")

F2JK_SYMBOLS = (
    "f2jk_ctx_create",
    "f2jk_ctx_destroy",
    "f2jk_set_B",
    "f2jk_get_shape",
    "f2jk_set_tuning",
    "f2jk_workspace_size",
    "f2jk_build_jk",
    "f2jk_build_jk_noalloc",
    "f2jk_build_j",
    "f2jk_build_k",
    "f2jk_last_timings",
    "f2jk_last_launches",
    "f2jk_work",
    "f2jk_group_workspace_size",
    "f2jk_group_build_jk",
    "f2jk_half_transform",
    "f2jk_last_error",
)


class NativeError(RuntimeError):
    pass


_jk = None
_ints = None


def _load(name):
    path = os.path.join(PKG, name)
    if not os.path.exists(path):
        raise NativeError(
            f"{name} is not built ({path} missing); run `python -m forte2_amd._build` "
            "or __graft_entry__.build()"
        )
    return ctypes.CDLL(path)


def jk_lib():
    """Load libf2jk.so (HIP) and declare its prototypes."""
    global _jk
    if _jk is not None:
        return _jk
    lib = _load("libf2jk.so")
    V, I, I64, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
    sz = ctypes.c_size_t
    lib.f2jk_ctx_create.argtypes = [I, ctypes.POINTER(V)]
    lib.f2jk_ctx_destroy.argtypes = [V]
    lib.f2jk_set_B.argtypes = [V, I, V, I64, I64]
    lib.f2jk_get_shape.argtypes = [V, c_int64_p, c_int64_p]
    lib.f2jk_set_tuning.argtypes = [V, ctypes.c_char_p, I64]
    lib.f2jk_workspace_size.argtypes = [V, I, c_int64_p, I, I, ctypes.POINTER(sz)]
    lib.f2jk_build_jk.argtypes = [V, I, c_voidp_p, c_int64_p, c_int64_p, c_voidp_p, c_voidp_p, I, I, V, sz, V]
    lib.f2jk_build_jk_noalloc.argtypes = [V, I, c_voidp_p, c_int64_p, c_int64_p, c_voidp_p, c_voidp_p, I, I, V]
    lib.f2jk_build_j.argtypes = [V, I, c_voidp_p, I64, c_voidp_p, I, I, V, sz, V]
    lib.f2jk_build_k.argtypes = [V, I, c_voidp_p, c_int64_p, c_int64_p, c_voidp_p, I, I, V, sz, V]
    lib.f2jk_last_timings.argtypes = [V, c_double_p, I]
    lib.f2jk_last_launches.argtypes = [V, ctypes.POINTER(I), I]
    lib.f2jk_work.argtypes = [V, I, c_int64_p, I, I, c_double_p, c_double_p]
    lib.f2jk_group_workspace_size.argtypes = [c_voidp_p, I, ctypes.POINTER(I), c_int64_p, I, I, ctypes.POINTER(sz)]
    lib.f2jk_group_build_jk.argtypes = [c_voidp_p, I, ctypes.POINTER(I), c_voidp_p, c_int64_p, c_int64_p, c_voidp_p,
                                        c_voidp_p, I, I, c_voidp_p, ctypes.POINTER(sz), V]
    lib.f2jk_half_transform.argtypes = [V, V, I64, I64, I64, I64, V, I64, I, V, sz, V]
    lib.f2jk_last_error.argtypes = []
    lib.f2jk_last_error.restype = ctypes.c_char_p
    for s in F2JK_SYMBOLS:
        if s != "f2jk_last_error":
            getattr(lib, s).restype = I
    _jk = lib
    return lib


def jk_check(rc, what=""):
    if rc != 0:
        msg = jk_lib().f2jk_last_error().decode()
        if rc == -3:
            raise MemoryError(f"{what}: {msg}")
        if rc in (-1, -4, -5):
            raise ValueError(f"{what}: {msg}")
        raise NativeError(f"{what} failed ({rc}): {msg}")


def ptr_array(ptrs):
    arr = (ctypes.c_void_p * len(ptrs))()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr


def i64_array(vals):
    arr = (ctypes.c_int64 * len(vals))()
    for i, v in enumerate(vals):
        arr[i] = int(v)
    return arr


def ints_lib():
    """Load libf2ints.so (host integral engine) and declare its prototypes."""
    global _ints
    if _ints is not None:
        return _ints
    lib = _load("libf2ints.so")
    from . import _ints_proto

    _ints_proto.declare(lib)
    _ints = lib
    return lib


F2PROBE_SYMBOLS = ("f2probe_sink_len", "f2probe_mfma", "f2probe_read", "f2probe_last_error")
F2PROBE_KIND = {"bf16": 0, "fp32": 1, "fp64": 2}
_probe = None


def probe_lib():
    """Load libf2probe.so (MFMA / HBM peak probes, ``include/f2probe.h``)."""
    global _probe
    if _probe is not None:
        return _probe
    lib = _load("libf2probe.so")
    V, I, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    lib.f2probe_sink_len.argtypes = [c_int64_p]
    lib.f2probe_mfma.argtypes = [I, I, V, I64, V, c_double_p, c_double_p]
    lib.f2probe_read.argtypes = [V, I64, I, V, I64, V, c_double_p]
    lib.f2probe_last_error.argtypes = []
    lib.f2probe_last_error.restype = ctypes.c_char_p
    for s in F2PROBE_SYMBOLS[:-1]:
        getattr(lib, s).restype = I
    _probe = lib
    return lib


def probe_check(rc, what=""):
    if rc != 0:
        msg = probe_lib().f2probe_last_error().decode()
        raise (ValueError if rc == -1 else NativeError)(f"{what} failed ({rc}): {msg}")
