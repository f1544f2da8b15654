"""ctypes prototypes of libf2ints.so (include/f2ints.h)."""

import ctypes

F2INTS_SYMBOLS = (
    "f2ints_basis_create",
    "f2ints_basis_destroy",
    "f2ints_basis_nbf",
    "f2ints_max_l",
    "f2ints_overlap",
    "f2ints_kinetic",
    "f2ints_nuclear",
    "f2ints_nuclear_repulsion",
    "f2ints_coulomb_2c",
    "f2ints_coulomb_3c",
    "f2ints_ket_pairs",
    "f2ints_coulomb_3c_pairs",
    "f2ints_coulomb_3c_pairs_t",
    "f2ints_coulomb_3c_pairs_aux_t",
    "f2ints_coulomb_3c_pairs_aux",
    "f2ints_coulomb_3c_sum",
    "f2ints_coulomb_4c",
    "f2ints_boys",
    "f2ints_cart2sph",
    "f2ints_set_threads",
    "f2ints_last_error",
)


def declare(lib):
    V, I, I64, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
    Ip, Dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    lib.f2ints_basis_create.argtypes = [I, Ip, Ip, Dp, Dp, Dp, ctypes.POINTER(V)]
    lib.f2ints_basis_destroy.argtypes = [V]
    lib.f2ints_basis_nbf.argtypes = [V]
    lib.f2ints_basis_nbf.restype = I64
    lib.f2ints_max_l.argtypes = []
    lib.f2ints_overlap.argtypes = [V, Dp]
    lib.f2ints_kinetic.argtypes = [V, Dp]
    lib.f2ints_nuclear.argtypes = [V, I, Dp, Dp, Dp]
    lib.f2ints_nuclear_repulsion.argtypes = [I, Dp, Dp]
    lib.f2ints_nuclear_repulsion.restype = D
    lib.f2ints_coulomb_2c.argtypes = [V, V, Dp]
    lib.f2ints_coulomb_3c.argtypes = [V, V, I, I, Dp]
    lib.f2ints_ket_pairs.argtypes = [V, Ip, Ip, I64]
    lib.f2ints_ket_pairs.restype = I64
    lib.f2ints_coulomb_3c_pairs.argtypes = [V, V, I64, I64, Dp, I64]
    lib.f2ints_coulomb_3c_pairs_t.argtypes = [V, V, I64, I64, Dp, I64]
    lib.f2ints_coulomb_3c_pairs_aux.argtypes = [V, V, I, I, I64, I64, Dp, I64]
    lib.f2ints_coulomb_3c_pairs_aux_t.argtypes = [V, V, I, I, I64, I64, Dp, I64]
    lib.f2ints_coulomb_3c_sum.argtypes = [V, V, Dp]
    lib.f2ints_coulomb_4c.argtypes = [V, Dp]
    lib.f2ints_boys.argtypes = [I, D, Dp]
    lib.f2ints_cart2sph.argtypes = [I, Dp]
    lib.f2ints_set_threads.argtypes = [I]
    lib.f2ints_last_error.argtypes = []
    lib.f2ints_last_error.restype = ctypes.c_char_p
    for s in F2INTS_SYMBOLS:
        if s not in ("f2ints_basis_nbf", "f2ints_nuclear_repulsion", "f2ints_last_error", "f2ints_ket_pairs"):
            getattr(lib, s).restype = I
