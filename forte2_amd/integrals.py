"""Host integrals through libf2ints.so (numpy in, numpy out).

Mirrors forte2/integrals/integrals.py: overlap (:198), kinetic (:224),
nuclear (:250), nuclear_repulsion (:156), coulomb_2c (:903-934),
coulomb_3c (:811-862) and the test-only coulomb_4c.
"""

import ctypes

import numpy as np

from . import _native

_dp = ctypes.POINTER(ctypes.c_double)


def _lib():
    return _native.ints_lib()


def _check(rc):
    if rc != 0:
        raise RuntimeError(_lib().f2ints_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(_dp)


def overlap(basis):
    n = basis.size
    S = np.empty((n, n))
    _check(_lib().f2ints_overlap(basis.handle, _ptr(S)))
    return S


def kinetic(basis):
    n = basis.size
    T = np.empty((n, n))
    _check(_lib().f2ints_kinetic(basis.handle, _ptr(T)))
    return T


def nuclear(basis, atoms):
    n = basis.size
    V = np.empty((n, n))
    Z = np.ascontiguousarray([float(z) for z, _ in atoms])
    xyz = np.ascontiguousarray(np.array([p for _, p in atoms], dtype=float).reshape(-1))
    _check(_lib().f2ints_nuclear(basis.handle, len(atoms), _ptr(Z), _ptr(xyz), _ptr(V)))
    return V


def nuclear_repulsion(atoms):
    Z = np.ascontiguousarray([float(z) for z, _ in atoms])
    xyz = np.ascontiguousarray(np.array([p for _, p in atoms], dtype=float).reshape(-1))
    return float(_lib().f2ints_nuclear_repulsion(len(atoms), _ptr(Z), _ptr(xyz)))


def coulomb_2c(basis1, basis2=None):
    b2 = basis1 if basis2 is None else basis2
    M = np.empty((basis1.size, b2.size))
    _check(_lib().f2ints_coulomb_2c(basis1.handle, b2.handle, _ptr(M)))
    return M


def coulomb_3c(aux, basis, shell_begin=0, shell_end=None, out=None):
    """(P|mn) for aux shells [shell_begin, shell_end): array (nP, nbf, nbf)."""
    offs = aux.shell_offsets()
    se = aux.nshells if shell_end is None else shell_end
    nP = offs[se] - offs[shell_begin]
    n = basis.size
    if out is None:
        out = np.empty((nP, n, n))
    elif out.shape != (nP, n, n) or not out.flags.c_contiguous or out.dtype != np.float64:
        raise ValueError("out has the wrong shape/layout")
    _check(_lib().f2ints_coulomb_3c(aux.handle, basis.handle, shell_begin, se, _ptr(out)))
    return out


def ket_pairs(basis):
    """(a, b) int arrays of the significant shell pairs a >= b, in the order
    of coulomb_3c_pairs."""
    lib = _lib()
    n = lib.f2ints_ket_pairs(basis.handle, None, None, 0)
    if n < 0:
        raise RuntimeError(lib.f2ints_last_error().decode())
    a = np.empty(n, dtype=np.int32)
    b = np.empty(n, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    lib.f2ints_ket_pairs(basis.handle, a.ctypes.data_as(ip), b.ctypes.data_as(ip), n)
    return a, b


def packed_columns(basis, q0, q1):
    """Packed columns of the significant pairs [q0, q1): na*nb for a > b,
    na(na+1)/2 for a == b (the layout of coulomb_3c_pairs)."""
    a, b = ket_pairs(basis)
    if not (0 <= q0 <= q1 <= len(a)):
        raise ValueError("bad pair range")
    ls = np.array([l for l, *_ in basis.shells], dtype=np.int64)
    na, nb = 2 * ls[a[q0:q1]] + 1, 2 * ls[b[q0:q1]] + 1
    return int(np.where(a[q0:q1] == b[q0:q1], na * (na + 1) // 2, na * nb).sum())


def coulomb_3c_pairs(aux, basis, q0, q1, out):
    """(P|ab) of all aux functions and the pairs [q0, q1) of ket_pairs,
    packed into out[:, :ncols] (out: (naux, ldo) C-contiguous float64)."""
    if out.ndim != 2 or out.shape[0] != aux.size or not out.flags.c_contiguous or out.dtype != np.float64:
        raise ValueError("out has the wrong shape/layout")
    _check(_lib().f2ints_coulomb_3c_pairs(aux.handle, basis.handle, int(q0), int(q1), _ptr(out), out.shape[1]))
    return out


def coulomb_3c_pairs_t(aux, basis, q0, q1, out):
    """coulomb_3c_pairs stored column-major: out[c, P] (out: (>= ncols, naux)
    C-contiguous float64), each packed column a contiguous row."""
    if out.ndim != 2 or out.shape[1] != aux.size or not out.flags.c_contiguous or out.dtype != np.float64:
        raise ValueError("out has the wrong shape/layout")
    if out.shape[0] < packed_columns(basis, q0, q1):
        raise ValueError("out has fewer rows than the pairs have packed columns")
    _check(_lib().f2ints_coulomb_3c_pairs_t(aux.handle, basis.handle, int(q0), int(q1), _ptr(out), out.shape[1]))
    return out


def coulomb_3c_pairs_aux(aux, basis, s0, s1, q0, q1, out):
    """coulomb_3c_pairs for the aux shells [s0, s1) only: out (rows = the
    functions of those shells, ldo) C-contiguous float64."""
    offs = aux.shell_offsets()
    if not (0 <= s0 < s1 <= len(offs) - 1):
        raise ValueError("bad aux shell range")
    if out.ndim != 2 or out.shape[0] != offs[s1] - offs[s0] or not out.flags.c_contiguous or out.dtype != np.float64:
        raise ValueError("out has the wrong shape/layout")
    _check(_lib().f2ints_coulomb_3c_pairs_aux(aux.handle, basis.handle, int(s0), int(s1), int(q0), int(q1),
                                               _ptr(out), out.shape[1]))
    return out


def coulomb_3c_pairs_aux_t(aux, basis, s0, s1, q0, q1, out):
    """coulomb_3c_pairs_aux stored column-major: out[c, P - first] (out:
    (>= ncols, ldo >= the block's rows) C-contiguous float64)."""
    offs = aux.shell_offsets()
    if not (0 <= s0 < s1 <= len(offs) - 1):
        raise ValueError("bad aux shell range")
    if out.ndim != 2 or out.shape[1] < offs[s1] - offs[s0] or not out.flags.c_contiguous or out.dtype != np.float64:
        raise ValueError("out has the wrong shape/layout")
    if out.shape[0] < packed_columns(basis, q0, q1):
        raise ValueError("out has fewer rows than the pairs have packed columns")
    _check(_lib().f2ints_coulomb_3c_pairs_aux_t(aux.handle, basis.handle, int(s0), int(s1), int(q0), int(q1),
                                                 _ptr(out), out.shape[1]))
    return out


def coulomb_3c_sum(aux, basis):
    """V[m, n] = sum_P (P|mn) over every function of `aux` (nbf x nbf)."""
    n = basis.size
    V = np.empty((n, n))
    _check(_lib().f2ints_coulomb_3c_sum(aux.handle, basis.handle, _ptr(V)))
    return V


def coulomb_4c(basis):
    n = basis.size
    out = np.empty((n, n, n, n))
    _check(_lib().f2ints_coulomb_4c(basis.handle, _ptr(out)))
    return out


def boys(nmax, T):
    out = np.empty(nmax + 1)
    _check(_lib().f2ints_boys(nmax, float(T), _ptr(out)))
    return out


def cart2sph(l):
    out = np.empty((2 * l + 1, (l + 1) * (l + 2) // 2))
    _check(_lib().f2ints_cart2sph(l, _ptr(out)))
    return out


def set_threads(n):
    _lib().f2ints_set_threads(int(n))
