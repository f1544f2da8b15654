"""Basis sets: BSE-schema JSON data -> shells -> the host integral engine.

Follows forte2's basis assembly (forte2/system/build_basis.py:104-212,
480-502): shells in atom order, then JSON order; a general contraction is
split into one shell per coefficient column, each carrying all exponents; the
primitive normalisation is embedded in the coefficients and the contraction
is normalised to unity (libint2's Shell::renorm convention, axial Cartesian
x^l normalised), unless raw coefficients are requested (SAP density basis,
build_basis.py:262-320).
"""

import ctypes
import itertools
import json
import math
import os
import threading

import numpy as np

from . import _native

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "basis")


def _dfact(n):
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def load_atom_basis(name, Z):
    """electron_shells of element Z from the shipped JSON (or a .json path)."""
    path = name if name.endswith(".json") else os.path.join(DATA, name.lower() + ".json")
    if not os.path.exists(path):
        raise ValueError(f"basis set {name!r} is not available (no {path})")
    with open(path) as f:
        data = json.load(f)
    if str(Z) not in data["elements"]:
        raise ValueError(f"Element {Z} not found in basis set {name}.")
    return data["elements"][str(Z)]["electron_shells"]


def normalize_shell(l, exps, coefs):
    """libint2 convention: c_p *= N(alpha_p, l), then scale to unit self-overlap."""
    exps = np.asarray(exps, dtype=float)
    c = np.asarray(coefs, dtype=float).copy()
    df = _dfact(2 * l - 1)
    c *= np.sqrt((4.0 * exps) ** l * (2.0 * exps / math.pi) ** 1.5 / df)
    g = exps[:, None] + exps[None, :]
    norm = np.sum(np.outer(c, c) * df * (math.pi / g) ** 1.5 / (2.0 * g) ** l)
    return c / math.sqrt(norm)


class BasisSet:
    """Ordered list of pure shells (l, exponents, coefficients, centre in bohr)."""

    def __init__(self, name="custom"):
        self.name = name
        self.shells = []
        self.atom_of_shell = []
        self._handle = None

    def add(self, l, exps, coefs, center, normalize=True, atom=-1):
        coefs = normalize_shell(l, exps, coefs) if normalize else np.asarray(coefs, dtype=float)
        self.shells.append((int(l), np.asarray(exps, dtype=float), coefs, np.asarray(center, dtype=float)))
        self.atom_of_shell.append(atom)
        self._handle = None

    @property
    def size(self):
        return sum(2 * l + 1 for l, *_ in self.shells)

    @property
    def nshells(self):
        return len(self.shells)

    @property
    def max_l(self):
        return max((l for l, *_ in self.shells), default=0)

    def shell_offsets(self):
        off = [0]
        for l, *_ in self.shells:
            off.append(off[-1] + 2 * l + 1)
        return off

    @property
    def handle(self):
        """f2ints_basis* of the host integral engine (created on first use;
        safe when threads ask for it at once: one handle per basis)."""
        if self._handle is None:
            with _HANDLE_LOCK:
                if self._handle is None:
                    self._make_handle()
        return self._handle.ptr

    def _make_handle(self):
        lib = _native.ints_lib()
        ls = (ctypes.c_int * self.nshells)(*[s[0] for s in self.shells])
        nps = (ctypes.c_int * self.nshells)(*[len(s[1]) for s in self.shells])
        xyz = np.ascontiguousarray(np.concatenate([s[3] for s in self.shells]) if self.shells else np.zeros(0))
        ex = np.ascontiguousarray(np.concatenate([s[1] for s in self.shells]) if self.shells else np.zeros(0))
        co = np.ascontiguousarray(np.concatenate([s[2] for s in self.shells]) if self.shells else np.zeros(0))
        h = ctypes.c_void_p()
        dp = ctypes.POINTER(ctypes.c_double)
        rc = lib.f2ints_basis_create(self.nshells, ls, nps, xyz.ctypes.data_as(dp), ex.ctypes.data_as(dp),
                                     co.ctypes.data_as(dp), ctypes.byref(h))
        if rc != 0:
            raise ValueError(lib.f2ints_last_error().decode())
        self._handle = _Handle(h)


_HANDLE_LOCK = threading.Lock()


class _Handle:
    def __init__(self, ptr):
        self.ptr = ptr

    def __del__(self):
        try:
            _native.ints_lib().f2ints_basis_destroy(self.ptr)
        except Exception:
            pass


def build_basis(name, atoms, normalize=True):
    """BasisSet for [(Z, xyz_bohr)] from the named basis (build_basis.py:104-212)."""
    bs = BasisSet(name)
    cache = {}
    for ia, (Z, xyz) in enumerate(atoms):
        if Z not in cache:
            cache[Z] = load_atom_basis(name, Z)
        for shell in cache[Z]:
            ams = [int(a) for a in shell["angular_momentum"]]
            exps = [float(e) for e in shell["exponents"]]
            for l, col in itertools.zip_longest(ams, shell["coefficients"], fillvalue=ams[-1]):
                bs.add(l, exps, [float(c) for c in col], xyz, normalize=normalize, atom=ia)
    return bs


def build_sap_density_basis(atoms, name="sap_helfem_large"):
    """SAP potential as an s-type charge-density basis (build_basis.py:262-320):
    c_k e^{-a_k r^2} -> -c_k (a_k/pi)^{3/2} e^{-a_k r^2}, unnormalised."""
    raw = build_basis(name, atoms, normalize=False)
    dens = BasisSet(name + "-density")
    for (l, exps, coefs, center), atom in zip(raw.shells, raw.atom_of_shell):
        if l != 0:
            raise ValueError(f"SAP basis '{name}' contains a shell with l={l}")
        dens.add(0, exps, -coefs * (exps / math.pi) ** 1.5, center, normalize=False, atom=atom)
    return dens
