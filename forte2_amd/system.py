"""Molecular system: geometry, basis sets, one-time host integrals.

Mirrors the parts of forte2's ``System`` (forte2/system/system.py:30-200,
405-484) that the SCF loops use: ``atoms``, ``Zsum``, ``nbf``, ``naux``,
``nmo``, ``nuclear_repulsion``, ``ints_overlap``, ``ints_hcore``,
``get_Xorth`` and the ``fock_builder`` slot (a device FockBuilder, created on
first use).  Integrals come from the host engine in libf2ints.so.
"""

import numpy as np

from . import basis as _basis
from . import geometry as _geom
from . import integrals as _ints


def canonical_orth(S, rtol=1e-8):
    """Canonical orthogonalisation X = U s^-1/2 over eigenvalues >= rtol*max
    (forte2/helpers/matrix_functions.py:95-152).  Returns (X, n_kept)."""
    w, U = np.linalg.eigh(S)
    if np.any(w < -1e-14):
        raise ValueError("The metric matrix must be positive semi-definite.")
    w[w < 0] = 0.0
    nd = int(np.searchsorted(w, rtol * w[-1]))
    return U[:, nd:] / np.sqrt(w[nd:]), len(w) - nd


class System:
    def __init__(self, xyz, basis_set, auxiliary_basis_set=None, unit="angstrom",
                 overlap_ortho_rtol=1e-8, df_ortho_rtol=None, device=None, jk_mem_thres_mb=None,
                 integral_backend="auto", cholesky_tei=False, cholesky_tol=1e-6):
        # B from a pivoted Cholesky decomposition of (mn|ls) instead of
        # density fitting: any auxiliary basis is ignored and naux is the rank,
        # known once B is built (forte2/system/system.py:123-124, 159-160, 341-352)
        if cholesky_tol < 0:
            raise ValueError("cholesky_tol must be non-negative.")
        self.cholesky_tei = bool(cholesky_tei)
        self.cholesky_tol = float(cholesky_tol)
        if self.cholesky_tei:
            auxiliary_basis_set = None
        self.xyz = xyz
        self.unit = unit
        self.atoms = _geom.parse_geometry(xyz, unit) if isinstance(xyz, str) else list(xyz)
        self.Zsum = sum(z for z, _ in self.atoms)
        self.basis_set = basis_set
        self.auxiliary_basis_set = auxiliary_basis_set
        self.basis = _basis.build_basis(basis_set, self.atoms)
        self.auxiliary_basis = (_basis.build_basis(auxiliary_basis_set, self.atoms)
                                if auxiliary_basis_set is not None else None)
        self.nbf = self.basis.size
        if self.cholesky_tei:
            self.naux = None
        else:
            self.naux = self.auxiliary_basis.size if self.auxiliary_basis is not None else 0
        self.df_ortho_rtol = df_ortho_rtol
        self.overlap_ortho_rtol = overlap_ortho_rtol
        self.nuclear_repulsion = _ints.nuclear_repulsion(self.atoms)
        self._S = None
        self._H = None
        self._V_sap = None
        self.Xorth, self.nmo = canonical_orth(self.ints_overlap(), overlap_ortho_rtol)
        self.two_component = False
        self.device = device
        self.jk_mem_thres_mb = jk_mem_thres_mb
        self.integral_backend = integral_backend
        self._fock_builder = None

    # ------------------------------------------------------------------ integrals
    def ints_overlap(self):
        if self._S is None:
            self._S = _ints.overlap(self.basis)
        return self._S.copy()

    def ints_kinetic(self):
        return _ints.kinetic(self.basis)

    def ints_nuclear(self):
        return _ints.nuclear(self.basis, self.atoms)

    def ints_hcore(self):
        if self._H is None:
            self._H = self.ints_kinetic() + self.ints_nuclear()
        return self._H.copy()

    def get_Xorth(self):
        return self.Xorth

    def coulomb_2c(self):
        if self.auxiliary_basis is None:
            raise ValueError("Auxiliary basis is not defined.")
        return _ints.coulomb_2c(self.auxiliary_basis)

    def coulomb_3c(self, shell_begin=0, shell_end=None, out=None):
        if self.auxiliary_basis is None:
            raise ValueError("Auxiliary basis is not defined.")
        return _ints.coulomb_3c(self.auxiliary_basis, self.basis, shell_begin, shell_end, out)

    def sap_potential(self):
        """V_SAP = sum_P (P|mn) over the SAP charge-density basis (scf_utils.py:10-51),
        summed over P inside the engine; computed once per System (it depends
        on the geometry and basis only)."""
        if self._V_sap is None:
            self._V_sap = _ints.coulomb_3c_sum(_basis.build_sap_density_basis(self.atoms), self.basis)
        return self._V_sap.copy()

    # ------------------------------------------------------------------ Fock builder slot
    @property
    def fock_builder(self):
        if self._fock_builder is None:
            self._fock_builder = self._init_fock_builder()
        return self._fock_builder

    def _init_fock_builder(self):
        """In-core builder, or the integral-direct one when B and the metric
        exceed jk_mem_thres_mb (MB of HBM) (forte2/system/system.py:405-434)."""
        from .jkbuilder import FockBuilder

        if self.jk_mem_thres_mb is not None and self.cholesky_tei:
            raise ValueError("If cholesky_tei is True, jk_mem_thres_mb must be None.")
        if self.jk_mem_thres_mb is None:
            return FockBuilder(self, device=self.device)
        if self.auxiliary_basis is None:
            raise ValueError("If jk_mem_thres_mb is set, auxiliary_basis_set must be provided to determine "
                             "the size of the B tensor.")
        b_mb = 8 * self.naux * self.nbf**2 / 1024**2
        metric_mb = 8 * self.naux**2 / 1024**2
        if b_mb + metric_mb > self.jk_mem_thres_mb:
            from .otf import FockBuilderOTF

            return FockBuilderOTF(self, jk_mem_thres_mb=self.jk_mem_thres_mb, backend=self.integral_backend,
                                  device=self.device)
        # the reference's third choice (no transposed copy) is the only layout here
        return FockBuilder(self, store_B_nPm=2 * b_mb + metric_mb <= self.jk_mem_thres_mb, device=self.device)

    @fock_builder.setter
    def fock_builder(self, fb):
        self._fock_builder = fb
