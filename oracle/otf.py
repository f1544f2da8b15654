"""numpy FP64 restatement of forte2's integral-direct FockBuilderOTF (test
infrastructure; forte2/jkbuilder/jkbuilder.py:598-1265, real orbitals).

The raw 3-centre integrals T = (P|mn) are passed whole (small systems) and
read block by block, as the reference reads its (P|mn) buffer: ``blocks`` is
the list of (p0, p1) aux-function ranges, ``iblk`` the occupied block size.

Pinned by tests/test_oracle_otf.py: J, K, the combined JK and the MO-basis B
equal the in-core oracle (oracle/jk.py, B = L^-1 T) for the Cholesky and the
truncated-eigen metric, over several block and column-block sizes; a
different algebra path (M^-1 by a Cholesky solve, M^-1/2 applied after the
contraction), so a dropped or transposed metric factor fails.
"""

import numpy as np
import scipy.linalg


class Metric:
    """M^-1/2 and M^-1 y of the Coulomb metric (jkbuilder.py:742-782)."""

    def __init__(self, M, df_ortho_rtol=None):
        self.rtol = df_ortho_rtol
        if df_ortho_rtol is None:
            try:
                self.L = scipy.linalg.cholesky(M, lower=True)
            except scipy.linalg.LinAlgError:
                raise ValueError("Density fitting Coulomb metric (P|Q) is not positive definite.")
            self.Mm12 = scipy.linalg.solve_triangular(self.L, np.eye(M.shape[0]), lower=True)
        else:
            w, U = np.linalg.eigh(M)
            w[w < 0] = 0.0
            nd = int(np.searchsorted(w, df_ortho_rtol * w[-1]))
            self.sevals, self.sevecs = w[nd:], U[:, nd:]
            self.Mm12 = (self.sevecs / np.sqrt(self.sevals)) @ self.sevecs.T

    def apply_Mm1(self, y):
        if self.rtol is None:
            return scipy.linalg.cho_solve((self.L, True), y)
        return self.sevecs @ ((self.sevecs.T @ y) / self.sevals)


def J_kernel(met, T, blocks, D):
    """Two passes over the blocks (jkbuilder.py:813-840)."""
    bP = np.zeros(T.shape[0])
    for p0, p1 in blocks:
        bP[p0:p1] = np.einsum("Pmn,nm->P", T[p0:p1], D, optimize=True)
    bP = met.apply_Mm1(bP)
    J = np.zeros_like(D)
    for p0, p1 in blocks:
        J += np.einsum("Pmn,P->mn", T[p0:p1], bP[p0:p1], optimize=True)
    return J


def K_kernel(met, T, blocks, C, iblk):
    """Occupied blocks of iblk columns; per block X = (Q|mn) C over all aux
    blocks, Y = M^-1/2 X, K += sum_Pi Y_Pmi Y_Pni (jkbuilder.py:842-875)."""
    Q, N, _ = T.shape
    K = np.zeros((N, N))
    for i0 in range(0, C.shape[1], iblk):
        c = C[:, i0:i0 + iblk]
        X = np.zeros((Q, N, c.shape[1]))
        for p0, p1 in blocks:
            X[p0:p1] = np.einsum("Qms,si->Qmi", T[p0:p1], c, optimize=True)
        Y = np.einsum("PQ,Qmi->Pmi", met.Mm12, X, optimize=True)
        K += np.einsum("Pmi,Pni->mn", Y, Y, optimize=True)
    return K


def JK_kernel(met, T, blocks, C, iblk):
    """The combined build (jkbuilder.py:945-1019): gamma in the first
    occupied-block sweep, J in the second (or in one more sweep)."""
    Q, N, _ = T.shape
    D = C @ C.T
    bP = np.zeros(Q)
    J = np.zeros((N, N))
    K = np.zeros((N, N))
    jpass = 0
    for i0 in range(0, C.shape[1], iblk):
        c = C[:, i0:i0 + iblk]
        X = np.zeros((Q, N, c.shape[1]))
        for p0, p1 in blocks:
            X[p0:p1] = np.einsum("Qms,si->Qmi", T[p0:p1], c, optimize=True)
            if jpass == 0:
                bP[p0:p1] = np.einsum("Pmn,nm->P", T[p0:p1], D, optimize=True)
            elif jpass == 1:
                J += np.einsum("Pmn,P->mn", T[p0:p1], bP[p0:p1], optimize=True)
        Y = np.einsum("PQ,Qmi->Pmi", met.Mm12, X, optimize=True)
        K += np.einsum("Pmi,Pni->mn", Y, Y, optimize=True)
        if jpass == 0:
            bP = met.apply_Mm1(bP)
        jpass += 1
    if jpass == 1:
        for p0, p1 in blocks:
            J += np.einsum("Pmn,P->mn", T[p0:p1], bP[p0:p1], optimize=True)
    return J, K


def B_tensor_gen_block(met, T, blocks, C1, C2):
    """B = M^-1/2 [(Q|mn) C1 C2 by blocks] (jkbuilder.py:1207-1226)."""
    B = np.zeros((T.shape[0], C1.shape[1], C2.shape[1]))
    for p0, p1 in blocks:
        B[p0:p1] = np.einsum("Pmn,mi,nj->Pij", T[p0:p1], C1, C2, optimize=True)
    return np.einsum("PQ,Qmn->Pmn", met.Mm12, B, optimize=True)
