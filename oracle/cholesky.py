"""numpy/LAPACK restatement of forte2's Cholesky-TEI B source (test infrastructure).

forte2 can take B from a pivoted Cholesky decomposition of the 4-index
electron-repulsion matrix V[(mn),(ls)] = (mn|ls) instead of density fitting
(System(cholesky_tei=True, cholesky_tol=...)).  The oracle does what the
reference does, with the same LAPACK routine (dpstrf through scipy).

Pinned by tests/test_oracle_cholesky.py: the reference's own checks
(tests/helpers/test_cholesky.py: B^T B = V for water cc-pVDZ at tol 1e-14, a
random positive semi-definite matrix with a zero row) and its golden RHF
energy (tests/scf/test_rhf.py:47-59, -76.021769351262 at tol 1e-10).
"""

import numpy as np
import scipy.linalg


def cholesky_wrapper(M, tol):
    """Upper Cholesky factor with complete pivoting, rows in the original
    column order, so that B^T B = M to within tol (dpstrf stops when the
    largest remaining diagonal element is <= tol; tol < 0 means LAPACK's own
    n * eps * max(diag)).  forte2/helpers/matrix_functions.py:239-270."""
    C, piv, rank, info = scipy.linalg.lapack.dpstrf(M, tol=tol, lower=False)
    if info < 0:
        raise ValueError(f"dpstrf failed with info={info}")
    piv = piv - 1
    inv = np.empty_like(piv)
    inv[piv] = np.arange(len(piv))
    return np.triu(C)[:rank, inv]


def build_B_cholesky(eri, tol):
    """B[P][m][n] from the (nbf, nbf, nbf, nbf) ERI tensor
    (forte2/jkbuilder/jkbuilder.py:90-117); naux = B.shape[0] is the rank."""
    nbf = eri.shape[0]
    B = cholesky_wrapper(np.ascontiguousarray(eri).reshape(nbf * nbf, nbf * nbf), tol)
    return B.reshape(B.shape[0], nbf, nbf)
