"""numpy FP64 restatement of forte2's FockBuilder algebra (test infrastructure).

Pinned by tests/test_oracle_jk.py: every function is checked against an
explicit 4-index brute force on small random inputs ((mn|rs) = sum_P B_Pmn
B_Prs contracted by plain Python loops), against linearity / symmetry /
positivity invariants, and -- through the SCF oracle -- against the reference's
golden energies (tests/test_oracle_scf.py).
"""

import numpy as np
import scipy.linalg


def build_B_density_fitting(M, Pmn):
    """B = L^-1 (Q|mn) with M = L L^T (forte2/jkbuilder/jkbuilder.py:135-159).

    Raises ValueError when the metric is not positive definite (:141-148).
    """
    try:
        L = scipy.linalg.cholesky(M, lower=True)
    except scipy.linalg.LinAlgError:
        raise ValueError("Density fitting Coulomb metric (P|Q) is not positive definite.")
    naux = M.shape[0]
    X = scipy.linalg.solve_triangular(L, np.eye(naux), lower=True)
    return np.einsum("PQ,Qmn->Pmn", X, Pmn, optimize=True)


def build_J(B, D):
    """[sum_P B_Pmn sum_rs B_Prs D_sr for D_k in D] (jkbuilder.py:161-166): the
    reference's one einsum as its two contractions, gamma_P = sum_rs B_Prs
    D_sr and then sum_P gamma_P B_Pmn (tensordot reaches BLAS; the einsum
    path takes ~30x longer at nbf 1536)."""
    out = []
    for Di in D:
        gamma = np.tensordot(B, Di, axes=([1, 2], [1, 0]))
        out.append(np.tensordot(gamma, B, axes=(0, 0)))
    return out


def build_K_Pmn(B, C):
    """K via the (P,m,n) layout (jkbuilder.py:168-182, real case): the same two
    contractions, Y[P,m,i] = sum_s B[P,m,s] C[s,i] and K[m,n] = sum_{P,i}
    Y[P,m,i] Y[P,n,i], written as GEMMs over the flattened (P,m) and (P,i)
    indices (numpy's einsum does not reach BLAS for the second one)."""
    Q, N = B.shape[0], B.shape[1]
    out = []
    for Ci in C:
        Y = (B.reshape(Q * N, B.shape[2]) @ Ci).reshape(Q, N, Ci.shape[1])
        Yt = np.ascontiguousarray(Y.transpose(1, 0, 2)).reshape(N, Q * Ci.shape[1])
        out.append(Yt @ Yt.T)
    return out


def build_K_nPm(B, C, BnPm=None):
    """K via the transposed (n,P,m) copy, the reference default (jkbuilder.py:74-80, 184-201)."""
    if BnPm is None:
        BnPm = np.ascontiguousarray(B.transpose((2, 0, 1)))
    Y = [BnPm @ Ci for Ci in C]
    return [np.tensordot(Yi, Yi, axes=([1, 2], [1, 2])) for Yi in Y]


def build_JK(B, C, BnPm=None):
    """(J, K) lists for C_k with D_k = C_k C_k^T (jkbuilder.py:209-251, real case)."""
    D = [np.einsum("mi,ni->mn", Ci, Ci, optimize=True) for Ci in C]
    return build_J(B, D), build_K_nPm(B, C, BnPm)


def build_JK_generalized(B, C, g1):
    """J/K of the density C g1 C^T via the Cholesky (or eigen) factor of g1 (jkbuilder.py:253-290)."""
    assert C.shape[1] == g1.shape[0]
    try:
        L = np.linalg.cholesky(g1)
        Cp = C @ L
    except np.linalg.LinAlgError:
        n, L = np.linalg.eigh(g1)
        assert np.all(n > -1.0e-11), "g1 must be positive semi-definite"
        n = np.maximum(n, 0)
        Cp = C @ L @ np.diag(np.sqrt(n))
    J, K = build_JK(B, [Cp])
    return J[0], K[0]


def build_core_fock(B, C_core, hcore):
    """h + 2 J - K of the core orbitals (jkbuilder.py:292-321)."""
    J, K = build_JK(B, [C_core])
    return hcore + 2.0 * J[0] - K[0]


def build_active_fock(B, C_act, g1):
    """J - K/2 of the active density (jkbuilder.py:323-349)."""
    J, K = build_JK_generalized(B, C_act, g1)
    return J - 0.5 * K


def eri_from_B(B):
    """(mn|rs) = sum_P B_Pmn B_Prs -- brute-force helper for small N."""
    return np.einsum("Pmn,Prs->mnrs", B, B)


def B_tensor_gen_block(B, C1, C2):
    """MO-basis B: B_Pij = sum_mn B_Pmn C1_mi C2_nj (jkbuilder.py:521-552,
    real orbitals)."""
    return np.einsum("Pmn,mi,nj->Pij", B, C1, C2, optimize=True)


def two_electron_integrals_gen_block(B, C1, C2, C3, C4):
    """<ij|kl> = sum_P B_P(C1, C3)_ik B_P(C2, C4)_jl (jkbuilder.py:382-420,
    the einsum "Pmn,Prs,mi,rj,nk,sl->ijkl")."""
    return np.einsum("Pmn,Prs,mi,rj,nk,sl->ijkl", B, B, C1, C2, C3, C4, optimize=True)
