"""Pins of the J/K oracle (oracle/jk.py) that need no GPU.

Each check is chosen so that a plausible slip in the oracle (a transposed
operand, a dropped symmetrisation, the wrong contraction index) fails it:
brute-force 4-index loops written independently of the einsum strings, and
the invariants that the mathematics fixes.
"""

import itertools

import numpy as np
import pytest

from oracle import jk as ojk


def _sym_B(rng, Q, N):
    B = rng.standard_normal((Q, N, N))
    return 0.5 * (B + B.transpose(0, 2, 1))


def _brute_JK(B, D):
    Q, N, _ = B.shape
    eri = np.zeros((N, N, N, N))
    for m, n, r, s in itertools.product(range(N), repeat=4):
        eri[m, n, r, s] = sum(B[P, m, n] * B[P, r, s] for P in range(Q))
    J = np.zeros((N, N))
    K = np.zeros((N, N))
    for m, n in itertools.product(range(N), repeat=2):
        J[m, n] = sum(eri[m, n, r, s] * D[s, r] for r in range(N) for s in range(N))
        # K_mn = sum_rs (m s | r n) D_sr  (jkbuilder.py:219-221 docstring)
        K[m, n] = sum(eri[m, s, r, n] * D[s, r] for r in range(N) for s in range(N))
    return J, K


@pytest.mark.parametrize("N,Q,o", [(4, 3, 2), (5, 7, 3)])
def test_jk_matches_brute_force(N, Q, o):
    rng = np.random.default_rng(12345)
    B = _sym_B(rng, Q, N)
    C = rng.standard_normal((N, o))
    J, K = ojk.build_JK(B, [C])
    Jb, Kb = _brute_JK(B, C @ C.T)
    assert np.abs(J[0] - Jb).max() < 1e-12
    assert np.abs(K[0] - Kb).max() < 1e-12
    # the (P,m,n) variant agrees with the default (n,P,m) one
    assert np.abs(ojk.build_K_Pmn(B, [C])[0] - K[0]).max() < 1e-12


def test_jk_invariants():
    rng = np.random.default_rng(7)
    N, Q = 9, 17
    B = _sym_B(rng, Q, N)
    C1 = rng.standard_normal((N, 3))
    C2 = rng.standard_normal((N, 4))
    (J1, J2), (K1, K2) = ojk.build_JK(B, [C1, C2])
    for M in (J1, J2, K1, K2):
        assert np.abs(M - M.T).max() < 1e-12
    # K(D) is PSD for PSD D: K = sum_P (B_P C)(B_P C)^T
    assert np.linalg.eigvalsh(K1).min() > -1e-10
    # linearity in D: J[D1 + D2] = J[D1] + J[D2]
    D12 = C1 @ C1.T + C2 @ C2.T
    (J12,) = ojk.build_J(B, [D12])
    assert np.abs(J12 - J1 - J2).max() < 1e-11
    Cc = np.hstack([C1, C2])
    _, (K12,) = ojk.build_JK(B, [Cc])
    assert np.abs(K12 - K1 - K2).max() < 1e-11
    # tr(D K) = sum_P ||C^T B_P C||_F^2 >= 0
    assert np.trace(C1 @ C1.T @ K1) >= 0


def test_generalized_matches_direct_density():
    rng = np.random.default_rng(3)
    N, Q, na = 6, 11, 3
    B = _sym_B(rng, Q, N)
    C = rng.standard_normal((N, na))
    A = rng.standard_normal((na, na))
    g1 = A @ A.T + 0.1 * np.eye(na)
    J, K = ojk.build_JK_generalized(B, C, g1)
    D = C @ g1 @ C.T
    Jb, Kb = _brute_JK(B, D)
    assert np.abs(J - Jb).max() < 1e-11
    assert np.abs(K - Kb).max() < 1e-11
    # singular g1 goes through the eigen branch
    g1s = np.zeros((na, na))
    g1s[0, 0] = 2.0
    J, K = ojk.build_JK_generalized(B, C, g1s)
    Jb, Kb = _brute_JK(B, C @ g1s @ C.T)
    assert np.abs(J - Jb).max() < 1e-11 and np.abs(K - Kb).max() < 1e-11


def test_B_build_reproduces_metric_contraction():
    # B = L^-1 (Q|mn): sum_P B_Pmn B_Prs = (mn|Q) M^-1 (Q|rs)
    rng = np.random.default_rng(11)
    Q, N = 8, 5
    A = rng.standard_normal((Q, Q))
    M = A @ A.T + Q * np.eye(Q)
    T = _sym_B(rng, Q, N)
    B = ojk.build_B_density_fitting(M, T)
    lhs = np.einsum("Pmn,Prs->mnrs", B, B)
    rhs = np.einsum("Qmn,QR,Rrs->mnrs", T, np.linalg.inv(M), T)
    assert np.abs(lhs - rhs).max() < 1e-12
    with pytest.raises(ValueError):
        ojk.build_B_density_fitting(-M, T)


def test_mo_blocks_match_brute_force():
    # B need not be symmetric here: a swapped m/n or r/s index would show
    rng = np.random.default_rng(7)
    Q, N = 3, 4
    B = rng.standard_normal((Q, N, N))
    C1, C2, C3, C4 = (rng.standard_normal((N, k)) for k in (2, 3, 1, 4))
    Bmo = ojk.B_tensor_gen_block(B, C1, C3)
    for P, i, k in itertools.product(range(Q), range(2), range(1)):
        ref = sum(B[P, m, n] * C1[m, i] * C3[n, k] for m in range(N) for n in range(N))
        assert abs(Bmo[P, i, k] - ref) < 1e-12
    V = ojk.two_electron_integrals_gen_block(B, C1, C2, C3, C4)
    assert V.shape == (2, 3, 1, 4)
    for i, j, k, l in itertools.product(range(2), range(3), range(1), range(4)):
        ref = sum(B[P, m, n] * B[P, r, s] * C1[m, i] * C2[r, j] * C3[n, k] * C4[s, l]
                  for P in range(Q) for m, n, r, s in itertools.product(range(N), repeat=4))
        assert abs(V[i, j, k, l] - ref) < 1e-11


def test_mo_blocks_physicist_convention():
    # <pq|rs> = (pr|qs): with C = 1, V is the AO chemist ERI with axes 1 and 2 swapped
    rng = np.random.default_rng(8)
    B = _sym_B(rng, 5, 4)
    eye = np.eye(4)
    V = ojk.two_electron_integrals_gen_block(B, eye, eye, eye, eye)
    assert np.abs(V - ojk.eri_from_B(B).transpose(0, 2, 1, 3)).max() < 1e-13
    assert np.abs(ojk.B_tensor_gen_block(B, eye, eye) - B).max() == 0.0


def test_blas_forms_equal_the_reference_einsum_strings():
    """build_J and build_K_Pmn are written as tensordot/GEMM contractions; on a
    non-symmetric B and D (so that a swapped index shows) they equal the
    reference's einsum strings (jkbuilder.py:161-166, 174-181)."""
    rng = np.random.default_rng(21)
    Q, N, o = 7, 9, 4
    B = rng.standard_normal((Q, N, N))
    D = rng.standard_normal((N, N))
    C = rng.standard_normal((N, o))
    J = ojk.build_J(B, [D])[0]
    assert np.abs(J - np.einsum("Pmn,Prs,sr->mn", B, B, D)).max() < 1e-12
    # the transposed D gives a different J on this B: the sr order matters
    assert np.abs(J - ojk.build_J(B, [D.T])[0]).max() > 1e-3
    Y = np.einsum("Pms,si->Pmi", B, C)
    K = ojk.build_K_Pmn(B, [C])[0]
    assert np.abs(K - np.einsum("Pmi,Pni->mn", Y, Y)).max() < 1e-12
    assert ojk.build_K_Pmn(B, [C[:, :0]])[0].shape == (N, N)
