"""numpy FP64 restatement of forte2's RHF/UHF SCF loop (test infrastructure).

Follows forte2/scf/scf_base.py:62-251 (loop, convergence test, final extra
iteration), forte2/scf/rhf.py:35-65 and forte2/scf/uhf.py:61-127 (Fock,
density, AO gradient, energy), forte2/helpers/diis.py:10-125 (Pulay CDIIS
with the SCF settings diis_start=1, nvec=8, min=2), forte2/helpers/
matrix_functions.py:9-152 (canonical orthogonalisation) and
forte2/scf/scf_utils.py:10-73 (SAP / core guesses).

Inputs are plain arrays: overlap S, core Hamiltonian H, nuclear repulsion
Vnn, the DF tensor B (any object with build_JK(list_of_C) -> (J, K)), and,
for the SAP guess, the SAP potential matrix V_SAP = sum_P (P|mn).

Pinned by tests/test_oracle_scf.py against the reference's golden energies
(water cc-pVDZ/cc-pVTZ-JKFIT RHF, H2 and HF cc-pVDZ/cc-pVTZ-JKFIT RHF, water
cc-pVQZ/cc-pVQZ-JKFIT RHF and UHF triplet with <S^2>) at the reference's
tolerance pytest.approx(abs=5e-8) (forte2/helpers/comparisons.py:5).
"""

from collections import deque

import numpy as np

from . import jk as ojk

MACHEPS = 1e-14


def canonical_orth(S, rtol=1e-8):
    """X = U_eta s_eta^{-1/2} over eigenvalues >= rtol * max (matrix_functions.py:9-31, 95-152)."""
    sevals, sevecs = np.linalg.eigh(S)
    if np.any(sevals < -MACHEPS):
        raise ValueError("The metric matrix must be positive semi-definite.")
    sevals[sevals < 0] = 0.0
    ndiscard = np.searchsorted(sevals, rtol * sevals[-1])
    U = sevecs[:, ndiscard:]
    return U / np.sqrt(sevals[ndiscard:])


class DIIS:
    """Pulay CDIIS (diis.py:10-125)."""

    def __init__(self, diis_start=1, diis_nvec=8, diis_min=2, do_diis=True):
        self.diis_start, self.diis_nvec, self.diis_min = diis_start, diis_nvec, diis_min
        self.do_diis = do_diis and diis_start >= 0 and diis_nvec >= 1
        self.p = deque(maxlen=diis_nvec)
        self.e = deque(maxlen=diis_nvec)
        self.iter = 0
        self.status = ""

    def update(self, p, e):
        self.status = ""
        if not self.do_diis:
            return p
        self.iter += 1
        if self.iter < self.diis_start:
            return p
        self.p.append(p.copy())
        self.e.append(e.copy())
        self.status = "S"
        n = len(self.p)
        if n < self.diis_min:
            return p
        Bm = -np.ones((n + 1, n + 1))
        Bm[-1, -1] = 0.0
        rhs = np.zeros(n + 1)
        rhs[-1] = -1.0
        for i in range(n):
            for j in range(i, n):
                Bm[i, j] = np.dot(self.e[i].ravel(), self.e[j].ravel())
                Bm[j, i] = Bm[i, j]
        Bm[:-1, :-1] /= np.abs(Bm[:-1, :-1]).max()
        try:
            x = np.linalg.solve(Bm, rhs)
        except np.linalg.LinAlgError:
            return p
        self.status += "/E"
        out = np.zeros_like(p)
        for i in range(n):
            out += x[i] * self.p[i]
        return out


def _eigh(X, F):
    e, c = np.linalg.eigh(X.T @ F @ X)
    return e, X @ c


def guess(X, H, V_sap=None):
    """SAP (V_sap given) or core-Hamiltonian guess (scf_utils.py:10-73)."""
    Hg = H if V_sap is None else H + V_sap
    return _eigh(X, Hg)[1]


class SCFResult:
    pass


def run_rhf(S, H, Vnn, jk, nel, X=None, V_sap=None, e_tol=1e-9, d_tol=1e-6, maxiter=100,
            diis=(1, 8, 2), do_diis=True, C0=None, trace=None, level_shift=None, level_shift_thresh=1e-5):
    """RHF (scf_base.py:121-251 with rhf.py:35-65): returns E, C, eps, D, iterations."""
    if nel % 2:
        raise AssertionError("RHF requires an even number of electrons.")
    na = nel // 2
    X = canonical_orth(S) if X is None else X
    dd = DIIS(*diis, do_diis=do_diis)
    C = guess(X, H, V_sap) if C0 is None else C0

    def fock(C):
        J, K = jk.build_JK([C[:, :na]])
        return H + 2.0 * J[0] - K[0]

    D = C[:, :na] @ C[:, :na].T
    F = fock(C)
    E = Vnn + np.sum(D * (H + F))
    Eold, Dold = E, D
    it = 1
    converged = False
    shift = level_shift  # scf_base.py:131; switched off below at :203-204
    for _ in range(maxiter):
        grad = X.T @ (F @ D @ S - S @ D @ F) @ X
        Fx = dd.update(F, grad)
        if shift is not None and shift >= 1e-4:  # rhf.py:90-95
            Fx = Fx + shift * (S - S @ D @ S)
        eps, C = _eigh(X, Fx)
        D = C[:, :na] @ C[:, :na].T
        F = fock(C)
        E = Vnn + np.sum(D * (H + F))
        dE = E - Eold
        if abs(dE) < level_shift_thresh:
            shift = None
        dD = np.linalg.norm(D - Dold)
        if trace is not None:
            trace.append((E, dE, dD))
        if abs(dE) < e_tol and dD < d_tol:
            # the final iteration diagonalises the new, non-extrapolated Fock
            # matrix (scf_base.py:214-221: F_canon was rebuilt at :196)
            eps, C = _eigh(X, F)
            D = C[:, :na] @ C[:, :na].T
            F = fock(C)
            E = Vnn + np.sum(D * (H + F))
            converged = True
            break
        Eold, Dold = E, D
        it += 1
    if not converged:
        raise RuntimeError(f"RHF did not converge in {maxiter} iterations.")
    r = SCFResult()
    r.E, r.C, r.eps, r.D, r.F, r.iter, r.na = E, C, eps, D, F, it, na
    return r


def run_uhf(S, H, Vnn, jk, nel, ms, X=None, V_sap=None, e_tol=1e-9, d_tol=1e-6, maxiter=100,
            diis=(1, 8, 2), do_diis=True, level_shift=None, level_shift_thresh=1e-5, guess_mix=False):
    """UHF (scf_base.py:121-251 with uhf.py:40-127)."""
    twicems = int(round(2 * ms))
    if nel % 2 != twicems % 2:
        raise ValueError(f"{nel} electrons is incompatible with ms={ms}!")
    na, nb = (nel + twicems) // 2, (nel - twicems) // 2
    X = canonical_orth(S) if X is None else X
    N = S.shape[0]
    dd = DIIS(*diis, do_diis=do_diis)
    C0 = guess(X, H, V_sap)
    C = [C0, C0]
    if twicems == 0 and guess_mix:
        # uhf.py:85-86 and scf_utils.py:76-104: HOMO/LUMO Givens rotation by
        # +pi/4 (alpha) and -pi/4 (beta)
        h = nel // 2 - 1
        q = np.pi / 4
        C = []
        for sq in (np.sin(q), -np.sin(q)):
            M = C0.copy()
            M[:, h] = np.cos(q) * C0[:, h] + sq * C0[:, h + 1]
            M[:, h + 1] = -sq * C0[:, h] + np.cos(q) * C0[:, h + 1]
            C.append(M)

    def dens(C):
        return [C[0][:, :na] @ C[0][:, :na].T, C[1][:, :nb] @ C[1][:, :nb].T]

    def fock(C):
        (Ja, Jb), K = jk.build_JK([C[0][:, :na], C[1][:, :nb]])
        return [H + Ja + Jb - k for k in K]

    def energy(D, F):
        return 0.5 * (np.sum((D[0] + D[1]) * H) + np.sum(D[0] * F[0]) + np.sum(D[1] * F[1]))

    def s2(C):
        Sij = C[0][:, :na].T @ S @ C[1][:, :nb]
        return ms * (ms + 1) + nb - np.sum(Sij * Sij)

    D = dens(C)
    F = fock(C)
    E = Vnn + energy(D, F)
    Eold, Dold = E, D
    it = 1
    converged = False
    shift = None if level_shift is None else (
        tuple(level_shift) if isinstance(level_shift, tuple) else (level_shift, level_shift))
    for _ in range(maxiter):
        grad = np.hstack([(X.T @ (f @ d @ S - S @ d @ f) @ X).ravel() for d, f in zip(D, F)])
        flat = dd.update(np.hstack([f.ravel() for f in F]), grad)
        Fx = [flat[: N * N].reshape(N, N), flat[N * N:].reshape(N, N)]
        if shift is not None and not all(x < 1e-4 for x in shift):  # uhf.py:150-159
            Fx = [f + x * (S - S @ d @ S) for x, f, d in zip(shift, Fx, D)]
        ea, Ca = _eigh(X, Fx[0])
        eb, Cb = _eigh(X, Fx[1])
        C = [Ca, Cb]
        D = dens(C)
        F = fock(C)
        E = Vnn + energy(D, F)
        dE = E - Eold
        if abs(dE) < level_shift_thresh:
            shift = None
        dD = sum(np.linalg.norm(d - do) for d, do in zip(D, Dold))
        if abs(dE) < e_tol and dD < d_tol:
            ea, Ca = _eigh(X, F[0])  # the new Fock matrices, as in RHF above
            eb, Cb = _eigh(X, F[1])
            C = [Ca, Cb]
            D = dens(C)
            F = fock(C)
            E = Vnn + energy(D, F)
            converged = True
            break
        Eold, Dold = E, D
        it += 1
    if not converged:
        raise RuntimeError(f"UHF did not converge in {maxiter} iterations.")
    r = SCFResult()
    r.E, r.C, r.eps, r.D, r.F, r.iter, r.S2 = E, C, [ea, eb], D, F, it, s2(C)
    return r


class DFJK:
    """build_JK over a host B tensor (oracle/jk.py), with the transposed copy cached."""

    def __init__(self, B):
        self.B = B
        self.BnPm = np.ascontiguousarray(B.transpose((2, 0, 1)))

    def build_JK(self, C):
        return ojk.build_JK(self.B, C, self.BnPm)
