"""Density-fitted B(P|mn) built on the device.

forte2 forms B on the host (forte2/jkbuilder/jkbuilder.py:119-159):
M = (P|Q), L = chol(M), X = L^-1, B = X (Q|mn) with an einsum.  Here the
3-centre integrals are evaluated by the host engine one aux-shell block at a
time and copied into the B tensor in HBM, then B = L^-1 (Q|mn) is applied in
place on the device by column blocks (triangular solve), so the host never
holds more than one block of (Q|mn).

With ``df_ortho_rtol`` set, X = M^-1/2 with eigenvalues below rtol*max
discarded (forte2/helpers/matrix_functions.py:52-92), as in the reference.
"""

import torch

from . import integrals as _ints


def metric_factor(M, df_ortho_rtol=None):
    """(kind, factor) with B = factor^-1 T ("chol", L) or B = factor T ("invsqrt", X)."""
    if df_ortho_rtol is None:
        L, info = torch.linalg.cholesky_ex(M)
        if int(info) != 0:
            raise ValueError(
                "Density fitting Coulomb metric (P|Q) is not positive definite.\n"
                "Please set df_ortho_rtol to a small positive value to orthogonalize the metric.")
        return "chol", L
    w, U = torch.linalg.eigh(M)
    if bool((w < -1e-14).any()):
        raise ValueError("The metric matrix must be positive semi-definite.")
    w = torch.clamp(w, min=0.0)
    keep = w >= df_ortho_rtol * w[-1]
    Uk, wk = U[:, keep], w[keep]
    return "invsqrt", (Uk / torch.sqrt(wk)) @ Uk.T


def _fitting_matrix(kind, fac, blk=1024):
    """X with B = X T: L^-1 (as the reference, jkbuilder.py:150), formed in
    column blocks so that no triangular solve has a huge right-hand side."""
    if kind != "chol":
        return fac
    Q = fac.shape[0]
    X = torch.empty_like(fac)
    for j0 in range(0, Q, blk):
        j1 = min(Q, j0 + blk)
        E = torch.zeros((Q, j1 - j0), dtype=fac.dtype, device=fac.device)
        E[j0:j1] = torch.eye(j1 - j0, dtype=fac.dtype, device=fac.device)
        X[:, j0:j1] = torch.linalg.solve_triangular(fac, E, upper=False)
    return X


def build_B_density_fitting(system, device, block_bytes=1 << 30, solve_bytes=1 << 30, progress=None):
    """B (naux, nbf, nbf) FP64 on `device`; `progress(msg)` is called per host block."""
    aux, bas = system.auxiliary_basis, system.basis
    if aux is None:
        raise ValueError("Auxiliary basis is not defined. Define auxiliary_basis_set.")
    Q, N = aux.size, bas.size
    M = torch.from_numpy(_ints.coulomb_2c(aux)).to(device)
    kind, fac = metric_factor(M, getattr(system, "df_ortho_rtol", None))
    B = torch.empty((Q, N, N), dtype=torch.float64, device=device)
    # stream (Q|mn) into B by aux-shell blocks
    offs = aux.shell_offsets()
    per_fn = N * N * 8
    s0 = 0
    host = None
    while s0 < aux.nshells:
        s1 = s0 + 1
        while s1 < aux.nshells and (offs[s1 + 1] - offs[s0]) * per_fn <= block_bytes:
            s1 += 1
        nP = offs[s1] - offs[s0]
        if host is None or host.shape[0] < nP:
            host = torch.empty((nP, N, N), dtype=torch.float64, pin_memory=True)
        blk = host[:nP]
        _ints.coulomb_3c(aux, bas, s0, s1, out=blk.numpy())
        B[offs[s0]:offs[s1]].copy_(blk, non_blocking=False)
        if progress is not None:
            progress(f"(Q|mn) aux functions {offs[s1]}/{Q} on device")
        s0 = s1
    X = _fitting_matrix(kind, fac)
    # B = X T in place, one contiguous column chunk of the (Q, N*N) view at a time
    Bv = B.view(Q, N * N)
    chunk = max(1, min(N * N, solve_bytes // (8 * Q)))
    for c0 in range(0, N * N, chunk):
        c1 = min(N * N, c0 + chunk)
        blk = Bv[:, c0:c1].contiguous()
        Bv[:, c0:c1] = X @ blk
        del blk
    return B



def host_df_integrals(system):
    """(M, T) on the host: the Coulomb metric and all of (Q|mn) (small jobs)."""
    return (_ints.coulomb_2c(system.auxiliary_basis),
            _ints.coulomb_3c(system.auxiliary_basis, system.basis))


def device_B_from_host(M, T, device, df_ortho_rtol=None, solve_bytes=1 << 30):
    """B = L^-1 T on the device from host (M, T) (same algebra as build_B_density_fitting)."""
    Q, N = T.shape[0], T.shape[1]
    kind, fac = metric_factor(torch.from_numpy(M).to(device), df_ortho_rtol)
    B = torch.from_numpy(T).to(device)
    X = _fitting_matrix(kind, fac)
    Bv = B.view(Q, N * N)
    chunk = max(1, min(N * N, solve_bytes // (8 * Q)))
    for c0 in range(0, N * N, chunk):
        c1 = min(N * N, c0 + chunk)
        Bv[:, c0:c1] = X @ Bv[:, c0:c1].contiguous()
    return B
