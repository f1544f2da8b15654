"""Density-fitted B(P|mn) built on the device.

forte2 forms B on the host (forte2/jkbuilder/jkbuilder.py:119-159):
M = (P|Q), L = chol(M), X = L^-1, B = X (Q|mn) with an einsum.  Here the
3-centre integrals are evaluated by the host engine one block of basis-shell
pairs (m >= n) at a time, for all Q, and each block is multiplied by X on
the device as soon as it arrives (host evaluation, copy and GEMM overlap);
the host never holds more than two blocks of (Q|mn).

With ``df_ortho_rtol`` set, X = M^-1/2 with eigenvalues below rtol*max
discarded (forte2/helpers/matrix_functions.py:52-92), as in the reference.
"""

import numpy as np
import torch

from . import integrals as _ints


def metric_factor(M, df_ortho_rtol=None):
    """(kind, factor) with B = factor^-1 T ("chol", L) or B = factor T ("invsqrt", X)."""
    if df_ortho_rtol is None:
        L, info = torch.linalg.cholesky_ex(M)
        # A pivot at the rounding level of the largest diagonal element counts
        # as zero: the LAPACK test (pivot <= 0) leaves an exactly singular
        # metric, such as two identical fitting functions, to the rounding
        # order of the factorisation, which differs between libraries.
        tiny = torch.finfo(M.dtype).eps * M.diagonal().max()
        if int(info) != 0 or bool((torch.diagonal(L) ** 2 <= tiny).any()):
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


def pair_columns(basis):
    """Column layout of the packed unique half of (P|mn) written by
    f2ints_coulomb_3c_pairs: the significant shell pairs (a >= b) in order, a
    pair a > b with nsa * nsb columns (row-major), a pair a == b with its
    components sa >= sb (row-major).  Returns (col0, m_of_col, n_of_col):
    col0[q] is the first column of pair q (length npairs + 1), and m >= n are
    the basis functions of every column."""
    import numpy as np

    a, b = _ints.ket_pairs(basis)
    offs = np.asarray(basis.shell_offsets(), dtype=np.int64)
    size = offs[1:] - offs[:-1]
    cols_m, cols_n, counts = [], [], []
    # the component pattern of each (la, lb, diagonal) kind, built once
    pattern = {}
    for A, Bs in zip(a, b):
        key = (int(size[A]), int(size[Bs]), bool(A == Bs))
        if key not in pattern:
            na, nb, diag = key
            sa, sb = np.meshgrid(np.arange(na), np.arange(nb), indexing="ij")
            keep = (sa >= sb) if diag else np.ones_like(sa, dtype=bool)
            pattern[key] = (sa[keep], sb[keep])
        sa, sb = pattern[key]
        cols_m.append(offs[A] + sa)
        cols_n.append(offs[Bs] + sb)
        counts.append(len(sa))
    col0 = np.zeros(len(a) + 1, dtype=np.int64)
    np.cumsum(counts, out=col0[1:])
    return col0, np.concatenate(cols_m), np.concatenate(cols_n)


def build_B_density_fitting(system, device, block_bytes=1 << 30, progress=None, stats=None):
    """B (naux, nbf, nbf) FP64 on `device`, B = X (Q|mn) with X = L^-1 (or
    M^-1/2), streamed by blocks of basis-shell pairs.

    The host engine evaluates only the unique half (m >= n shell pairs) of
    (Q|mn), one block of pairs at a time for all Q, on a worker thread into
    one of two pinned staging buffers.  Meanwhile the device copies the
    previous block in, multiplies it by X (a (Q, Q) x (Q, ncols) GEMM, half
    the flops of the full square) and scatters the columns to (m, n) and
    (n, m) of B.  The first block is evaluated while the metric is
    factored.  `stats`, if a dict, receives the total seconds, the seconds
    the host engine spent on its worker thread (overlapped with the rest),
    the number of blocks and of packed columns, and the times to the metric,
    the fitting matrix and the end of the pipeline."""
    import concurrent.futures
    import time

    import numpy as np

    aux, bas = system.auxiliary_basis, system.basis
    if aux is None:
        raise ValueError("Auxiliary basis is not defined. Define auxiliary_basis_set.")
    Q, N = aux.size, bas.size
    t_start = time.perf_counter()
    col0, mcol, ncol_ = pair_columns(bas)
    npairs, ncols = len(col0) - 1, int(col0[-1])
    # pair blocks of at most block_bytes of packed columns
    width = max(1, block_bytes // (8 * Q))
    blocks, q = [], 0
    while q < npairs:
        q1 = q + 1
        while q1 < npairs and col0[q1 + 1] - col0[q] <= width:
            q1 += 1
        blocks.append((q, q1, int(col0[q]), int(col0[q1])))
        q = q1
    pin = torch.device(device).type == "cuda"
    wmax = max((c1 - c0 for _, _, c0, c1 in blocks), default=1)
    host = [torch.empty(Q * wmax, dtype=torch.float64, pin_memory=pin) for _ in range(2)]
    host_s = [0.0]

    def packed(flat, w):
        # contiguous (w, Q) block of a flat buffer: one row of naux values per
        # packed column, the engine's column-major output
        return flat[: Q * w].view(w, Q)

    def evaluate(k, slot):
        q0, q1, c0, c1 = blocks[k]
        t0 = time.perf_counter()
        _ints.coulomb_3c_pairs_t(aux, bas, q0, q1, packed(host[slot], c1 - c0).numpy())
        host_s[0] += time.perf_counter() - t0

    aux.handle, bas.handle  # engine handles made here, before two threads use them
    with concurrent.futures.ThreadPoolExecutor(max_workers=1) as pool:
        # the first pair block is evaluated while the metric is factored
        fut = pool.submit(evaluate, 0, 0) if blocks else None
        M = torch.from_numpy(_ints.coulomb_2c(aux)).to(device)
        t_m2c = time.perf_counter()
        kind, fac = metric_factor(M, getattr(system, "df_ortho_rtol", None))
        X = _fitting_matrix(kind, fac)
        del M, fac
        t_fit = time.perf_counter()
        B = torch.zeros((Q, N, N), dtype=torch.float64, device=device)
        Bv = B.view(Q, N * N)
        idx_mn = torch.from_numpy(mcol * N + ncol_).to(device)
        idx_nm = torch.from_numpy(ncol_ * N + mcol).to(device)
        dev_in = torch.empty(Q * wmax, dtype=torch.float64, device=device)
        stream = torch.cuda.current_stream(device) if pin else None
        copied = [None, None]
        for k, (q0, q1, c0, c1) in enumerate(blocks):
            slot = k % 2
            fut.result()
            if k + 1 < len(blocks):
                nxt = (k + 1) % 2
                if copied[nxt] is not None:
                    copied[nxt].synchronize()
                fut = pool.submit(evaluate, k + 1, nxt)
            Tt = packed(dev_in, c1 - c0)
            Tt.copy_(packed(host[slot], c1 - c0), non_blocking=pin)
            if pin:
                ev = torch.cuda.Event()
                ev.record(stream)
                copied[slot] = ev
            Bp = X @ Tt.T
            Bv.index_copy_(1, idx_nm[c0:c1], Bp)
            Bv.index_copy_(1, idx_mn[c0:c1], Bp)
            del Bp
            if progress is not None:
                progress(f"B: shell pairs {q1}/{npairs} ({c1}/{ncols} packed columns)")
    if pin:
        torch.cuda.synchronize(device)
    if stats is not None:
        stats.update(total_s=time.perf_counter() - t_start, host_s=host_s[0], blocks=len(blocks),
                     packed_columns=ncols, metric_2c_s=t_m2c - t_start, fitting_matrix_s=t_fit - t_m2c,
                     pipeline_s=time.perf_counter() - t_fit)
    return B


def host_df_integrals(system):
    """Host part of B formation for a job prepared off the serving thread:
    (M, packed) with M = (P|Q) and packed = (Tt, m, n), Tt the unique half of
    (Q|mn) over the significant shell pairs as an (ncols, naux) array (one row
    per packed column) and m, n the basis functions of its columns
    (pair_columns)."""
    import numpy as np

    aux, bas = system.auxiliary_basis, system.basis
    col0, m, n = pair_columns(bas)
    Tt = np.empty((int(col0[-1]), aux.size))
    if Tt.shape[0]:
        _ints.coulomb_3c_pairs_t(aux, bas, 0, len(col0) - 1, Tt)
    return _ints.coulomb_2c(aux), (Tt, m, n)


def device_B_from_host(M, packed, device, nbf, df_ortho_rtol=None, block_bytes=1 << 30):
    """B = X T on the device from the host_df_integrals output (same algebra
    as build_B_density_fitting): the packed columns are multiplied by X in
    blocks and scattered to (m, n) and (n, m)."""
    Tt, m, n = packed
    Q, N = Tt.shape[1], int(nbf)
    kind, fac = metric_factor(torch.from_numpy(M).to(device), df_ortho_rtol)
    X = _fitting_matrix(kind, fac)
    B = torch.zeros((Q, N, N), dtype=torch.float64, device=device)
    Bv = B.view(Q, N * N)
    idx_mn = torch.from_numpy(m * N + n).to(device)
    idx_nm = torch.from_numpy(n * N + m).to(device)
    ncols = Tt.shape[0]
    width = max(1, block_bytes // (8 * Q))
    for c0 in range(0, ncols, width):
        c1 = min(ncols, c0 + width)
        Bp = X @ torch.from_numpy(Tt[c0:c1]).to(device).T
        Bv.index_copy_(1, idx_nm[c0:c1], Bp)
        Bv.index_copy_(1, idx_mn[c0:c1], Bp)
    return B
