"""B(P|mn) from a pivoted Cholesky decomposition of the electron-repulsion
matrix, on the device (the reference's Cholesky-TEI B source).

forte2 builds B either by density fitting or, with
``System(cholesky_tei=True, cholesky_tol=tol)``, by factoring the 4-index
matrix V[(mn),(ls)] = (mn|ls) as V ~= B^T B with complete pivoting
(forte2/jkbuilder/jkbuilder.py:90-117, forte2/helpers/matrix_functions.py:239-270,
LAPACK dpstrf).  The number of Cholesky vectors (the rank at tol) becomes naux.

Here the host engine evaluates (mn|ls) once (libf2ints.so) and the
factorisation runs on the device:

* d = diag(V); at step k the pivot p is the largest remaining d (the first
  one on ties), the new vector is L_k = (V_p - sum_{j<k} L_jp L_j) / sqrt(d_p),
  and d -= L_k^2 with d_p = 0.  The decomposition stops when the largest
  remaining d is <= tol (tol < 0: n * eps * max(diag V), dpstrf's default),
  as dpstrf does; every accepted step is the same arithmetic as an unblocked
  dpstrf, so the rank is the same up to rounding of d near tol.
* The steps run in blocks without reading anything back: a step whose pivot
  is already <= tol writes a zero vector, and one read of the block's pivots
  per block finds where the decomposition ended.
* B is made exactly symmetric in (m, n) at the end (the J kernels read the
  lower triangle only).
"""

import time

import numpy as np
import torch

from . import integrals as _ints


def pivoted_cholesky(V, tol, block=32):
    """Rows L (rank x n) with L^T L ~= V for a symmetric positive
    semi-definite V (n x n, FP64, any torch device); see the module text."""
    if V.dim() != 2 or V.shape[0] != V.shape[1]:
        raise ValueError("V must be square")
    n = V.shape[0]
    d = V.diagonal().clone()
    if tol < 0:
        tol = n * float(torch.finfo(V.dtype).eps) * float(d.max())
    # the rank is usually well below n: the vectors' storage grows by doubling
    L = torch.zeros((min(n, 8 * block), n), dtype=V.dtype, device=V.device)
    piv = torch.empty(n, dtype=V.dtype, device=V.device)
    zero = torch.zeros((), dtype=V.dtype, device=V.device)
    k = 0
    while k < n:
        kb = min(block, n - k)
        if k + kb > L.shape[0]:
            grown = torch.zeros((min(n, 2 * L.shape[0]), n), dtype=V.dtype, device=V.device)
            grown[:k] = L[:k]
            L = grown
        for j in range(k, k + kb):
            dm, p = torch.max(d, 0)
            piv[j] = dm
            pv = p.view(1)
            row = V.index_select(0, pv)[0]
            if j:
                row = row - L[:j].index_select(1, pv)[:, 0] @ L[:j]
            scale = torch.where(dm > tol, torch.rsqrt(torch.clamp(dm, min=torch.finfo(V.dtype).tiny)), zero)
            row = row * scale
            L[j] = row
            d.sub_(row * row)
            d.index_fill_(0, pv, 0.0)
        done = (piv[k:k + kb] <= tol).nonzero()
        if done.numel():
            return L[: k + int(done[0])]
        k += kb
    return L[:n]


def build_B_cholesky(system, tol, device, stats=None):
    """B[P][m][n] (rank x nbf x nbf, FP64, on `device`) from the system's
    (mn|ls) (forte2/jkbuilder/jkbuilder.py:90-117)."""
    t0 = time.perf_counter()
    nbf = system.nbf
    eri = _ints.coulomb_4c(system.basis)
    t1 = time.perf_counter()
    V = torch.from_numpy(np.ascontiguousarray(eri).reshape(nbf * nbf, nbf * nbf)).to(device)
    del eri
    L = pivoted_cholesky(V, tol)
    del V
    B = L.reshape(L.shape[0], nbf, nbf)
    B = 0.5 * (B + B.transpose(1, 2))
    if device is not None and torch.device(device).type == "cuda":
        torch.cuda.synchronize(device)
    if stats is not None:
        stats.update(total_s=time.perf_counter() - t0, host_s=t1 - t0, rank=int(B.shape[0]),
                     pairs=nbf * nbf)
    return B.contiguous()
