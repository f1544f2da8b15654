"""Integral-direct J/K on the device: the drop-in for forte2's FockBuilderOTF
(forte2/jkbuilder/jkbuilder.py:598-1265).

The 3-centre integrals (P|mn) are never stored whole.  They are evaluated by
the host engine one block of auxiliary shells at a time, copied to HBM, and
contracted there:

* J (two sweeps over the blocks): gamma_P = sum_mn (P|mn) D_mn, then
  gamma = M^-1 gamma (Cholesky solve, or the truncated eigen-decomposition
  when df_ortho_rtol is set), then J_mn = sum_P gamma_P (P|mn)
  (jkbuilder.py:813-840);
* K: X_Qmi = sum_n (Q|mn) C_ni on MFMA (the half-transform kernel of the
  in-core K build, f2jk_half_transform), Y = M^-1/2 X, K = sum_Pi Y_Pmi Y_Pni
  (jkbuilder.py:842-875), by blocks of occupied columns;
* build_JK: the first sweep of K also forms gamma, the second finishes J
  (jkbuilder.py:945-1019).

Differences from the reference, none of which changes a result beyond
summation order:

* the columns of all densities of one call share the sweeps (the reference
  repeats every sweep per density), packed into blocks of at most
  ``iblksize`` columns that never straddle two densities;
* the host evaluates the next block on a worker thread while the device
  contracts the current one (two pinned staging buffers);
* ``cache_mb`` keeps up to that many MB of (P|mn) blocks resident in HBM, so
  that later sweeps and builds skip their host evaluation (0: none, as the
  reference).

The block sizes follow the reference's memory rule (jkbuilder.py:676-740)
with the budget ``jk_mem_thres_mb`` (MB of HBM), so the same budgets are
accepted and rejected.
"""

import concurrent.futures
import math
import time

import numpy as np
import torch

from . import integrals as _ints
from .df import _fitting_matrix, metric_factor, pair_columns
from .jkbuilder import FockBuilder, _device_of


def block_sizes(jk_mem_thres_mb, nbf, naux, nocc_guess, max_shell):
    """(iblksize, pblksize) of a memory budget (jkbuilder.py:676-716, real
    orbitals: one (P|mn) block of 8 pblk N^2 bytes plus two (Q, N, iblk)
    buffers of 8 bytes per element)."""
    budget = jk_mem_thres_mb * 1024**2
    per_iblk = (2 * 8 + 8) * nbf * naux
    iblk = min(max(1, nocc_guess), math.ceil(budget / per_iblk))
    pblk = min(naux, math.floor((budget - 2 * 8 * nbf * naux * iblk) / (8 * nbf**2)))
    if pblk < max_shell:
        need = math.ceil((2 * 8 + 8) * max_shell * naux**2 / 1024**2)
        raise ValueError(
            f"[FockBuilderOTF]: Memory threshold {jk_mem_thres_mb} is too low to even hold the largest shell "
            f"of the auxiliary basis (of size {max_shell})). Please increase the memory threshold to {need} MB.")
    return iblk, pblk


def aux_shell_blocks(offsets, pblksize):
    """[(s0, s1, p0, p1)]: consecutive aux shells whose functions fit pblksize
    (jkbuilder.py:784-794)."""
    out, s0, nsh = [], 0, len(offsets) - 1
    while s0 < nsh:
        s1 = s0
        while s1 < nsh and offsets[s1 + 1] - offsets[s0] <= pblksize:
            s1 += 1
        out.append((s0, s1, offsets[s0], offsets[s1]))
        s0 = s1
    return out


def column_sweeps(noccs, iblksize):
    """Sweeps of column segments: [[(k, i0, i1), ...], ...].  Each density's
    columns are cut into segments of at most iblksize, and segments are
    packed, in order, into sweeps of at most iblksize columns."""
    segs = [(k, i0, min(o, i0 + iblksize)) for k, o in enumerate(noccs) for i0 in range(0, o, iblksize)]
    sweeps, cur, w = [], [], 0
    for k, i0, i1 in segs:
        if cur and w + (i1 - i0) > iblksize:
            sweeps.append(cur)
            cur, w = [], 0
        cur.append((k, i0, i1))
        w += i1 - i0
    if cur:
        sweeps.append(cur)
    return sweeps


class FockBuilderOTF:
    """Integral-direct J/K builder (forte2/jkbuilder/jkbuilder.py:598-640).

    Parameters
    ----------
    system : System
        Basis, auxiliary basis and df_ortho_rtol.
    jk_mem_thres_mb : float
        HBM budget (MB) that fixes the block sizes, as in the reference.
    backend : {"auto", "libint2", "libcint"}
        Accepted for signature compatibility (jkbuilder.py:614-618); every
        value uses the host integral engine of this package.
    cache_mb : float
        HBM (MB) for (P|mn) blocks kept between sweeps (0: recompute all).
    """

    def __init__(self, system, jk_mem_thres_mb=4000, backend="auto", *, cache_mb=0, device=None):
        if backend.lower() not in ["auto", "libint2", "libcint"]:
            raise ValueError(f"Invalid backend {backend}. Must be one of 'auto', 'libint2', or 'libcint'.")
        if system.auxiliary_basis is None:
            raise ValueError("Auxiliary basis is not defined. Define auxiliary_basis_set.")
        self.backend = backend.lower()
        self.system = system
        self.device = _device_of(device if device is not None else getattr(system, "device", None))
        self.nbf = system.nbf
        self.jk_mem_thres_mb = jk_mem_thres_mb
        self.auxbasis = system.auxiliary_basis
        self.basis = system.basis
        self.naux = self.auxbasis.size
        self.aux_sh_offsets = self.auxbasis.shell_offsets()
        self.metric_ortho_rtol = getattr(system, "df_ortho_rtol", None)
        self._build_metric()
        max_shell = max(b - a for a, b in zip(self.aux_sh_offsets[:-1], self.aux_sh_offsets[1:]))
        self.iblksize, self.pblksize = block_sizes(jk_mem_thres_mb, self.nbf, self.naux,
                                                   int(system.Zsum) // 2, max_shell)
        self.blocks = aux_shell_blocks(self.aux_sh_offsets, self.pblksize)
        N = self.nbf
        # The host evaluates each aux block over the significant shell pairs
        # (m >= n) only, packed; the device scatters the columns to (m, n)
        # and (n, m).  Every other element of a block is zero (the full-square
        # routine screens the same pairs), so the block buffers are zeroed
        # once and only the pair positions are ever written.
        col0, mcol, ncol = pair_columns(self.basis)
        self._npairs, self._ncols = len(col0) - 1, int(col0[-1])
        self._idx_mn = torch.from_numpy(mcol * N + ncol).to(self.device)
        self._idx_nm = torch.from_numpy(ncol * N + mcol).to(self.device)
        self._dev_block = torch.zeros((self.pblksize, N, N), dtype=torch.float64, device=self.device)
        # staging of a block: one row of pblksize values per packed column
        # (the engine's column-major output: each thread fills whole rows)
        w = max(1, self._ncols)
        self._dev_packed = torch.empty((w, self.pblksize), dtype=torch.float64, device=self.device)
        pin = self.device.type == "cuda"
        self._host = [torch.empty((w, self.pblksize), dtype=torch.float64, pin_memory=pin) for _ in range(2)]
        self._pool = concurrent.futures.ThreadPoolExecutor(max_workers=1)
        # a device FockBuilder whose B is re-pointed at each block
        self._kx = FockBuilder(B=self._dev_block, device=self.device)
        self.cache_bytes = int(cache_mb * 1024**2)
        self._cache = {}
        self.host_blocks_computed = 0  # blocks evaluated by the host engine so far
        # With a single aux block, the later sweeps of one build (the J sweep
        # after the K/gamma sweep) reuse the block left on the device instead
        # of evaluating it again; each build still evaluates it once.
        self._keep_resident = False
        self._resident = False
        self.host_seconds = 0.0        # wall time of those evaluations (worker thread)

    # ------------------------------------------------------------------ metric
    def _build_metric(self):
        """M = (P|Q) on the device; M^-1/2 for K and a factor for M^-1 y
        (jkbuilder.py:742-775)."""
        M = torch.from_numpy(_ints.coulomb_2c(self.auxbasis)).to(self.device)
        kind, fac = metric_factor(M, self.metric_ortho_rtol)
        self._kind = kind
        if kind == "chol":
            self.L = fac
            self.Mm12 = _fitting_matrix("chol", fac)
        else:
            w, U = torch.linalg.eigh(M)
            w = torch.clamp(w, min=0.0)
            keep = w >= self.metric_ortho_rtol * w[-1]
            self.sevals, self.sevecs = w[keep], U[:, keep]
            self.Mm12 = fac

    def _apply_Mm1(self, y):
        """x = M^-1 y (jkbuilder.py:777-782): Cholesky solve, or U s^-1 U^T y
        over the kept eigenpairs."""
        if self._kind == "chol":
            return torch.cholesky_solve(y, self.L, upper=False)
        return self.sevecs @ ((self.sevecs.T @ y) / self.sevals[:, None])

    # ------------------------------------------------------------------ (P|mn) blocks
    def _host_block(self, k, slot):
        s0, s1, p0, p1 = self.blocks[k]
        t0 = time.perf_counter()
        if self._ncols:
            _ints.coulomb_3c_pairs_aux_t(self.auxbasis, self.basis, s0, s1, 0, self._npairs,
                                         self._host[slot].numpy())
        self.host_seconds += time.perf_counter() - t0

    def _sweep(self):
        """Yields (p0, p1, T) for every aux-shell block, T = (P|mn) on the
        device, (p1 - p0, N, N).  T is valid until the next block is
        requested (stream-ordered: work queued on it before then is safe).
        However the sweep ends (also when the consumer stops early), the
        pending host block is waited for and the copies out of the staging
        buffers are complete before the next sweep writes them again."""
        N = self.nbf
        if self._resident and len(self.blocks) == 1 and 0 not in self._cache:
            # the only block is still on the device from this build's first sweep
            p0, p1 = self.blocks[0][2], self.blocks[0][3]
            yield p0, p1, self._dev_block[: p1 - p0]
            return
        todo = [k for k in range(len(self.blocks)) if k not in self._cache]
        copied = [None, None]  # event: the device copy out of that staging buffer is done
        fut = self._pool.submit(self._host_block, todo[0], 0) if todo else None
        it = 0
        stream = torch.cuda.current_stream(self.device)
        try:
            for k, (s0, s1, p0, p1) in enumerate(self.blocks):
                if k in self._cache:
                    yield p0, p1, self._cache[k]
                    continue
                slot = it % 2
                fut.result()
                self.host_blocks_computed += 1
                if it + 1 < len(todo):
                    nxt = (it + 1) % 2
                    if copied[nxt] is not None:
                        copied[nxt].synchronize()
                    fut = self._pool.submit(self._host_block, todo[it + 1], nxt)
                nP = p1 - p0
                if self.cache_bytes - sum(t.numel() * 8 for t in self._cache.values()) >= nP * N * N * 8:
                    T = torch.zeros((nP, N, N), dtype=torch.float64, device=self.device)
                    self._cache[k] = T
                else:
                    T = self._dev_block[:nP]
                self._dev_packed.copy_(self._host[slot], non_blocking=True)
                packed = self._dev_packed[: self._ncols, :nP].T
                ev = torch.cuda.Event()
                ev.record(stream)
                copied[slot] = ev
                Tv = T.view(nP, N * N)
                Tv.index_copy_(1, self._idx_nm, packed)
                Tv.index_copy_(1, self._idx_mn, packed)
                self._resident = self._keep_resident and T.data_ptr() == self._dev_block.data_ptr()
                it += 1
                yield p0, p1, T
        finally:
            if fut is not None:
                fut.result()
            for ev in copied:
                if ev is not None:
                    ev.synchronize()

    def _half_transform(self, T, C, out):
        """out[P, m, i] = sum_n T_Pmn C_ni on the device (MFMA kernel)."""
        self._kx._set_B(T)
        return self._kx.half_transform(C, 0, T.shape[0], out=out)

    # ------------------------------------------------------------------ helpers
    def _as_dev(self, x):
        if isinstance(x, torch.Tensor):
            return x.to(device=self.device, dtype=torch.float64), False
        return torch.from_numpy(np.asarray(x, dtype=np.float64)).to(self.device), True

    @staticmethod
    def _check_precision(precision):
        if precision not in (None, "fp64"):
            raise ValueError("the integral-direct builder computes in fp64 only")

    def _mats(self, X, what):
        if not isinstance(X, (list, tuple)):
            raise TypeError(f"{what} must be a list of matrices")
        out, to_np = [], False
        for x in X:
            t, was_np = self._as_dev(x)
            to_np = to_np or was_np
            if t.dim() != 2 or t.shape[0] != self.nbf:
                raise ValueError(f"{what} must have {self.nbf} rows")
            out.append(t)
        return out, to_np

    # ------------------------------------------------------------------ J and K
    def _run(self, Cs, Ds, want_k):
        """J of the densities Ds (None: no J) and K of the orbitals Cs (if
        want_k), sharing the sweeps over (P|mn)."""
        self._keep_resident, self._resident = True, False
        try:
            return self._run_sweeps(Cs, Ds, want_k)
        finally:
            self._keep_resident = self._resident = False

    def _run_sweeps(self, Cs, Ds, want_k):
        N, Q, dev = self.nbf, self.naux, self.device
        nd_j = len(Ds) if Ds is not None else 0
        J = [torch.zeros((N, N), dtype=torch.float64, device=dev) for _ in range(nd_j)]
        K = [torch.zeros((N, N), dtype=torch.float64, device=dev) for _ in range(len(Cs))] if want_k else []
        Dflat = torch.stack([d.reshape(N * N) for d in Ds], 1) if nd_j else None  # (N^2, nd)
        gamma = torch.zeros((Q, nd_j), dtype=torch.float64, device=dev) if nd_j else None
        sweeps = column_sweeps([int(c.shape[1]) for c in Cs], self.iblksize) if want_k else []
        jpass = 0  # 0: gamma pending; 1: J pending; 2: done
        if nd_j == 0:
            jpass = 2
        for sw in sweeps:
            Csw = torch.cat([Cs[k][:, i0:i1] for k, i0, i1 in sw], 1).contiguous()
            ncol = int(Csw.shape[1])
            X = torch.empty((Q, N, ncol), dtype=torch.float64, device=dev)
            for p0, p1, T in self._sweep():
                self._half_transform(T, Csw, X[p0:p1])
                if jpass == 0:
                    gamma[p0:p1] = T.view(p1 - p0, N * N) @ Dflat
                elif jpass == 1:
                    self._j_acc(J, gamma[p0:p1], T)
            Y = (self.Mm12 @ X.view(Q, N * ncol)).view(Q, N, ncol)
            del X
            c0 = 0
            for k, i0, i1 in sw:
                w = i1 - i0
                Yk = Y[:, :, c0:c0 + w].permute(1, 0, 2).reshape(N, Q * w)
                K[k] += Yk @ Yk.T
                c0 += w
            del Y
            if jpass == 0:
                gamma = self._apply_Mm1(gamma)
            jpass = min(jpass + 1, 2)
        if jpass == 0:
            for p0, p1, T in self._sweep():
                gamma[p0:p1] = T.view(p1 - p0, N * N) @ Dflat
            gamma = self._apply_Mm1(gamma)
            jpass = 1
        if jpass == 1:
            for p0, p1, T in self._sweep():
                self._j_acc(J, gamma[p0:p1], T)
        return J, K

    def _j_acc(self, J, g, T):
        nP, N = T.shape[0], self.nbf
        acc = (g.T @ T.view(nP, N * N)).view(-1, N, N)
        for k in range(len(J)):
            J[k] += acc[k]

    def _jk_call(self, C, want_j, want_k, precision=None):
        self._check_precision(precision)
        Cs, to_np = self._mats(C, "C")
        Ds = [c @ c.T for c in Cs] if want_j else None
        J, K = self._run(Cs, Ds, want_k)
        if to_np:
            J = [j.cpu().numpy() for j in J]
            K = [k.cpu().numpy() for k in K]
        return J, K

    def build_J(self, D, precision=None):
        """J_k[D_k] by two sweeps (jkbuilder.py:1157-1158, :813-840)."""
        self._check_precision(precision)
        Ds, to_np = self._mats(D, "D")
        J, _ = self._run([], [d.contiguous() for d in Ds], False)
        return [j.cpu().numpy() for j in J] if to_np else J

    def build_K(self, C, precision=None):
        """K_k[C_k] (jkbuilder.py:1160-1172, :842-875)."""
        return self._jk_call(C, False, True, precision)[1]

    def build_JK(self, C, precision=None):
        """(J_k, K_k) with D_k = C_k C_k^T (jkbuilder.py:1174-1205, :945-1019)."""
        return self._jk_call(C, True, True, precision)

    build_JK_generalized = FockBuilder.build_JK_generalized
    build_core_fock = FockBuilder.build_core_fock
    build_active_fock = FockBuilder.build_active_fock
    build_generalized_fock = FockBuilder.build_generalized_fock

    # ------------------------------------------------------------------ MO-basis blocks
    def B_tensor_gen_block(self, C1, C2):
        """B_Pij = sum_Q M^-1/2_PQ sum_mn (Q|mn) C1_mi C2_nj (jkbuilder.py:1207-1226)."""
        (c1, c2), to_np = self._mats([C1, C2], "C1 and C2")
        N, Q = self.nbf, self.naux
        n1, n2 = int(c1.shape[1]), int(c2.shape[1])
        B = torch.zeros((Q, n1, n2), dtype=torch.float64, device=self.device)
        if B.numel():
            c1t = c1.T.contiguous()
            for p0, p1, T in self._sweep():
                X = self._half_transform(T, c2, torch.empty((p1 - p0, N, n2), dtype=torch.float64,
                                                            device=self.device))
                B[p0:p1] = torch.matmul(c1t, X)
            B = (self.Mm12 @ B.view(Q, n1 * n2)).view(Q, n1, n2)
        return B.cpu().numpy() if to_np else B

    def two_electron_integrals_gen_block(self, C1, C2, C3, C4):
        """<ij|kl> = sum_P B13_Pik B24_Pjl (jkbuilder.py:1251-1254)."""
        to_np = not all(isinstance(c, torch.Tensor) for c in (C1, C2, C3, C4))
        same = C2 is C1 and C4 is C3
        C1, C2, C3, C4 = (self._as_dev(c)[0] for c in (C1, C2, C3, C4))
        B13 = self.B_tensor_gen_block(C1, C3)
        B24 = B13 if same else self.B_tensor_gen_block(C2, C4)
        Q, n1, n3 = B13.shape
        n2, n4 = B24.shape[1], B24.shape[2]
        V = (B13.reshape(Q, n1 * n3).T @ B24.reshape(Q, n2 * n4)).reshape(n1, n3, n2, n4).permute(0, 2, 1, 3)
        V = V.contiguous()
        return V.cpu().numpy() if to_np else V

    def two_electron_integrals_block(self, C):
        """<pq|rs> over one orbital set (jkbuilder.py:1256-1258)."""
        c, to_np = self._as_dev(C)
        V = self.two_electron_integrals_gen_block(c, c, c, c)
        return V.cpu().numpy() if to_np else V

    def release(self):
        """Give back the HBM and pinned staging now."""
        self._pool.shutdown(wait=True)
        self._kx.release()
        self._cache.clear()
        self._dev_block = self._dev_packed = None
        self._host = None
