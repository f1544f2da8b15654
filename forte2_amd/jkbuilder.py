"""Device-resident density-fitted J/K builder (drop-in for forte2's FockBuilder).

Mirrors the duck-typed surface of ``forte2.jkbuilder.FockBuilder``
(forte2/jkbuilder/jkbuilder.py:21-595) that the RHF/UHF/MCSCF loops call:
``build_JK``, ``build_J``, ``build_K``, ``build_JK_generalized``,
``build_core_fock``, ``build_active_fock``, ``build_generalized_fock`` and the
attributes ``system``, ``nbf``, ``naux``, ``B_Pmn``.

The 3-index tensor B(P|mn) lives in HBM as a torch tensor and every
contraction runs in the hand-written HIP kernels of ``libf2jk.so`` through its
C ABI.  Inputs may be numpy arrays (results come back as numpy, like the
reference) or torch tensors on the GPU (results stay on the device).
"""

import ctypes

import numpy as np
import torch

from . import _native as nat


def _device_of(device):
    if device is None:
        return torch.device("cuda", torch.cuda.current_device())
    d = torch.device(device)
    if d.type != "cuda":
        raise ValueError("FockBuilder runs on a HIP device only")
    if d.index is None:
        d = torch.device("cuda", torch.cuda.current_device())
    return d


class FockBuilder:
    """J/K builder with B(P|mn) resident in HBM.

    Parameters
    ----------
    system : System, optional
        Molecule and basis sets; B is built from it on first use
        (forte2/jkbuilder/jkbuilder.py:52-72, 119-159).
    store_B_nPm : bool
        Accepted for signature compatibility with the reference
        (forte2/jkbuilder/jkbuilder.py:46-50).  The device kernels read the
        (P, m, n) layout directly, so no transposed copy is ever made.
    B : array or tensor (naux, nbf, nbf), optional
        A ready B tensor (e.g. from a previous run); copied to the device.
    precision : {"fp64", "fp32", "bf16x3"}
        Arithmetic of ``build_*`` calls (see include/f2jk.h).  fp32/bf16x3
        create an FP32 copy of B next to the FP64 one.
    j_sparse : bool
        J passes over the lower-triangle elements where B is nonzero for some
        P only (B is scanned once per copy; its zero pattern must not change).
    """

    def __init__(self, system=None, store_B_nPm=True, *, B=None, precision="fp64", device=None, j_sparse=True):
        self.system = system
        self.store_B_nPm = store_B_nPm
        self.device = _device_of(device)
        self._B64 = None
        self._B32 = None
        self._ctx = None
        self._ws = None
        self._manual = set()  # launch-shape knobs set by hand
        self._tuned = None    # precision whose tuned launch shape is applied
        self.knobs = {}       # launch-shape knobs set on the context (by hand or from the table)
        self.j_sparse = bool(j_sparse)
        self.timing = False
        self.last_timings = None
        self.last_launches = None
        self.progress = None
        self.B_stats = None  # timings of the lazy B formation
        if precision not in nat.PREC:
            raise ValueError(f"unknown precision {precision!r}")
        self.precision = precision
        if B is not None:
            self._set_B(B)
        elif system is not None:
            self.nbf = system.nbf
        else:
            raise ValueError("FockBuilder needs a system or a B tensor")

    # ------------------------------------------------------------------ B
    def _set_B(self, B):
        if isinstance(B, np.ndarray):
            B = torch.from_numpy(np.ascontiguousarray(B, dtype=np.float64))
        B = B.to(device=self.device, dtype=torch.float64).contiguous()
        if B.dim() != 3 or B.shape[1] != B.shape[2]:
            raise ValueError("B must have shape (naux, nbf, nbf)")
        self._B64 = B
        self._B32 = None
        self.naux, self.nbf = int(B.shape[0]), int(B.shape[1])
        if self._ctx is not None:
            self._register(nat.F2JK_F64, self._B64)

    @property
    def B_Pmn(self):
        """B(P|mn) as a device tensor (built lazily from the system)."""
        if self._B64 is None:
            self.B_stats = {}
            if getattr(self.system, "cholesky_tei", False):
                # Cholesky vectors of (mn|ls); naux is their number
                # (forte2/jkbuilder/jkbuilder.py:66-70, 90-117)
                from .cholesky import build_B_cholesky

                self._set_B(build_B_cholesky(self.system, self.system.cholesky_tol, self.device,
                                             stats=self.B_stats))
                self.system.naux = self.naux
            else:
                from .df import build_B_density_fitting

                self._set_B(build_B_density_fitting(self.system, device=self.device, progress=self.progress,
                                                    stats=self.B_stats))
        return self._B64

    def B_fp32(self):
        """The FP32 copy of B used by the fp32/bf16x3 precision forms."""
        if self._B32 is None:
            self._B32 = self.B_Pmn.to(torch.float32)
            if self._ctx is not None:
                self._register(nat.F2JK_F32, self._B32)
        return self._B32

    # ------------------------------------------------------------------ native context
    @property
    def ctx(self):
        if self._ctx is None:
            lib = nat.jk_lib()
            h = ctypes.c_void_p()
            nat.jk_check(lib.f2jk_ctx_create(self.device.index, ctypes.byref(h)), "f2jk_ctx_create")
            self._ctx = h
            self._register(nat.F2JK_F64, self.B_Pmn)
            if self._B32 is not None:
                self._register(nat.F2JK_F32, self._B32)
            # J reads only the elements where B is nonzero for some P (B is
            # fixed once formed; the scan happens here and when the FP32 copy
            # is registered)
            if self.j_sparse:
                self._tune("j_sparse", 1)
        return self._ctx

    def _register(self, dtype, B):
        lib = nat.jk_lib()
        nat.jk_check(
            lib.f2jk_set_B(self._ctx, dtype, ctypes.c_void_p(B.data_ptr()), B.shape[0], B.shape[1]),
            "f2jk_set_B",
        )

    def set_tuning(self, key, value):
        """Set a launch-shape knob (f2jk_set_tuning).  A knob set here is kept:
        the autotuned table no longer changes it."""
        self._manual.add(key)
        self._tune(key, value)

    def _tune(self, key, value):
        nat.jk_check(nat.jk_lib().f2jk_set_tuning(self.ctx, key.encode(), int(value)), "f2jk_set_tuning")
        self.knobs[key] = int(value)

    def _apply_table(self, prec):
        """Launch shape of this size class and precision from the autotuned
        table (forte2_amd/data/tuning_gfx950.json), for knobs not set by hand."""
        if self._tuned == prec:
            return
        self._tuned = prec
        from . import autotune

        e = autotune.lookup(self.nbf, self.naux, prec)
        if e is not None:
            for k in autotune.KEYS:
                if k in e and k not in self._manual:
                    self._tune(k, e[k])

    def release(self):
        """Give back the device memory now: destroy the native context and drop
        B, its FP32 copy and the workspace (the builder is unusable after).
        A builder that is part of a reference cycle (System <-> FockBuilder)
        would otherwise keep B alive until Python's cycle collector runs."""
        if self._ctx is not None:
            nat.jk_lib().f2jk_ctx_destroy(self._ctx)
            self._ctx = None
        self._B64 = self._B32 = self._ws = None

    def __del__(self):
        try:
            if self._ctx is not None:
                nat.jk_lib().f2jk_ctx_destroy(self._ctx)
        except Exception:
            pass

    def _workspace(self, nbytes):
        if self._ws is None or self._ws.numel() < nbytes:
            self._ws = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
        return self._ws

    def _prec(self, precision):
        p = self.precision if precision is None else precision
        if p not in nat.PREC:
            raise ValueError(f"unknown precision {p!r}")
        if p != "fp64":
            self.B_fp32()
        self._apply_table(p)
        return nat.PREC[p]

    # ------------------------------------------------------------------ helpers
    def _as_dev(self, x):
        if isinstance(x, torch.Tensor):
            return x.to(device=self.device, dtype=torch.float64), False
        return torch.from_numpy(np.asarray(x, dtype=np.float64)).to(self.device), True

    @staticmethod
    def _rowmajor(x):
        """(tensor, leading dimension) with unit column stride."""
        if x.dim() != 2:
            raise ValueError("expected a 2-D matrix")
        if x.stride(1) != 1 or x.stride(0) < max(1, x.shape[1]):
            x = x.contiguous()
        return x, max(int(x.stride(0)), 1)

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _collect_timings(self):
        if self.timing:
            ms = (ctypes.c_double * len(nat.PHASES))()
            nat.jk_check(nat.jk_lib().f2jk_last_timings(self.ctx, ms, len(nat.PHASES)), "timings")
            self.last_timings = dict(zip(nat.PHASES, list(ms)))
            cnt = (ctypes.c_int * len(nat.PHASES))()
            nat.jk_check(nat.jk_lib().f2jk_last_launches(self.ctx, cnt, len(nat.PHASES)), "launches")
            self.last_launches = dict(zip(nat.PHASES, list(cnt)))

    # ------------------------------------------------------------------ builds
    def _jk_call(self, C, want_j, want_k, precision=None):
        if not isinstance(C, (list, tuple)):
            raise TypeError("C must be a list of coefficient matrices")
        ctx = self.ctx
        prec = self._prec(precision)
        N = self.nbf
        to_np = False
        Cs, nocc, ldc = [], [], []
        for Ci in C:
            t, was_np = self._as_dev(Ci)
            to_np = to_np or was_np
            if t.dim() != 2 or t.shape[0] != N:
                raise ValueError(f"C must have {N} rows")
            t, ld = self._rowmajor(t)
            Cs.append(t)
            nocc.append(int(t.shape[1]))
            ldc.append(ld)
        nd = len(Cs)
        J = [torch.empty((N, N), dtype=torch.float64, device=self.device) for _ in range(nd)] if want_j else []
        K = [torch.empty((N, N), dtype=torch.float64, device=self.device) for _ in range(nd)] if want_k else []
        flags = (nat.F2JK_WANT_J if want_j else 0) | (nat.F2JK_WANT_K if want_k else 0)
        if self.timing:
            flags |= nat.F2JK_TIMING
        lib = nat.jk_lib()
        nocc_a = nat.i64_array(nocc)
        nbytes = ctypes.c_size_t()
        nat.jk_check(lib.f2jk_workspace_size(ctx, nd, nocc_a, prec, flags, ctypes.byref(nbytes)), "workspace")
        ws = self._workspace(nbytes.value)
        cp = nat.ptr_array([c.data_ptr() for c in Cs])
        jp = nat.ptr_array([j.data_ptr() for j in J]) if want_j else None
        kp = nat.ptr_array([k.data_ptr() for k in K]) if want_k else None
        ldc_a = nat.i64_array(ldc)
        if want_j:
            rc = lib.f2jk_build_jk(ctx, nd, cp, nocc_a, ldc_a, jp, kp, prec, flags,
                                   ctypes.c_void_p(ws.data_ptr()), ws.numel(), self._stream())
            nat.jk_check(rc, "f2jk_build_jk")
        else:
            rc = lib.f2jk_build_k(ctx, nd, cp, nocc_a, ldc_a, kp, prec, flags,
                                  ctypes.c_void_p(ws.data_ptr()), ws.numel(), self._stream())
            nat.jk_check(rc, "f2jk_build_k")
        self._collect_timings()
        if to_np:
            J = [j.cpu().numpy() for j in J]
            K = [k.cpu().numpy() for k in K]
        return J, K

    def build_J(self, D, precision=None):
        """J_k = sum_P B_P <B_P, D_k> for each D_k (jkbuilder.py:161-166)."""
        if not isinstance(D, (list, tuple)):
            raise TypeError("D must be a list of density matrices")
        ctx = self.ctx
        prec = self._prec(precision)
        N = self.nbf
        to_np = False
        Ds = []
        for Di in D:
            t, was_np = self._as_dev(Di)
            to_np = to_np or was_np
            if t.shape != (N, N):
                raise ValueError(f"D must be {N}x{N}")
            Ds.append(t.contiguous())
        nd = len(Ds)
        J = [torch.empty((N, N), dtype=torch.float64, device=self.device) for _ in range(nd)]
        flags = nat.F2JK_WANT_J | (nat.F2JK_TIMING if self.timing else 0)
        lib = nat.jk_lib()
        nbytes = ctypes.c_size_t()
        nat.jk_check(lib.f2jk_workspace_size(ctx, nd, None, prec, flags, ctypes.byref(nbytes)), "workspace")
        ws = self._workspace(nbytes.value)
        rc = lib.f2jk_build_j(ctx, nd, nat.ptr_array([d.data_ptr() for d in Ds]), N,
                              nat.ptr_array([j.data_ptr() for j in J]), prec, flags,
                              ctypes.c_void_p(ws.data_ptr()), ws.numel(), self._stream())
        nat.jk_check(rc, "f2jk_build_j")
        self._collect_timings()
        return [j.cpu().numpy() for j in J] if to_np else J

    def build_K(self, C, precision=None):
        """K_k = sum_{P,i} (B_P C_k)_mi (B_P C_k)_ni (jkbuilder.py:184-207)."""
        return self._jk_call(C, False, True, precision)[1]

    def build_JK(self, C, precision=None):
        """(J_k, K_k) for each C_k with D_k = C_k C_k^T (jkbuilder.py:209-251)."""
        return self._jk_call(C, True, True, precision)

    # ------------------------------------------------------------------ generalized Fock
    def build_JK_generalized(self, C, g1, precision=None):
        """J/K of the density C g1 C^T via a factor of g1 (jkbuilder.py:253-290)."""
        t, to_np = self._as_dev(C)
        g, _ = self._as_dev(g1)
        if t.shape[1] != g.shape[0]:
            raise AssertionError("C and g1 must have compatible dimensions")
        try:
            L = torch.linalg.cholesky(g)
            Cp = t @ L
        except RuntimeError:
            n, U = torch.linalg.eigh(g)
            if bool((n < -1.0e-11).any()):
                raise AssertionError("g1 must be positive semi-definite")
            Cp = t @ U @ torch.diag(torch.sqrt(torch.clamp(n, min=0.0)))
        J, K = self._jk_call([Cp], True, True, precision)
        if to_np:
            return J[0].cpu().numpy(), K[0].cpu().numpy()
        return J[0], K[0]

    # The assembly methods below call build_JK / build_JK_generalized with the
    # reference's own arguments when no precision is given, so a builder whose
    # build methods are replaced (as forte2's tests do) assembles the same way.
    def build_core_fock(self, C_core, hcore=None, precision=None):
        """F_core = h + 2 J[C_core] - K[C_core], h + J - K for two-component
        systems (jkbuilder.py:292-321)."""
        if hcore is None:
            hcore = self.system.ints_hcore()
        J, K = self.build_JK([C_core]) if precision is None else self.build_JK([C_core], precision=precision)
        if getattr(self.system, "two_component", False):
            return hcore + J[0] - K[0]
        return hcore + 2.0 * J[0] - K[0]

    def build_active_fock(self, C_act, g1, precision=None):
        """F_act = J[g1] - K[g1] / 2, J - K for two-component systems
        (jkbuilder.py:323-349)."""
        if precision is None:
            J, K = self.build_JK_generalized(C_act, g1)
        else:
            J, K = self.build_JK_generalized(C_act, g1, precision=precision)
        if getattr(self.system, "two_component", False):
            return J - K
        return J - 0.5 * K

    def build_generalized_fock(self, C_core, C_act, g1, hcore=None, precision=None):
        """F = F_core + F_act (jkbuilder.py:351-380)."""
        return (self.build_core_fock(C_core, hcore=hcore, precision=precision)
                + self.build_active_fock(C_act, g1, precision=precision))

    # ------------------------------------------------------------------ MO-basis blocks
    def half_transform(self, C, p0=0, p1=None, out=None):
        """X[P - p0, m, i] = sum_n B_Pmn C_ni for p0 <= P < p1, FP64 on the
        device (f2jk_half_transform, the MFMA kernel of the K build).  C is a
        device tensor (nbf, ncol); out, if given, a contiguous FP64 device
        tensor (p1 - p0, nbf, ncol)."""
        lib = nat.jk_lib()
        p1 = self.naux if p1 is None else int(p1)
        t, ld = self._rowmajor(C)
        N, ncol = self.nbf, int(t.shape[1])
        if t.shape[0] != N:
            raise ValueError(f"C must have {N} rows")
        shape = (p1 - p0, N, ncol)
        if out is None:
            out = torch.empty(shape, dtype=torch.float64, device=self.device)
        elif tuple(out.shape) != shape or out.dtype != torch.float64 or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous float64 tensor of shape {shape}")
        nbytes = ctypes.c_size_t()
        nat.jk_check(lib.f2jk_workspace_size(self.ctx, 1, nat.i64_array([ncol]), nat.PREC["fp64"], nat.F2JK_WANT_K,
                                             ctypes.byref(nbytes)), "workspace")
        ws = self._workspace(nbytes.value)
        nat.jk_check(lib.f2jk_half_transform(self.ctx, ctypes.c_void_p(t.data_ptr()), ncol, ld, p0, p1,
                                             ctypes.c_void_p(out.data_ptr()), ncol, nat.PREC["fp64"],
                                             ctypes.c_void_p(ws.data_ptr()), ws.numel(), self._stream()),
                     "f2jk_half_transform")
        return out

    def _x_blocks(self, C, block_bytes=1 << 30):
        """Yields (P0, P1, X) with X = half_transform(C, P0, P1) for blocks of P
        of at most block_bytes."""
        N, Q, ncol = self.nbf, self.naux, int(C.shape[1])
        step = max(1, min(Q, block_bytes // (8 * N * ncol)))
        for P0 in range(0, Q, step):
            P1 = min(Q, P0 + step)
            yield P0, P1, self.half_transform(C, P0, P1)

    def B_tensor_gen_block(self, C1, C2):
        """B[P, i, j] = sum_mn B_Pmn C1_mi C2_nj: the MO-basis B tensor
        (jkbuilder.py:521-552).  X = B C2 on MFMA, then C1^T X as a batched
        GEMM, by blocks of P.  Real orbitals (conj is the identity)."""
        self.B_Pmn
        c1, to_np = self._as_dev(C1)
        c2, to_np2 = self._as_dev(C2)
        if c1.dim() != 2 or c2.dim() != 2 or c1.shape[0] != self.nbf or c2.shape[0] != self.nbf:
            raise ValueError(f"C1 and C2 must have {self.nbf} rows")
        out = torch.empty((self.naux, c1.shape[1], c2.shape[1]), dtype=torch.float64, device=self.device)
        if out.numel() == 0:
            return out.cpu().numpy() if (to_np or to_np2) else out
        c1t = c1.T.contiguous()
        for P0, P1, X in self._x_blocks(c2):
            out[P0:P1] = torch.matmul(c1t, X)
        return out.cpu().numpy() if (to_np or to_np2) else out

    def two_electron_integrals_gen_block(self, C1, C2, C3, C4):
        """V[i, j, k, l] = <ij|kl> = sum_P B13[P, i, k] B24[P, j, l] with
        B13 = B_tensor_gen_block(C1, C3), B24 = B_tensor_gen_block(C2, C4)
        (jkbuilder.py:382-420): one GEMM over P, then (i, k, j, l) -> (i, j, k, l)."""
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
        """<pq|rs> over one orbital set (jkbuilder.py:422-443)."""
        c, to_np = self._as_dev(C)
        V = self.two_electron_integrals_gen_block(c, c, c, c)
        return V.cpu().numpy() if to_np else V

    # ------------------------------------------------------------------ measurement
    def work(self, nocc, precision="fp64", want_j=True, want_k=True):
        """(J bytes, K flops) of one build: Bytes_J = 2 Q N(N+1)/2 s_B + 8 nd (2Q + 2N^2),
        Flops_K = sum_k 2 Q N^2 o_k + N(N+1) Q o_k."""
        jb, kf = ctypes.c_double(), ctypes.c_double()
        flags = (nat.F2JK_WANT_J if want_j else 0) | (nat.F2JK_WANT_K if want_k else 0)
        nat.jk_check(nat.jk_lib().f2jk_work(self.ctx, len(nocc), nat.i64_array(nocc), nat.PREC[precision],
                                            flags, ctypes.byref(jb), ctypes.byref(kf)), "f2jk_work")
        return jb.value, kf.value
