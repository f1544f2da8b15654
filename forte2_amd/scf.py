"""RHF / UHF on the device (FP64 linear algebra in HBM, J/K through libf2jk).

Mirrors the reference's SCF driver and its defaults (forte2/scf/scf_base.py:
62-251; rhf.py:35-65; uhf.py:40-127; helpers/diis.py:10-125):
  * usage ``RHF(charge=0)(system).run()`` / ``UHF(charge, ms=..)(system).run()``;
  * DIIS start 1 / nvec 8 / min 2, e_tol 1e-9, d_tol 1e-6 (Frobenius norm of
    the density change summed over spins), maxiter 100, SAP ("minao") or
    core-Hamiltonian guess, one extra diagonalisation + Fock build after
    convergence, RuntimeError on non-convergence;
  * canonical orthogonalisation X (eigen-truncated at 1e-8), eigh of X^T F X,
    DIIS and the energy all run on the device in FP64.

Precision form (``precision="bf16x3"`` or ``"fp32"``): the J/K builds of the
early iterations run in that precision on full densities; when |dE| drops
below ``switch_inc`` (3e-3), or the energy rises by less than ten times that
(the low-precision energy has reached its noise floor: at (H2O)_64 a BF16x3
build is 2e-3 Eh off FP64 and consecutive |dE| stall near 3e-4 with
alternating signs), an FP64 reference Fock matrix F_ref(D_ref) is built
and later iterations use incremental builds F = F_ref + G_lp(D - D_ref) whose
error scales with |D - D_ref|.  Once |dE| < ``switch_fp64`` (1e-6) the tail
runs in one of two ways (``fp64_tail``):
  * ``"reref"`` (default): one more FP64 reference build at the current
    density, then incremental builds again until convergence.  |D - D_ref| is
    now of the order of the remaining density change, so the low-precision
    part of F is ~1e-5 of an already small G(D - D_ref).  That part is
    rounded afresh every cycle (the factors of D - D_ref change), so the
    incremental energies scatter (a few 1e-10 Eh at (H2O)_64, up to ~1e-8 Eh
    for alkane cations in cc-pVTZ) and the energy test of these cycles is
    |dE| < max(e_tol, ``tail_e_tol``) (1e-8, the accuracy asked of the final
    energy), with the density test d_tol unchanged.  The converged
    iteration is followed, as in every form, by the extra diagonalisation
    and an FP64 build, so the reported energy is an FP64 energy at a density
    within ~|dF|/gap of the FP64 one (its error is second order in that).  If these incremental
    cycles have not converged ``inc_max_tail`` (12) cycles after the tail's
    reference, one more FP64 reference is taken at the current density, up
    to ``tail_refs`` (1) references in the tail; after that the remaining
    cycles use FP64 builds.  (A UHF cation with a small gap, C4H10+ in
    cc-pVTZ, keeps its density change near 1e-5 under the rounding of the
    incremental cycles whatever the reference, so it needs FP64 cycles at
    the end; a second reference only adds a build there.);
  * ``"full"``: the remaining cycles, including the converged one, use FP64
    builds.

The difference density of each spin is factored exactly as
D - D_ref = P A^T + A P^T with the reference orbitals rotated onto the new
ones (orthogonal Procrustes), P their midpoint and A their difference, so
K(D - D_ref) = (K[U+] - K[U-]) / 2 with U+- = sqrt(t) P +- A / sqrt(t).
Every run records, per iteration, the build used ("lp", "ref", "inc",
"fp64"); with ``lp_error_check`` it also builds the low-precision Fock
matrix once more at the first reference density and records its energy
error against FP64 (one extra low-precision build).
"""

import time

import numpy as np
import torch


# ---------------------------------------------------------------- helpers
def canonical_orth(S, rtol=1e-8):
    w, U = torch.linalg.eigh(S)
    if bool((w < -1e-14).any()):
        raise ValueError("The metric matrix must be positive semi-definite.")
    w = torch.clamp(w, min=0.0)
    keep = w >= rtol * w[-1]
    return U[:, keep] / torch.sqrt(w[keep])


class DeviceDIIS:
    """Pulay CDIIS on device tensors (forte2/helpers/diis.py:67-125)."""

    def __init__(self, diis_start=1, diis_nvec=8, diis_min=2, do_diis=True):
        if diis_min > diis_nvec:
            raise ValueError("diis_min must be less than or equal to diis_nvec")
        self.start, self.nvec, self.min = diis_start, diis_nvec, diis_min
        self.do = bool(do_diis) and diis_start >= 0 and diis_nvec >= 1
        self.reset()

    def reset(self):
        self.p, self.e = [], []
        self.G = None  # Gram matrix <e_i, e_j> of the stored error vectors (host FP64)
        self.iter = 0
        self.status = ""

    def update(self, p, e):
        self.status = ""
        if not self.do:
            return p
        self.iter += 1
        if self.iter < self.start:
            return p
        e = e.reshape(-1).clone()
        # one new row of the Gram matrix: dot products of the new error vector
        # with the stored ones (memory-bound reductions, one sync)
        row = torch.stack([torch.dot(e, ei) for ei in self.e] + [torch.dot(e, e)]).cpu().numpy()
        n_old = len(self.e)
        G = np.zeros((n_old + 1, n_old + 1))
        if n_old:
            G[:n_old, :n_old] = self.G
        G[n_old, :] = row
        G[:, n_old] = row
        self.p.append(p.clone())
        self.e.append(e)
        if len(self.p) > self.nvec:
            self.p.pop(0)
            self.e.pop(0)
            G = G[1:, 1:]
        self.G = G
        self.status = "S"
        n = len(self.p)
        if n < self.min:
            return p
        # the (n+1) x (n+1) CDIIS system of diis.py:104-117, normalised by max|B|
        Bm = -np.ones((n + 1, n + 1))
        Bm[-1, -1] = 0.0
        with np.errstate(invalid="ignore", divide="ignore"):  # all-zero errors: caught below
            Bm[:n, :n] = G / np.abs(G).max()
        rhs = np.zeros(n + 1)
        rhs[-1] = -1.0
        try:
            if not np.isfinite(Bm).all():
                raise np.linalg.LinAlgError
            x = np.linalg.solve(Bm, rhs)
        except np.linalg.LinAlgError:
            return p  # singular DIIS system: skip extrapolation (diis.py:115-120)
        if not np.isfinite(x).all():
            return p
        self.status += "/E"
        out = self.p[0] * float(x[0])
        for xi, pi in zip(x[1:n], self.p[1:]):
            out.add_(pi, alpha=float(xi))
        return out


def _polar(M, tol=5e-14, max_iter=24):
    """Orthogonal polar factor of a square M (the R of the orthogonal
    Procrustes problem).  M = Co^T S Cn has the cosines of the angles between
    two occupied spaces as its singular values, near 1 in an SCF's late
    cycles, where the Newton-Schulz iteration X <- X (3 I - X^T X) / 2
    converges quadratically (a 1e-2 deviation from orthogonality reaches
    roundoff in four steps) with small GEMMs only; checked every four steps.
    When it has not reached ||X^T X - I||_max < tol the SVD gives the factor
    (its latency on the device is ~40 ms at 320 x 320, the iteration's ~0.2 ms)."""
    n = M.shape[0]
    eye = torch.eye(n, dtype=M.dtype, device=M.device)
    X = M
    for k in range(max_iter):
        X = 0.5 * X @ (3.0 * eye - X.T @ X)
        if k % 4 == 3:
            err = float((X.T @ X - eye).abs().max())
            if not err < 1.0:  # diverging (a singular value beyond sqrt(3)) or NaN
                break
            if err < tol:
                return X
    U, _, Vh = torch.linalg.svd(M)
    return U @ Vh


def _procrustes_split(Cn, Co, S):
    """U+, U- with Cn Cn^T - Co Co^T = (U+ U+^T - U- U-^T) / 2."""
    if Cn.shape[1] == 0:
        return None
    M = Co.T @ S @ Cn
    Coa = Co @ _polar(M)
    P = 0.5 * (Cn + Coa)
    A = Cn - Coa
    na, npn = torch.linalg.norm(A), torch.linalg.norm(P)
    if float(na) == 0.0:
        return None
    t = float(na / npn)
    s = t ** 0.5
    return s * P + A / s, s * P - A / s


def guess_mix(C, homo, angle=np.pi / 4):
    """Break the alpha/beta symmetry of an ms = 0 guess (scf_utils.py:76-104):
    the HOMO and LUMO columns are rotated by +angle for alpha and -angle for
    beta (Szabo and Ostlund 3.8.7)."""
    c, s = float(np.cos(angle)), float(np.sin(angle))
    out = []
    for sg in (1.0, -1.0):
        M = C.clone()
        h, l = C[:, homo], C[:, homo + 1]
        M[:, homo] = c * h + sg * s * l
        M[:, homo + 1] = -sg * s * h + c * l
        out.append(M)
    return out


# ---------------------------------------------------------------- SCF base
class _SCF:
    def __init__(self, charge, do_diis=True, diis_start=1, diis_nvec=8, diis_min=2, e_tol=1e-9,
                 d_tol=1e-6, maxiter=100, guess_type="minao", die_if_not_converged=True,
                 precision="fp64", switch_inc=3e-3, switch_fp64=1e-6, incremental=True,
                 fp64_tail="reref", inc_max_tail=12, tail_refs=1, tail_e_tol=1e-8, lp_error_check=False,
                 device=None, log=None, level_shift=None, level_shift_thresh=1e-5):
        self.charge = charge
        self.level_shift, self.level_shift_thresh = level_shift, level_shift_thresh
        self.do_diis, self.diis_start, self.diis_nvec, self.diis_min = do_diis, diis_start, diis_nvec, diis_min
        self.e_tol, self.d_tol, self.maxiter = e_tol, d_tol, maxiter
        self.guess_type = guess_type
        self.die_if_not_converged = die_if_not_converged
        if precision not in ("fp64", "fp32", "bf16x3"):
            raise ValueError(f"unknown precision {precision!r}")
        self.precision = precision
        self.switch_inc, self.switch_fp64, self.incremental = switch_inc, switch_fp64, incremental
        if fp64_tail not in ("reref", "full"):
            raise ValueError(f"unknown fp64_tail {fp64_tail!r}")
        self.fp64_tail, self.inc_max_tail, self.tail_e_tol = fp64_tail, int(inc_max_tail), float(tail_e_tol)
        self.tail_refs = int(tail_refs)
        self.lp_error_check = bool(lp_error_check)
        self.device = device
        self.log = log
        self.converged = False
        self.executed = False
        self.V_sap = None  # SAP guess potential, computed from the system when not given
        # per-phase wall times (s) when `profile` is set: J/K builds, the
        # diagonalisation and density, the rest (DIIS, gradient, energy, host
        # syncs); each phase is closed by a device synchronisation
        self.profile = False
        self.phase_seconds = None

    def __call__(self, system):
        self.system = system
        self.nel = system.Zsum - self.charge
        if self.nel < 0:
            raise AssertionError("Number of electrons must be non-negative.")
        self._parse_state()
        self.C = None
        self._validate_level_shift()
        return self

    def _validate_level_shift(self):
        """The checks of scf_base.py:101-111: non-negative; a pair only for
        UHF, where one number is used for both spins."""
        ls = self.level_shift
        if ls is None:
            return
        if isinstance(ls, tuple):
            if self.method != "UHF":
                raise ValueError("Tuple level_shift is only valid for UHF.")
            if len(ls) != 2:
                raise ValueError("Tuple level_shift must have length 2 for UHF.")
            return
        if ls < 0.0:
            raise ValueError("level_shift must be non-negative.")
        if self.method == "UHF":
            self.level_shift = (float(ls), float(ls))

    def _apply_level_shift(self, F):
        """F_s + shift_s (S - S D_s S) on the current densities (rhf.py:90-95,
        uhf.py:150-159); shifts under 1e-4 leave F as it is."""
        cur = self._current_level_shift
        if cur is None:
            return F
        shifts = cur if isinstance(cur, tuple) else (cur,) * len(F)
        if all(x < 1e-4 for x in shifts):
            return F
        return [f + x * (self.S - self.S @ d @ self.S) for x, f, d in zip(shifts, F, self.D)]

    # ------------------------------------------------------------------ setup
    def _setup(self):
        sysm = self.system
        fb = sysm.fock_builder
        # duck-typed, as forte2's own SCF (scf_base.py:143): build_JK/J/K and a device
        if not all(hasattr(fb, a) for a in ("build_JK", "build_J", "build_K", "device")):
            raise TypeError("system.fock_builder lacks build_JK / build_J / build_K / device")
        self.fb = fb
        dev = fb.device
        self.dev = dev
        f64 = torch.float64
        self.S = torch.as_tensor(sysm.ints_overlap(), dtype=f64).to(dev)
        self.H = torch.as_tensor(sysm.ints_hcore(), dtype=f64).to(dev)
        self.Vnn = float(sysm.nuclear_repulsion)
        self.X = canonical_orth(self.S, getattr(sysm, "overlap_ortho_rtol", 1e-8))
        self._XT = None
        self.nbf = sysm.nbf
        self.nmo = int(self.X.shape[1])
        self.naux = sysm.naux

    def _eigh(self, F):
        e, c = torch.linalg.eigh(self.X.T @ F @ self.X)
        return e, self.X @ c

    def _eigh_batch(self, F):
        """eigh of X^T F_s X for a (nspin, N, N) stack in one batched call: the
        library's eigensolver is latency-bound at these sizes, so the pair
        costs about one single call (both operands of every product are
        contiguous)."""
        if self._XT is None:
            self._XT = self.X.T.contiguous()
        e, c = torch.linalg.eigh(torch.matmul(self._XT, torch.matmul(F, self.X)))
        return e, torch.matmul(self.X, c)

    def _guess(self):
        if self.guess_type == "minao":
            Vh = self.V_sap if self.V_sap is not None else self.system.sap_potential()
            V = torch.as_tensor(Vh, dtype=torch.float64).to(self.dev)
            return self._eigh(self.H + V)[1]
        if self.guess_type == "hcore":
            return self._eigh(self.H)[1]
        raise RuntimeError(f"Unknown initial guess type: {self.guess_type}")

    # ------------------------------------------------------------------ precision schedule
    # The SCF loop is a generator: it yields the J/K builds it needs as
    # ("JK", C_occ list, precision), ("J", D list, precision) or
    # ("K", C list, precision) and receives their results.  run() executes
    # every request at once; the serving layer batches the FP64 "JK"
    # requests of many jobs into grouped launches.
    def _fock_gen(self, stage):
        """Yields the builds of the Fock matrices of `stage`; returns them."""
        Cocc = self._occ()
        if stage in ("fp64", "lp", "ref"):
            prec = self.precision if stage == "lp" else "fp64"
            J, K = yield ("JK", Cocc, prec)
            F = self._assemble(J, K)
            if stage == "ref":
                self._ref = ([c.clone() for c in Cocc], [d.clone() for d in self.D], [f.clone() for f in F])
            return F
        # incremental: F = F_ref + G_lp(D - D_ref)
        Cref, Dref, Fref = self._ref
        dD = [d - dr for d, dr in zip(self.D, Dref)]
        Jd = (yield ("J", [sum(dD)] if len(dD) > 1 else dD, self.precision))[0]
        Kd = []
        for cn, co in zip(Cocc, Cref):
            split = self._timed("split", _procrustes_split, cn, co, self.S)
            if split is None:
                Kd.append(torch.zeros_like(Jd))
            else:
                Kp, Km = yield ("K", list(split), self.precision)
                Kd.append(0.5 * (Kp - Km))
        return self._assemble_delta(Fref, Jd, Kd)

    def _execute(self, req):
        kind, args, prec = req
        if kind == "JK":
            return self.fb.build_JK(args, precision=prec)
        if kind == "J":
            return self.fb.build_J(args, precision=prec)
        return self.fb.build_K(args, precision=prec)

    def _timed(self, phase, fn, *args):
        if not self.profile:
            return fn(*args)
        torch.cuda.synchronize(self.dev)
        t0 = time.perf_counter()
        out = fn(*args)
        torch.cuda.synchronize(self.dev)
        self.phase_seconds[phase] += time.perf_counter() - t0
        return out

    def run(self):
        self.phase_seconds = {"build": 0.0, "diag": 0.0, "split": 0.0} if self.profile else None
        gen = self.iterate()
        try:
            req = next(gen)
            while True:
                req = gen.send(self._timed("build", self._execute, req))
        except StopIteration:
            pass
        if self.profile:
            self.phase_seconds["other"] = self.time - sum(self.phase_seconds.values())
        return self

    def iterate(self):
        """The SCF loop of scf_base.py:121-251 as a generator of J/K build requests."""
        t0 = time.monotonic()
        self._setup()
        diis = DeviceDIIS(self.diis_start, self.diis_nvec, self.diis_min, self.do_diis)
        self.iter = 0
        self.history = []
        self._last_dE = float("inf")
        stage = "fp64" if self.precision == "fp64" else "lp"
        if self.C is None:
            self.C = self._guess_orbitals()
        self.D = self._density()
        F = yield from self._fock_gen(stage)
        self.E = self.Vnn + self._energy(F)
        Eold, Dold = self.E, self.D
        self.iter += 1
        self.history.append((self.iter, stage, self.E))
        self.lp_error_vs_fp64 = None
        self._current_level_shift = self.level_shift
        # in the converging tail; its incremental cycles since its last
        # reference; the references taken in it
        tail, tail_inc, tail_ref_count = False, 0, 0
        for _ in range(self.maxiter):
            grad = self._ao_grad(F)
            Fx = self._apply_level_shift(self._diis_update(diis, F, grad))
            self.eps, self.C = self._timed("diag", self._diagonalize, Fx)
            self.D = self._density()
            # stage transitions of the precision form (decided on the previous dE)
            next_stage = "inc" if stage == "ref" else stage
            if stage == "lp" and (abs(self._last_dE) < self.switch_inc
                                  or 0.0 < self._last_dE < 10.0 * self.switch_inc):
                next_stage = "ref" if self.incremental else "fp64"
            elif next_stage == "inc" and not tail and abs(self._last_dE) < self.switch_fp64:
                tail = True
                if self.fp64_tail == "full":
                    next_stage = "fp64"
                elif stage != "ref":
                    next_stage = "ref"
                else:  # a reference taken just now is the tail's reference
                    tail_ref_count = 1
            elif next_stage == "inc" and tail:
                tail_inc += 1
                if tail_inc > self.inc_max_tail:
                    next_stage = "ref" if tail_ref_count < self.tail_refs else "fp64"
            if next_stage == "ref" and tail:
                tail_ref_count += 1
                tail_inc = 0
            if stage == "lp" and next_stage == "fp64":
                diis.reset()  # low-precision Fock matrices leave the DIIS subspace
            F = yield from self._fock_gen(next_stage)
            self.E = self.Vnn + self._energy(F)
            if next_stage == "ref" and not tail:
                # first reference: the DIIS subspace restarts from
                # FP64-accurate Fock matrices (a later re-reference keeps it)
                if self.lp_error_check:
                    # the low-precision build's energy error at this density
                    Flp = yield from self._fock_gen("lp")
                    self.lp_error_vs_fp64 = abs((self.Vnn + self._energy(Flp)) - self.E)
                diis.reset()
            stage = next_stage
            dE = self.E - Eold
            self._last_dE = dE
            if abs(dE) < self.level_shift_thresh:  # scf_base.py:203-204
                self._current_level_shift = None
            dD = float(sum(torch.linalg.norm(d - do) for d, do in zip(self.D, Dold)))  # one host sync
            self.history.append((self.iter + 1, stage, self.E))
            if self.log:
                self.log(f"{self.iter + 1:4d} {self.E:20.12f} {dE:12.4e} {dD:12.4e} {stage:>5s} {diis.status}")
            # convergence is tested on FP64 builds and on the tail's cycles
            # (incremental against the tail's reference, or that reference)
            e_tol = max(self.e_tol, self.tail_e_tol) if stage != "fp64" else self.e_tol
            if (stage == "fp64" or tail) and abs(dE) < e_tol and dD < self.d_tol:
                # final iteration: diagonalise the new, non-extrapolated Fock
                # matrix (scf_base.py:214-221; F_canon was rebuilt at :196)
                self.eps, self.C = self._diagonalize(F)
                self.D = self._density()
                F = yield from self._fock_gen("fp64")
                self.E = self.Vnn + self._energy(F)
                self.converged = True
                break
            Eold, Dold = self.E, self.D
            self.iter += 1
        else:
            if self.die_if_not_converged:
                raise RuntimeError(f"{self.method} did not converge in {self.maxiter} iterations.")
        self.F = F
        self.S2 = self._spin()
        self.time = time.monotonic() - t0
        self.executed = True
        return self



class RHF(_SCF):
    """Restricted Hartree-Fock (forte2/scf/rhf.py:13-65)."""

    method = "RHF"

    def _parse_state(self):
        if self.nel % 2:
            raise AssertionError("RHF requires an even number of electrons.")
        self.na = self.nb = self.nel // 2
        self.ms = 0

    def _guess_orbitals(self):
        return [self._guess()]

    def _occ(self):
        return [self.C[0][:, : self.na]]

    def _density(self):
        c = self.C[0][:, : self.na]
        return [c @ c.T]

    def _assemble(self, J, K):
        return [self.H + 2.0 * J[0] - K[0]]

    def _assemble_delta(self, Fref, Jd, Kd):
        return [Fref[0] + 2.0 * Jd - Kd[0]]

    def _energy(self, F):
        return float(torch.sum(self.D[0] * (self.H + F[0])))

    def _ao_grad(self, F):
        f, d = F[0], self.D[0]
        return self.X.T @ (f @ d @ self.S - self.S @ d @ f) @ self.X

    def _diis_update(self, diis, F, grad):
        return [diis.update(F[0], grad)]

    def _diagonalize(self, F):
        e, c = self._eigh(F[0])
        return [e], [c]

    def _spin(self):
        return 0.0


class UHF(_SCF):
    """Unrestricted Hartree-Fock (forte2/scf/uhf.py:14-127)."""

    method = "UHF"

    def __init__(self, charge, ms=None, guess_mix=False, **kw):
        super().__init__(charge, **kw)
        self.ms = ms
        self.guess_mix = guess_mix

    def _parse_state(self):
        if self.ms is None:
            raise ValueError("Spin projection (ms) must be set for UHF.")
        tw = int(round(self.ms * 2))
        if abs(tw - self.ms * 2) > 1e-12:
            raise AssertionError("ms must be a multiple of 0.5.")
        if self.nel % 2 != tw % 2:
            raise ValueError(f"{self.nel} electrons is incompatible with ms={self.ms}!")
        self.na, self.nb = (self.nel + tw) // 2, (self.nel - tw) // 2
        self.twicems = tw
        if self.na < 0 or self.nb < 0:
            raise AssertionError("UHF requires non-negative numbers of alpha and beta electrons.")

    def _guess_orbitals(self):
        C = self._guess()
        if self.twicems == 0 and self.guess_mix:
            return guess_mix(C, self.nel // 2 - 1)
        return [C, C.clone()]

    def _occ(self):
        return [self.C[0][:, : self.na], self.C[1][:, : self.nb]]

    def _density(self):
        a, b = self._occ()
        return [a @ a.T, b @ b.T]

    def _assemble(self, J, K):
        Jt = J[0] + J[1]
        return [self.H + Jt - k for k in K]

    def _assemble_delta(self, Fref, Jd, Kd):
        return [f + Jd - k for f, k in zip(Fref, Kd)]

    def _energy(self, F):
        Da, Db = self.D
        return float(0.5 * (torch.sum((Da + Db) * self.H) + torch.sum(Da * F[0]) + torch.sum(Db * F[1])))

    def _ao_grad(self, F):
        return torch.cat([(self.X.T @ (f @ d @ self.S - self.S @ d @ f) @ self.X).reshape(-1)
                          for d, f in zip(self.D, F)])

    def _diis_update(self, diis, F, grad):
        N = self.nbf
        flat = diis.update(torch.cat([f.reshape(-1) for f in F]), grad)
        return [flat[: N * N].reshape(N, N), flat[N * N:].reshape(N, N)]

    def _diagonalize(self, F):
        # both spins in one batched eigh (scf_base.py:113-116 per spin)
        e, c = self._eigh_batch(torch.stack(F))
        return [e[0], e[1]], [c[0], c[1]]

    def _spin(self):
        a, b = self._occ()
        Sij = a.T @ self.S @ b
        return float(self.ms * (self.ms + 1) + self.nb - torch.sum(Sij * Sij))
