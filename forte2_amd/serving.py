"""Continuous-batching request layer for streams of SCF jobs on one GPU.

A request is an RHF (or UHF) job: a geometry, basis sets and charge/spin.  The
server keeps each admitted job's B, density, Fock and DIIS history on the
device.  It drives all active jobs in lock-step ticks.  At each tick, every
job that needs a Fock matrix contributes its occupied orbitals to one grouped
J/K build (``f2jk_group_build_jk``: one persistent launch per phase for up to
12 jobs, in the job's precision: FP64, or the FP32 / BF16x3 precision form).
Then every job advances one SCF step on the device.

Admission is continuous:
  * a request becomes eligible at its arrival time;
  * it is admitted, in arrival order, as soon as a slot is free and its HBM
    footprint (B = naux*nbf^2*8 bytes, plus the J/K workspace) fits the
    budget left by the jobs in flight;
  * its one-time host work (integrals, (Q|mn), SAP guess potential) runs in a
    worker thread pool while the GPU serves the other jobs;
  * its B is formed on the device when it joins the next tick.

A job's J/K do not depend on what else is in the group (the grouped kernels
reuse the single-job decomposition), so a result is bitwise the same alone
and in any mix.

Per request the server records arrival, admission, first tick and completion
times.  latency = completion - arrival.  ``summary()`` reports p50/p99
latency and jobs/s.
"""

import concurrent.futures as cf
import ctypes
import dataclasses
import time

import numpy as np
import torch

from . import _native as nat
from . import autotune
from . import df as _df
from .jkbuilder import FockBuilder
from .scf import RHF, UHF
from .system import System

MAX_GROUP = 12


@dataclasses.dataclass
class SCFRequest:
    id: int
    xyz: str
    basis: str = "cc-pvdz"
    aux: str = "cc-pvtz-jkfit"
    charge: int = 0
    ms: float = None           # None -> RHF, else UHF with this ms
    arrival: float = 0.0       # seconds after the start of the run


@dataclasses.dataclass
class SCFResult:
    id: int
    E: float = None
    iterations: int = 0
    converged: bool = False
    nbf: int = 0
    naux: int = 0
    arrival: float = 0.0
    admitted: float = 0.0
    started: float = 0.0
    finished: float = 0.0
    error: str = None

    @property
    def latency(self):
        return self.finished - self.arrival


def job_footprint(nbf, naux, nocc_total, precision="fp64"):
    """HBM bytes a job holds: B (plus its FP32 copy in the precision forms) and
    an upper bound on its J/K workspace."""
    N, Q = nbf, naux
    opad = 32 * (int(nocc_total) // 16 + 2)
    B = (8 if precision == "fp64" else 12) * Q * N * N
    ws = 8 * Q * N * opad + 8 * N * N * (4 + 80) + 8 * 2048 * 4 * 1024 + (64 << 20)
    return B + ws


class _Job:
    """One SCF run driven one J/K build at a time.

    The SCF loop itself is ``RHF/UHF.iterate()``: a generator that yields its
    build requests ("JK" / "J" / "K", arguments, precision) and receives
    their results, so the served job runs exactly the loop of ``run()``.
    """

    def __init__(self, req, system, scf, fb):
        self.req, self.system, self.scf, self.fb = req, system, scf, fb
        self.error = None
        self.gen = scf.iterate()
        self.pending = next(self.gen)  # set-up, guess and the first build request

    def send(self, result):
        """Deliver the result of the pending request; True when the job has finished."""
        try:
            self.pending = self.gen.send(result)
            return False
        except StopIteration:
            self.pending = None
            return True


class SCFServer:
    """Serve a stream of SCF requests on one GPU with continuous batching.

    Parameters
    ----------
    device : torch device (default cuda:0)
    hbm_budget : bytes of HBM the jobs in flight may hold (default: 85% of free)
    max_active : jobs in flight at once (slots)
    prep_threads : host threads computing the one-time integrals of admitted jobs
    grouped : batch the J/K builds of a tick into grouped launches
    e_tol, d_tol, maxiter : SCF convergence (forte2 semantics)
    precision : "fp64", or a precision form ("bf16x3", "fp32") for every job
    """

    def __init__(self, device=None, hbm_budget=None, max_active=12, prep_threads=4, grouped=True,
                 e_tol=1e-8, d_tol=1e-6, maxiter=100, precision="fp64", log=None):
        self.device = torch.device(device) if device is not None else torch.device("cuda", 0)
        if hbm_budget is None:
            # 85% of the free HBM, less room for the transient of forming one
            # job's B on the device (the metric factor and the solve blocks)
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.85 * free) - (4 << 30)
        self.hbm_budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.prep_threads = int(prep_threads)
        self.grouped = grouped
        self.e_tol, self.d_tol, self.maxiter = e_tol, d_tol, maxiter
        self.precision = precision
        self.log = log
        self.results = {}
        self.ticks = 0
        self.wall = 0.0
        self.group_sizes = []
        # grouped builds the library declined (ValueError / MemoryError /
        # out of HBM: the jobs continue alone), grouped builds that failed
        # otherwise (a HIP launch or runtime error, or a bug: the jobs of the
        # group end with the error recorded) and single-job builds that failed
        # in the library or the HIP runtime
        self.group_declines = 0
        self.group_failures = 0
        self.build_failures = 0

    # ------------------------------------------------------------------ host preparation
    @staticmethod
    def _prepare_host(req):
        """Host-only work of a request (runs in a worker thread)."""
        system = System(req.xyz, req.basis, req.aux)
        system.ints_hcore()  # cached on the system: no host integrals on the serving thread
        M, T = _df.host_df_integrals(system)
        V_sap = system.sap_potential()
        return system, M, T, V_sap

    def _activate(self, req, prepared):
        system, M, T, V_sap = prepared
        system.device = self.device
        B = _df.device_B_from_host(M, T, self.device, system.nbf, getattr(system, "df_ortho_rtol", None))
        # J over all elements (j_sparse off): a job's J must not depend on
        # whether a step ran grouped (the group reads every element) or alone,
        # and the element scan would synchronise the device at admission
        fb = FockBuilder(system, B=B, device=self.device, j_sparse=False)
        # launch shape of the job's size class (tuned for grouped builds when
        # batching); the X chunk is pinned so that the job's builds that run
        # alone keep the plan the grouped path was asked about
        autotune.apply(fb, self.precision, grouped=self.grouped)
        if self.grouped:
            fb.set_tuning("xchunk_bytes", 1 << 32)
            # its builds that run alone contract in the grouped tile (128), so
            # a job's K does not depend on which path a step took
            fb.set_tuning("ks_tile", 128)
        system.fock_builder = fb
        kw = dict(e_tol=self.e_tol, d_tol=self.d_tol, maxiter=self.maxiter, precision=self.precision)
        scf = RHF(req.charge, **kw) if req.ms is None else UHF(req.charge, ms=req.ms, **kw)
        scf(system)
        scf.V_sap = V_sap
        return _Job(req, system, scf, fb)

    # ------------------------------------------------------------------ grouped J/K
    def _groupable(self, jb, kind, prec, nd):
        """Whether the library takes this job into a grouped build (asked once per
        kind/precision/density count: odd nbf, or an X larger than one chunk, run
        the job on its own)."""
        key = (kind, prec, nd)
        memo = jb.__dict__.setdefault("_groupable", {})
        if key not in memo:
            lib = nat.jk_lib()
            occ = jb.pending[1]
            if prec != "fp64":
                jb.fb.B_fp32()
            flags = nat.F2JK_WANT_K | (nat.F2JK_WANT_J if kind == "JK" else 0)
            sizes = (ctypes.c_size_t * 1)()
            rc = lib.f2jk_group_workspace_size(nat.ptr_array([jb.fb.ctx.value]), 1, (ctypes.c_int * 1)(nd),
                                               nat.i64_array([c.shape[1] for c in occ]), nat.PREC[prec],
                                               flags, sizes)
            memo[key] = rc == 0
        return memo[key]

    def _build_group(self, jobs):
        """The pending "JK" (or "K") builds of `jobs` (same kind, precision and
        density count) in grouped launches; returns per job (J list, K list), or
        the K list for "K"."""
        lib = nat.jk_lib()
        kind, _, prec = jobs[0].pending
        want_j = kind == "JK"
        flags = nat.F2JK_WANT_K | (nat.F2JK_WANT_J if want_j else 0)
        ctxs, nds, Cs, nocc, ldc, Js, Ks = [], [], [], [], [], [], []
        outs = []
        for jb in jobs:
            occ = [c.contiguous() for c in jb.pending[1]]
            N = jb.fb.nbf
            J = [torch.empty((N, N), dtype=torch.float64, device=self.device) for _ in occ] if want_j else []
            K = [torch.empty((N, N), dtype=torch.float64, device=self.device) for _ in occ]
            if prec != "fp64":
                jb.fb.B_fp32()  # the FP32 copy the precision forms read
            jb._keep = occ
            ctxs.append(jb.fb.ctx.value)
            nds.append(len(occ))
            for d, c in enumerate(occ):
                Cs.append(c.data_ptr())
                nocc.append(c.shape[1])
                ldc.append(max(1, c.stride(0)))
                Js.append(J[d].data_ptr() if want_j else 0)
                Ks.append(K[d].data_ptr())
            outs.append((J, K) if want_j else K)
        n = len(jobs)
        ctx_a = nat.ptr_array(ctxs)
        nd_a = (ctypes.c_int * n)(*nds)
        nocc_a = nat.i64_array(nocc)
        sizes = (ctypes.c_size_t * n)()
        rc = lib.f2jk_group_workspace_size(ctx_a, n, nd_a, nocc_a, nat.PREC[prec], flags, sizes)
        nat.jk_check(rc, "f2jk_group_workspace_size")
        ws = [jb.fb._workspace(sizes[i]) for i, jb in enumerate(jobs)]
        ws_a = nat.ptr_array([w.data_ptr() for w in ws])
        wsb = (ctypes.c_size_t * n)(*[w.numel() for w in ws])
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        rc = lib.f2jk_group_build_jk(ctx_a, n, nd_a, nat.ptr_array(Cs), nocc_a, nat.i64_array(ldc),
                                     nat.ptr_array(Js), nat.ptr_array(Ks), nat.PREC[prec], flags,
                                     ws_a, wsb, stream)
        nat.jk_check(rc, "f2jk_group_build_jk")
        return outs

    def _tick(self, active):
        """Serve the pending build of every active job, then let each take its next step.

        "JK" builds (J and K from C) and "K" builds (the factored difference
        densities of the incremental precision form) are batched into grouped
        launches by (kind, precision, density count); J builds from a
        difference density, and jobs the grouped path does not take, run on
        their own.
        """
        done, results, solo, failed = [], {}, [], set()
        groups = {}
        for jb in active:
            kind, args, prec = jb.pending
            nd = len(args)
            if self.grouped and kind in ("JK", "K") and nd <= 2 and self._groupable(jb, kind, prec, nd):
                groups.setdefault((kind, prec, nd), []).append(jb)
            else:
                solo.append(jb)
        for (kind, prec, nd), jobs in groups.items():
            for g0 in range(0, len(jobs), MAX_GROUP):
                grp = jobs[g0:g0 + MAX_GROUP]
                try:
                    outs = self._build_group(grp)
                except (ValueError, MemoryError, torch.OutOfMemoryError):
                    # the library declines the group (shape it does not take,
                    # workspace it cannot get): these jobs run alone from now on
                    self.group_declines += 1
                    for jb in grp:
                        jb._groupable[(kind, prec, nd)] = False
                    solo.extend(grp)
                    continue
                except Exception as e:
                    # a launch or runtime failure (NativeError) or a fault of
                    # the group path itself: the work of these jobs is lost, so
                    # they end with the error; nothing is retried alone
                    self.group_failures += 1
                    for jb in grp:
                        jb.error = repr(e)
                        failed.add(id(jb))
                    continue
                for jb, out in zip(grp, outs):
                    results[id(jb)] = out
                self.group_sizes.append(len(grp))
        for jb in solo:
            try:
                results[id(jb)] = jb.scf._execute(jb.pending)
            except nat.NativeError as e:  # a launch or runtime failure: this job ends, counted
                self.build_failures += 1
                jb.error = repr(e)
                failed.add(id(jb))
                continue
            except Exception as e:  # this job's build failed: it ends, the others go on
                jb.error = repr(e)
                failed.add(id(jb))
                continue
            self.group_sizes.append(1)
        for jb in active:
            if id(jb) in failed:
                done.append(jb)
                continue
            try:
                if jb.send(results[id(jb)]):
                    done.append(jb)
            except RuntimeError as e:  # non-convergence: this job fails, the others go on
                jb.error = str(e)
                done.append(jb)
            except Exception as e:  # any other failure of this job's step: likewise
                jb.error = repr(e)
                done.append(jb)
        self.ticks += 1
        return done

    def _retire(self, jb):
        """Release a finished job's device memory at once (its B, workspace and
        SCF state): System and FockBuilder refer to each other, so dropping the
        job would leave them to the cycle collector."""
        if jb.fb is not None:
            jb.fb.release()
        if jb.system is not None:
            jb.system.fock_builder = None
        jb.gen = jb.scf = jb.system = jb.fb = jb.pending = None
        jb.__dict__.pop("_keep", None)
        torch.cuda.empty_cache()

    # ------------------------------------------------------------------ main loop
    def run(self, requests):
        """Serve `requests` (arrival times in seconds from now); returns {id: SCFResult}."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.id))
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        queue, preparing, active = [], {}, []
        ready = []  # (req, fp, prepared) whose activation ran out of HBM: retried as jobs finish
        used = 0
        pool = cf.ThreadPoolExecutor(max_workers=self.prep_threads)
        pi = 0
        try:
            while pi < len(pending) or queue or preparing or active or ready:
                t = now()
                while pi < len(pending) and pending[pi].arrival <= t:
                    queue.append(pending[pi])
                    pi += 1
                # admission in arrival order: slot and HBM budget
                while queue and len(active) + len(preparing) < self.max_active:
                    req = queue[0]
                    try:
                        fp = self._estimate(req)
                    except Exception as e:  # unreadable request (geometry, basis name): this job only
                        queue.pop(0)
                        self.results[req.id] = SCFResult(req.id, arrival=req.arrival, admitted=now(),
                                                         finished=now(), error=repr(e))
                        continue
                    if fp > self.hbm_budget:  # can never fit: rejected, the stream goes on
                        queue.pop(0)
                        self.results[req.id] = SCFResult(req.id, arrival=req.arrival, admitted=now(),
                                                         finished=now(),
                                                         error=f"footprint {fp} B exceeds the HBM budget")
                        continue
                    if used + fp > self.hbm_budget:
                        break
                    queue.pop(0)
                    used += fp
                    res = SCFResult(req.id, arrival=req.arrival, admitted=now())
                    self.results[req.id] = res
                    preparing[pool.submit(self._prepare_host, req)] = (req, fp)
                # jobs whose host work is done join the next tick
                done_prep = []
                for fut in [f for f in preparing if f.done()]:
                    req, fp = preparing.pop(fut)
                    try:
                        done_prep.append((req, fp, fut.result()))
                    except Exception as e:  # host preparation failed: this job only
                        res = self.results[req.id]
                        res.error, res.finished = repr(e), now()
                        used -= fp
                retry, ready = ready, []
                for req, fp, prepared in retry + done_prep:
                    res = self.results[req.id]
                    try:
                        job = self._activate(req, prepared)
                        job.footprint = fp
                        job.error = None
                        res.started = now()
                        res.nbf, res.naux = job.fb.nbf, job.fb.naux
                        active.append(job)
                    except torch.OutOfMemoryError as e:
                        torch.cuda.empty_cache()
                        if active:  # memory frees up as jobs finish: try again then
                            ready.append((req, fp, prepared))
                        else:
                            res.error, res.finished = repr(e), now()
                            used -= fp
                    except Exception as e:  # failed job: others unaffected
                        res.error = repr(e)
                        res.finished = now()
                        used -= fp
                if active:
                    for jb in self._tick(active):
                        active.remove(jb)
                        res = self.results[jb.req.id]
                        res.E, res.iterations = getattr(jb.scf, "E", None), getattr(jb.scf, "iter", 0)
                        res.converged, res.error = bool(jb.scf.converged), jb.error
                        res.finished = now()
                        used -= jb.footprint
                        self._retire(jb)
                    if self.log:
                        self.log(f"t={now():7.2f}s active={len(active)} preparing={len(preparing)} "
                                 f"queued={len(queue)} done={sum(1 for r in self.results.values() if r.finished)}")
                elif preparing:
                    cf.wait(list(preparing), timeout=0.05, return_when=cf.FIRST_COMPLETED)
                elif pi < len(pending):
                    time.sleep(max(0.0, min(0.05, pending[pi].arrival - now())))
        finally:
            pool.shutdown(wait=True)
        self.wall = now()
        return self.results

    def _estimate(self, req):
        from . import basis as _basis
        from . import geometry as _geom

        atoms = _geom.parse_geometry(req.xyz)
        nbf = _basis.build_basis(req.basis, atoms).size
        naux = _basis.build_basis(req.aux, atoms).size
        nel = sum(z for z, _ in atoms) - req.charge
        return job_footprint(nbf, naux, nel, self.precision)

    def summary(self):
        ok = [r for r in self.results.values() if r.error is None and r.converged]
        lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
        # where a request's latency goes: waiting for a slot or HBM (arrival ->
        # admission), host integrals and B formation (-> start on the device),
        # the SCF on the device (-> completion)
        parts = {
            "queue": np.array([r.admitted - r.arrival for r in ok]) if ok else np.zeros(1),
            "prep": np.array([r.started - r.admitted for r in ok]) if ok else np.zeros(1),
            "device": np.array([r.finished - r.started for r in ok]) if ok else np.zeros(1),
        }
        split = {f"{k}_{stat}_s": float(f(v)) for k, v in parts.items()
                 for stat, f in (("mean", np.mean), ("p99", lambda x: np.percentile(x, 99)))}
        return {
            "requests": len(self.results),
            "completed": len(ok),
            "failed": len(self.results) - len(ok),
            "wall_s": self.wall,
            "jobs_per_s": len(ok) / self.wall if self.wall > 0 else 0.0,
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_mean_s": float(lat.mean()),
            "ticks": self.ticks,
            "mean_group_size": float(np.mean(self.group_sizes)) if self.group_sizes else 0.0,
            "group_declines": self.group_declines,
            "group_failures": self.group_failures,
            "build_failures": self.build_failures,
            **split,
        }


def serving_stream(nreq=512, rate=4.0, nmax=24, seed=7, basis="cc-pvdz", aux="cc-pvtz-jkfit"):
    """The serving workload: (H2O)_n with n ~ U{1..nmax}, seeded geometries in a
    cube of side 3.6 n^(1/3) A, Poisson arrivals at `rate` per second."""
    from . import geometry as _geom

    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(nreq):
        n = int(rng.integers(1, nmax + 1))
        xyz = _geom.water_cluster(n, seed=int(rng.integers(0, 2 ** 31)))
        reqs.append(SCFRequest(i, xyz, basis, aux, arrival=t))
        t += float(rng.exponential(1.0 / rate))
    return reqs


def alkane_stream(nreq=48, rate=0.1, nmin=2, nmax=20, cation_fraction=0.5, seed=11,
                  basis="cc-pvtz", aux="cc-pvtz-jkfit"):
    """The mixed-size UHF workload: all-trans C_nH_(2n+2) with n ~ U{nmin..nmax},
    each either neutral (UHF, ms = 0) or the cation (UHF doublet, ms = 1/2) with
    probability `cation_fraction`; two densities per J/K build.  Poisson
    arrivals at `rate` per second."""
    from . import geometry as _geom

    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(nreq):
        n = int(rng.integers(nmin, nmax + 1))
        cation = bool(rng.random() < cation_fraction)
        reqs.append(SCFRequest(i, _geom.alkane(n), basis, aux, charge=1 if cation else 0,
                               ms=0.5 if cation else 0.0, arrival=t))
        t += float(rng.exponential(1.0 / rate))
    return reqs
