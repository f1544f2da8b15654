"""Benchmark of the density-fitted J/K build on one MI355X.

Workload: (H2O)_64 with the cc-pVDZ orbital basis and cc-pVTZ-JKFIT auxiliary
basis shapes (nbf N=1536, naux Q=8896, nocc o=320; the cc-pVTZ-JKFIT fit set
stands in for the unshipped cc-pVDZ-JKFIT).  B(P|mn) (168 GB FP64 + 84 GB FP32
copy) and the occupied coefficients are synthetic and already resident in HBM
when the timed region starts.  One step = one RHF J/K build (D = C C^T, J, K).

The headline value is the builds/s of the precision form (``--mixed``,
default bf16x3); the FP64 plain form is timed on the same inputs, and the line
reports the speed-up and the max-abs J/K difference between the two.

Printed: one JSON line with metric/value/unit, the roofline of the dominant
kernel (K half-transform, measured with HIP events on the launch stream), the
achievable MFMA and HBM-read ceilings measured on the device by peak probes
beside the published ones, and the CPU baseline (the numpy oracle on a bounded
aux-slice sample, scaled).  Two more legs of the metric run outside the
timed region:
  * accuracy, before the synthetic B is made (so that its B formation runs on
    device memory nothing has used yet): the real (H2O)_64 RHF (seed 64,
    14.5 A cube) in FP64, in the precision form, and in the precision form
    with FP64 cycles in its tail for comparison;
    ``final_energy_error_vs_fp64_Eh`` (bar: 1e-8 Eh) and the SCF speed-up
    (``--scf`` runs this leg alone; ``--scf-alkane N`` the C_nH_(2n+2)+ UHF
    in cc-pVTZ instead);
  * latency, after the synthetic B has left HBM: the first 48 requests of the
    seeded (H2O)_n serving stream
    (n ~ U{1..24}) at 2 requests/s through the continuous-batching server;
    ``latency_p50_s`` / ``latency_p99_s`` and jobs/s.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "JK builds/s + MFMA util, (H2O)_64 mixed vs FP64 at 1e-8 Eh; p99 SCF latency"
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}
HBM_PEAK_GBS = 8000.0
# HBM bytes per kernel launch measured by `bench.py --measure-traffic` (rocprofv3
# FETCH_SIZE / WRITE_SIZE, each scaled by what it reads for a plain device copy
# of known size), keyed by workload and precision, together with the SHA-256 of
# the kernel sources it was measured on: a table measured on other sources is
# not reported (traffic null).
TRAFFIC_TABLE = os.path.join(ROOT, "forte2_amd", "data", "traffic_gfx950.json")
KERNEL_SOURCES = ("forte2_amd/csrc/f2jk.hip", "forte2_amd/csrc/f2jk_kernels.h", "forte2_amd/csrc/f2jk_group.h")


def kernel_sources_sha256():
    import hashlib

    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def measured_traffic(waters, precision, kernel_prefix):
    """(HBM bytes per launch of the first kernel whose name starts with
    kernel_prefix, note) from the traffic table, or (None, reason)."""
    try:
        with open(TRAFFIC_TABLE) as f:
            tab = json.load(f)
    except (OSError, ValueError):
        return None, "no traffic table"
    if tab.get("kernel_sources_sha256") != kernel_sources_sha256():
        return None, "traffic table measured on other kernel sources (rerun bench.py --measure-traffic)"
    ent = tab.get("entries", {}).get(f"{waters}:{precision}", {})
    for name, k in ent.get("kernels", {}).items():
        if name.startswith(kernel_prefix):
            return k["read_bytes"] + k["write_bytes"], (
                f"{name}: {k['read_bytes']:.4g} B read + {k['write_bytes']:.4g} B written per launch "
                f"(rocprofv3 FETCH_SIZE x{tab['fetch_factor']:.3g}, WRITE_SIZE x{tab['write_factor']:.3g}, "
                f"factors from a plain {tab['copy_bytes']} B copy)")
    return None, f"no entry for {waters}:{precision}"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--waters", type=int, default=64, help="(H2O)_n cluster size")
    ap.add_argument("--mixed", default="bf16x3", choices=["fp32", "bf16x3"])
    ap.add_argument("--cpu-sample-aux", type=int, default=384,
                    help="aux functions in the CPU baseline sample (0 = skip)")
    ap.add_argument("--profile-steps", action="store_true",
                    help="run only the timed steps of one precision (for rocprofv3)")
    ap.add_argument("--profile-precision", default=None, choices=["fp64", "fp32", "bf16x3"],
                    help="precision of --profile-steps (default: --mixed)")
    ap.add_argument("--measure-traffic", action="store_true",
                    help="run rocprofv3 FETCH_SIZE / WRITE_SIZE passes over --profile-steps and a plain "
                         "copy as child processes and write the per-launch HBM traffic table")
    ap.add_argument("--traffic-out", default=None, help="output of --measure-traffic (default: the package table)")
    ap.add_argument("--set", action="append", default=[], metavar="KNOB=VALUE",
                    help="launch-shape knob of the headline builds (f2jk_set_tuning), over the tuned table")
    ap.add_argument("--scf-leg", type=int, default=1,
                    help="1: also run the real (H2O)_n RHF in FP64 and in the precision form after the "
                         "microbenchmark and report the final-energy error (0: skip)")
    ap.add_argument("--serve-leg-requests", type=int, default=48,
                    help="requests of the serving leg of the default run (0: skip)")
    ap.add_argument("--serve-leg-rate", type=float, default=2.0, help="arrival rate of the serving leg (1/s)")
    ap.add_argument("--scf", action="store_true",
                    help="instead of the J/K microbenchmark, run a real (H2O)_n RHF in FP64 and in the "
                         "precision form and report energies, errors and SCF latency")
    ap.add_argument("--seed", type=int, default=None, help="cluster seed for --scf (default n)")
    ap.add_argument("--scf-alkane", type=int, default=None,
                    help="--scf: the C_nH_(2n+2)+ UHF in cc-pVTZ/cc-pVTZ-JKFIT instead of (H2O)_n")
    ap.add_argument("--scf-profile", action="store_true",
                    help="--scf: also time the SCF phases (builds, diagonalisation, rest; device syncs added)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 4 GiB device copy (x3), the reference for HBM byte counters")
    ap.add_argument("--peaks", action="store_true",
                    help="run only the MFMA and HBM-read peak probes and print their JSON")
    ap.add_argument("--otf", type=int, default=None, metavar="N",
                    help="time the integral-direct builder on (H2O)_N instead (cc-pVDZ/cc-pVTZ-JKFIT)")
    ap.add_argument("--otf-mem", type=float, default=4000.0, help="jk_mem_thres_mb of the integral-direct builder")
    ap.add_argument("--otf-cache", type=float, default=16000.0,
                    help="MB of (P|mn) blocks kept in HBM for the cached leg of --otf")
    ap.add_argument("--serve", choices=["water", "alkane"], default=None,
                    help="instead of the J/K microbenchmark, serve a stream of SCF requests and report "
                         "p50/p99 latency and jobs/s: (H2O)_n RHF cc-pVDZ, or C_nH_(2n+2) (+ cations) UHF cc-pVTZ")
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--serve-rate", type=float, default=1.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-nmax", type=int, default=None, help="largest cluster/chain (24 water, 20 alkane)")
    ap.add_argument("--serve-precision", default="fp64", choices=["fp64", "fp32", "bf16x3"])
    ap.add_argument("--serve-slots", type=int, default=12, help="jobs in flight")
    ap.add_argument("--serve-integral-threads", type=int, default=0,
                    help="threads of each host integral call (0: all host cores; concurrent "
                         "preparations share them)")
    ap.add_argument("--serve-prep-threads", type=int, default=2,
                    help="host threads preparing admitted jobs (integral threads = cores / this)")
    return ap.parse_args()


def copy_calibration():
    """A plain device-to-device copy of 4 GiB, three times: per launch it reads
    and writes exactly 4 GiB, which calibrates FETCH_SIZE / WRITE_SIZE."""
    import torch

    n = (4 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_per_launch": 4 << 30, "launches": 3}))


def measured_peaks(read_buf=None, read_reps=3):
    """Achievable ceilings of this device from libf2probe.so: TFLOP/s of
    register-only MFMA chains (bf16 16x16x32, the BF16x3 kernels' instruction;
    fp32 16x16x4; fp64 16x16x4), eight waves per SIMD, ~20 ms per launch, and
    the HBM GB/s of one streaming read of read_buf (default: a fresh 16 GiB
    buffer).  Timed with HIP events on the current stream."""
    import ctypes

    import torch

    from forte2_amd import _native as nat

    lib = nat.probe_lib()
    n = ctypes.c_int64()
    nat.probe_check(lib.f2probe_sink_len(ctypes.byref(n)), "f2probe_sink_len")
    sink = torch.empty(n.value, dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    out = {}
    for kind, iters in (("bf16", 100000), ("fp32", 50000), ("fp64", 25000)):
        fl, ms = ctypes.c_double(), ctypes.c_double()
        nat.probe_check(lib.f2probe_mfma(nat.F2PROBE_KIND[kind], iters, ctypes.c_void_p(sink.data_ptr()), n.value,
                                         ctypes.c_void_p(st), ctypes.byref(fl), ctypes.byref(ms)), "f2probe_mfma")
        out[f"mfma_{kind}_TFLOPs"] = fl.value / (ms.value * 1e-3) / 1e12
        out[f"mfma_{kind}_ms_per_launch"] = ms.value
    own = read_buf is None
    if own:
        read_buf = torch.empty((16 << 30) // 4, dtype=torch.float32, device="cuda")
    nbytes = read_buf.numel() * read_buf.element_size()
    ms = ctypes.c_double()
    nat.probe_check(lib.f2probe_read(ctypes.c_void_p(read_buf.data_ptr()), nbytes, read_reps,
                                     ctypes.c_void_p(sink.data_ptr()), n.value, ctypes.c_void_p(st),
                                     ctypes.byref(ms)), "f2probe_read")
    out["hbm_read_GBps"] = nbytes / (ms.value * 1e-3) / 1e9
    out["hbm_read_bytes"] = nbytes
    if own:
        del read_buf
    return out


def scf_leg(n, mixed, seed=None, log=None, profile=False, alkane=None):
    """Real (H2O)_n cc-pVDZ/cc-pVTZ-JKFIT RHF (or, with `alkane`, the
    C_nH_(2n+2)+ cation UHF in cc-pVTZ/cc-pVTZ-JKFIT): integrals on the host
    streamed into HBM, then the FP64 plain form and the precision form on
    the same B, both to e_tol 1e-10.  Returns energies, iterations, SCF
    seconds and the final energy error of the precision form against FP64."""
    import functools

    import torch

    from forte2_amd import geometry as geo
    from forte2_amd import integrals
    from forte2_amd.scf import RHF, UHF
    from forte2_amd.system import System

    log = log or (lambda msg: None)
    dev = torch.device("cuda", 0)
    box = {16: 9.0, 64: 14.5}.get(n, 3.6 * n ** (1.0 / 3.0))
    seed = n if seed is None else seed
    integrals.set_threads(host_cores())
    t0 = time.perf_counter()
    if alkane:
        s = System(geo.alkane(alkane), "cc-pvtz", "cc-pvtz-jkfit", device=dev)
        SCF = functools.partial(UHF, 1, ms=0.5)
        name = f"C{alkane}H{2 * alkane + 2}+ UHF (ms 0.5) cc-pVTZ/cc-pVTZ-JKFIT"
    else:
        s = System(geo.water_cluster(n, box, seed=seed), "cc-pvdz", "cc-pvtz-jkfit", device=dev)
        SCF = functools.partial(RHF, 0)
        name = f"(H2O)_{n} RHF cc-pVDZ/cc-pVTZ-JKFIT, box {box} A, seed {seed}"
    fb = s.fock_builder
    last = [0.0]

    def progress(msg):
        if time.perf_counter() - last[0] > 5.0:
            last[0] = time.perf_counter()
            log(msg)

    fb.progress = progress
    fb.B_Pmn
    torch.cuda.synchronize()
    t_b = time.perf_counter() - t0
    log(f"{name}: N={s.nbf} Q={s.naux}: B formed in {t_b:.1f} s")
    out = {"workload": name,
           "nbf": s.nbf, "naux": s.naux, "nocc": s.Zsum // 2, "B_formation_s": t_b,
           "B_formation_detail": fb.B_stats, "e_tol": 1e-10}
    # one-time work outside the timed SCFs: the guess potential and the core
    # Hamiltonian (host integrals cached on the System), the eigensolver's
    # first call, one build of each precision (the context, the element scan
    # of each B copy, the workspace)
    s.sap_potential()
    s.ints_hcore()
    torch.linalg.eigh(torch.eye(s.nbf, dtype=torch.float64, device=dev))
    Cw = torch.eye(s.nbf, s.Zsum // 2, dtype=torch.float64, device=dev)
    for prec in ("fp64", mixed):
        fb.build_JK([Cw], precision=prec)
    # and the first call of every device routine the SCF loop uses (DIIS,
    # gradient, energy, the difference-density split): two short SCFs whose
    # results are discarded
    for prec, it in (("fp64", 2), (mixed, 12)):
        SCF(precision=prec, maxiter=it, die_if_not_converged=False)(s).run()
    torch.cuda.synchronize()
    runs = {}
    # the plain FP64 SCF, the precision form (FP64 re-reference tail) and, for
    # comparison, the precision form with FP64 cycles in the tail
    for name, prec, tail in (("fp64", "fp64", "reref"), (mixed, mixed, "reref"),
                             (mixed + "_fp64_cycle_tail", mixed, "full")):
        fb.timing = False
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        lines = []
        # the FP64-cycle-tail run also measures the low-precision build's
        # energy error at the first reference density (one extra build)
        r = SCF(e_tol=1e-10, precision=prec, fp64_tail=tail, log=lines.append,
                lp_error_check=(tail == "full"))(s)
        r.profile = profile
        r.run()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        stages = "".join(h[1][0] for h in r.history)
        runs[name] = {"E": r.E, "iterations": r.iter, "scf_seconds": dt, "stages": stages,
                      "converged": bool(r.converged), "lp_error_vs_fp64_Eh": r.lp_error_vs_fp64,
                      "dE_per_iteration": [h[2] - g[2] for g, h in zip(r.history, r.history[1:])],
                      "iterations_log": lines}
        if profile:
            runs[name]["phase_seconds"] = r.phase_seconds
        log(f"{name}: E={r.E:.12f} iters={r.iter} {dt:.2f} s stages={stages}")
    out["runs"] = runs
    # per-phase times of one build on this real B (the converged occupied
    # orbitals; J reads only the elements where B is nonzero for some P)
    Cocc = r.C[0][:, : s.Zsum // 2].contiguous()
    out["real_B_build_ms"] = {prec: kernel_times(fb, Cocc, prec)[0] for prec in ("fp64", mixed)}
    log("real-B builds (ms): " + json.dumps({p: round(v["total"], 2) for p, v in out["real_B_build_ms"].items()}))
    out["final_energy_error_vs_fp64_Eh"] = abs(runs[mixed]["E"] - runs["fp64"]["E"])
    out["scf_speedup_vs_fp64"] = runs["fp64"]["scf_seconds"] / runs[mixed]["scf_seconds"]
    full = runs[mixed + "_fp64_cycle_tail"]
    out["fp64_cycle_tail"] = {"final_energy_error_vs_fp64_Eh": abs(full["E"] - runs["fp64"]["E"]),
                              "scf_speedup_vs_fp64": runs["fp64"]["scf_seconds"] / full["scf_seconds"]}
    fb.release()
    s.fock_builder = None
    del s, fb
    torch.cuda.empty_cache()
    return out


def scf_mode(args):
    out = scf_leg(args.waters, args.mixed, args.seed,
                  log=lambda m: print(f"[scf] {m}", file=sys.stderr, flush=True), profile=args.scf_profile,
                  alkane=args.scf_alkane)
    print(json.dumps(out))


def host_cores():
    """Host threads this process may use: OMP_NUM_THREADS when set (the share of
    the host given to one GPU), else the affinity mask."""
    n = len(os.sched_getaffinity(0))
    try:
        return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))
    except ValueError:
        return n


def serve_leg(nreq, rate, nmax=24, precision="fp64", slots=12, prep_threads=2, log=None):
    """A bounded slice of the serving workload: the first nreq requests of the
    seeded (H2O)_n stream (n ~ U{1..nmax}, seed 7) at Poisson rate `rate`,
    served with continuous batching; p50/p99 latency and jobs/s."""
    import torch

    from forte2_amd import integrals
    from forte2_amd.serving import SCFServer, serving_stream

    integrals.set_threads(host_cores())
    reqs = serving_stream(nreq=nreq, rate=rate, nmax=nmax)
    srv = SCFServer(max_active=slots, prep_threads=prep_threads, precision=precision, log=log)
    res = srv.run(reqs)
    torch.cuda.synchronize()
    s = srv.summary()
    return {
        "workload": f"{len(reqs)} RHF requests, (H2O)_n n~U{{1..{nmax}}}, cc-pVDZ/cc-pVTZ-JKFIT, "
                    f"Poisson arrivals at {rate}/s (seed 7), {slots} slots, {precision}",
        "rate_per_s": rate, "requests": s["requests"], "completed": s["completed"],
        "jobs_per_s": s["jobs_per_s"], "latency_p50_s": s["latency_p50_s"],
        "latency_p99_s": s["latency_p99_s"], "latency_mean_s": s["latency_mean_s"],
        "mean_group_size": s["mean_group_size"], "wall_s": s["wall_s"],
        "largest_nbf": max((r.nbf for r in res.values()), default=0),
        "group_failures": s["group_failures"], "build_failures": s["build_failures"],
    }


def serve_mode(args):
    """A stream of SCF requests through the continuous-batching server."""
    import numpy as np
    import torch

    from forte2_amd import integrals
    from forte2_amd.serving import SCFServer, alkane_stream, serving_stream

    if args.serve == "water":
        nmax = args.serve_nmax or 24
        reqs = serving_stream(nreq=args.serve_requests, rate=args.serve_rate, nmax=nmax)
        workload = (f"{len(reqs)} RHF requests, (H2O)_n n~U{{1..{nmax}}}, cc-pVDZ/cc-pVTZ-JKFIT, "
                    f"Poisson arrivals at {args.serve_rate}/s (seed 7)")
    else:
        nmax = args.serve_nmax or 20
        reqs = alkane_stream(nreq=args.serve_requests, rate=args.serve_rate, nmax=nmax)
        ncat = sum(1 for r in reqs if r.charge == 1)
        workload = (f"{len(reqs)} UHF requests, C_nH_(2n+2) n~U{{2..{nmax}}} ({ncat} cations, doublet; "
                    f"neutrals ms=0), cc-pVTZ/cc-pVTZ-JKFIT, Poisson arrivals at {args.serve_rate}/s (seed 11)")
    cores = host_cores()
    ithreads = args.serve_integral_threads or cores
    integrals.set_threads(ithreads)
    last = [0.0]

    def log(msg):
        t = time.perf_counter()
        if t - last[0] >= 10.0:
            last[0] = t
            print(f"[serve] {msg}", file=sys.stderr, flush=True)

    print(f"[serve] {workload}; host cores {cores}, {args.serve_prep_threads} prep threads x "
          f"{ithreads} integral threads", file=sys.stderr, flush=True)
    srv = SCFServer(max_active=args.serve_slots, prep_threads=args.serve_prep_threads,
                    precision=args.serve_precision, log=log)
    res = srv.run(reqs)
    torch.cuda.synchronize()
    s = srv.summary()
    lat = sorted(r.latency for r in res.values() if r.error is None and r.converged)
    queue = [r.admitted - r.arrival for r in res.values() if r.error is None]
    gpu = [r.finished - r.started for r in res.values() if r.error is None and r.started > 0]
    fails = {r.id: r.error for r in res.values() if r.error is not None or not r.converged}
    print(json.dumps({
        "metric": "SCF serving latency p50/p99 and jobs/s",
        "value": s["jobs_per_s"], "unit": "jobs/s", "n_gpus": 1, "higher_is_better": True,
        "latency_p50_s": s["latency_p50_s"], "latency_p99_s": s["latency_p99_s"],
        "latency_mean_s": s["latency_mean_s"], "latency_max_s": lat[-1] if lat else None,
        "queue_wait_mean_s": float(np.mean(queue)) if queue else None,
        "device_residence_mean_s": float(np.mean(gpu)) if gpu else None,
        "requests": s["requests"], "completed": s["completed"], "failed": s["failed"],
        "failures": dict(list(fails.items())[:5]), "wall_s": s["wall_s"], "ticks": s["ticks"],
        "mean_group_size": s["mean_group_size"], "precision": args.serve_precision,
        "config": {"workload": workload, "slots": args.serve_slots,
                   "prep_threads": args.serve_prep_threads, "integral_threads": ithreads},
    }))


def otf_mode(args):
    """The integral-direct builder (FockBuilderOTF) on a real (H2O)_n geometry:
    (P|mn) evaluated by the host engine every sweep, contracted on the device.
    A second leg keeps the (P|mn) blocks resident in HBM (cache)."""
    import numpy as np
    import torch
    from threadpoolctl import threadpool_limits

    from forte2_amd import geometry as geo
    from forte2_amd import integrals
    from forte2_amd.otf import FockBuilderOTF
    from forte2_amd.system import System
    from forte2_amd.workloads import orthonormal_C
    from oracle import otf as ootf

    n = args.otf
    dev = torch.device("cuda", 0)
    cores = host_cores()
    integrals.set_threads(cores)
    s = System(geo.water_cluster(n, 9.0 if n == 16 else None, seed=n), "cc-pvdz", "cc-pvtz-jkfit", device=dev)
    N, Q, o = s.nbf, s.naux, s.Zsum // 2
    C = orthonormal_C(N, o, seed=n, device=dev)
    print(f"[otf] (H2O)_{n}: N={N} Q={Q} o={o}, {cores} host cores", file=sys.stderr, flush=True)

    def leg(cache_mb):
        fb = FockBuilderOTF(s, jk_mem_thres_mb=args.otf_mem, cache_mb=cache_mb, device=dev)
        for _ in range(args.warmup):
            fb.build_JK([C])
        torch.cuda.synchronize()
        h0 = fb.host_seconds
        t0 = time.perf_counter()
        for _ in range(args.steps):
            J, K = fb.build_JK([C])
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / args.steps
        info = {"ms_per_build": t * 1e3, "builds_per_s": 1.0 / t,
                "host_integral_s_per_build": (fb.host_seconds - h0) / args.steps,
                "blocks": len(fb.blocks), "pblksize": fb.pblksize, "iblksize": fb.iblksize}
        fb.release()
        return info, J[0], K[0]

    plain, J, K = leg(0.0)
    print(f"[otf] direct: {plain['ms_per_build']:.1f} ms/build", file=sys.stderr, flush=True)
    cached, Jc, Kc = leg(args.otf_cache)
    print(f"[otf] cached: {cached['ms_per_build']:.1f} ms/build", file=sys.stderr, flush=True)
    sweeps = -(-o // plain["iblksize"])
    # CPU: the same host integrals per sweep plus the oracle's numpy contractions
    with threadpool_limits(limits=cores):
        t0 = time.perf_counter()
        T = s.coulomb_3c()
        t_int = time.perf_counter() - t0
        met = ootf.Metric(s.coulomb_2c())
        Ch = C.cpu().numpy()
        blocks = [(0, Q)]
        t0 = time.perf_counter()
        Jr, Kr = ootf.JK_kernel(met, T, blocks, Ch, plain["iblksize"])
        t_con = time.perf_counter() - t0
    t_cpu = (sweeps + (1 if sweeps == 1 else 0)) * t_int + t_con
    err = max(float(np.abs(J.cpu().numpy() - Jr).max()), float(np.abs(K.cpu().numpy() - Kr).max()))
    print(json.dumps({
        "metric": "integral-direct JK builds/s", "value": plain["builds_per_s"], "unit": "builds/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": plain["ms_per_build"],
        "higher_is_better": True, "dtype": "fp64", "data": "synthetic geometry",
        "config": {"workload": f"(H2O)_{n} RHF J/K build, integral-direct, cc-pVDZ/cc-pVTZ-JKFIT "
                               f"(nbf {N}, naux {Q}, nocc {o}), jk_mem_thres_mb {args.otf_mem}"},
        "direct": plain, "cached": dict(cached, cache_mb=args.otf_cache),
        "max_abs_diff_vs_oracle": err,
        "cpu_baseline": {"value": 1.0 / t_cpu, "unit": "builds/s", "cores": cores, "kind": "port",
                         "sample": f"full build: {sweeps + (1 if sweeps == 1 else 0)} host evaluations of (P|mn) "
                                   f"({t_int:.2f} s each) + oracle numpy contractions ({t_con:.2f} s)"},
    }))


def measure_traffic(args):
    """HBM bytes per kernel launch from rocprofv3 counters, one counter per pass,
    each pass a child process: FETCH_SIZE and WRITE_SIZE over `--profile-steps`
    (2 builds of each precision at the default workload) and over a plain
    4 GiB device copy, whose known bytes give the factors that turn the counters
    into bytes.  Writes the table bench.py reads (with the SHA-256 of the kernel
    sources it was measured on).  This process itself never touches the GPU."""
    import csv
    import glob
    import subprocess
    import tempfile

    work = tempfile.mkdtemp(prefix="f2jk_pmc_")
    commands = []

    def pass_(counter, tag, bench_args, regex=None):
        d = os.path.join(work, tag)
        cmd = ["rocprofv3", "--pmc", counter]
        if regex:
            cmd += ["--kernel-include-regex", regex]
        cmd += ["-f", "csv", "-d", d, "-o", "run", "--", sys.executable, "-u", os.path.join(ROOT, "bench.py")]
        cmd += bench_args
        commands.append(" ".join(cmd[:cmd.index("--") + 1] + ["python3", "-u", "bench.py"] + bench_args))
        print(f"[traffic] {counter} {tag}", file=sys.stderr, flush=True)
        subprocess.run(cmd, check=True, timeout=900, cwd=ROOT)
        per = {}
        for fn in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(fn) as f:
                for row in csv.DictReader(f):
                    if row.get("Counter_Name") != counter:
                        continue
                    name = row["Kernel_Name"]
                    name = name[5:] if name.startswith("void ") else name
                    name = name.split("(")[0]
                    disp = row.get("Dispatch_Id") or row.get("Correlation_Id")
                    per.setdefault(name, {}).setdefault(disp, 0.0)
                    per[name][disp] += float(row["Counter_Value"])
        if not per:
            raise RuntimeError(f"no {counter} rows under {d}")
        return {k: (sum(v.values()) / len(v), len(v)) for k, v in per.items()}

    copy_bytes = 4 << 30
    fc = pass_("FETCH_SIZE", "fetch_copy", ["--copy-calibration"])
    wc = pass_("WRITE_SIZE", "write_copy", ["--copy-calibration"])
    copy_name = max(fc, key=lambda k: fc[k][0])
    fetch_factor = copy_bytes / 1024.0 / fc[copy_name][0]
    write_factor = copy_bytes / 1024.0 / wc[copy_name][0]
    entries = {}
    for prec in ("fp64", args.mixed):
        run = ["--profile-steps", "--profile-precision", prec, "--steps", "2", "--warmup", "0",
               "--waters", str(args.waters)]
        fe = pass_("FETCH_SIZE", f"fetch_{prec}", run, regex="f2jk")
        wr = pass_("WRITE_SIZE", f"write_{prec}", run, regex="f2jk")
        kern = {}
        for name in sorted(set(fe) | set(wr)):
            f_kib, nl = fe.get(name, (0.0, 0))
            w_kib, _ = wr.get(name, (0.0, 0))
            kern[name] = {"launches": nl, "fetch_kib": f_kib, "write_kib": w_kib,
                          "read_bytes": f_kib * 1024.0 * fetch_factor, "write_bytes": w_kib * 1024.0 * write_factor}
        entries[f"{args.waters}:{prec}"] = {"kernels": kern}
    tab = {"kernel_sources_sha256": kernel_sources_sha256(), "fetch_factor": fetch_factor,
           "write_factor": write_factor, "copy_bytes": copy_bytes, "copy_kernel": copy_name,
           "commands": commands, "entries": entries}
    out = args.traffic_out or TRAFFIC_TABLE
    with open(out, "w") as f:
        json.dump(tab, f, indent=1, sort_keys=True)
    print(json.dumps({"traffic_table": out, "fetch_factor": fetch_factor, "write_factor": write_factor}))


def timed_builds(fb, C, prec, steps, warmup):
    import torch

    for _ in range(warmup):
        fb.build_JK([C], precision=prec)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        J, K = fb.build_JK([C], precision=prec)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, J[0], K[0]


def kernel_times(fb, C, prec, reps=3):
    """Per-phase HIP-event times (ms) of one build, averaged over reps."""
    import torch

    fb.timing = True
    acc, launches = {}, None
    for _ in range(reps):
        fb.build_JK([C], precision=prec)
        for k, v in fb.last_timings.items():
            acc[k] = acc.get(k, 0.0) + v / reps
        launches = fb.last_launches
    fb.timing = False
    torch.cuda.synchronize()
    return acc, launches


def cpu_baseline(B, C, q_sample):
    """The numpy oracle (forte2 FockBuilder.build_JK restated) on an aux slice.

    J and K are sums over the aux index P, so the time of a build scales
    linearly with naux; builds/s of the full problem = sample builds/s *
    q_sample / naux.
    """
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle import jk as ojk

    cores = host_cores()
    Bs = B[:q_sample].cpu().numpy()
    Ch = C.cpu().numpy()
    D = np.einsum("mi,ni->mn", Ch, Ch, optimize=True)  # formed as the reference forms it (jkbuilder.py:229)

    def median_time(fn):
        fn()  # warm-up
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    with threadpool_limits(limits=cores):
        BnPm = np.ascontiguousarray(Bs.transpose((2, 0, 1)))
        t = median_time(lambda: ojk.build_JK(Bs, [Ch], BnPm))
        tj = median_time(lambda: ojk.build_J(Bs, [D]))
        tk = median_time(lambda: ojk.build_K_nPm(Bs, [Ch], BnPm))
    Q = B.shape[0]
    return {
        "value": (1.0 / t) * q_sample / Q,
        "unit": "builds/s",
        "cores": cores,
        "kind": "port",
        "sample": f"oracle build_JK (numpy FP64, B_nPm path) on B[:{q_sample}] of the same "
                  f"(H2O)_n tensor (naux {Q}), median of 5 after 1 warm-up, "
                  f"{t:.2f} s/sample-build, scaled by {q_sample}/{Q}",
        "j_only_builds_per_s": (1.0 / tj) * q_sample / Q,
        "k_only_builds_per_s": (1.0 / tk) * q_sample / Q,
        "cpu_model": cpu_model(),
    }


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.scf:
        scf_mode(args)
        return
    if args.serve:
        serve_mode(args)
        return
    if args.otf is not None:
        otf_mode(args)
        return
    if args.copy_calibration:
        copy_calibration()
        return
    if args.measure_traffic:
        measure_traffic(args)
        return
    if args.peaks:
        print(json.dumps({"measured_peaks": measured_peaks()}))
        return
    import torch

    from forte2_amd.jkbuilder import FockBuilder
    from forte2_amd.workloads import orthonormal_C, synthetic_B, water_cluster_shape

    dev = torch.device("cuda", 0)
    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)
    scf = None
    if args.scf_leg and not args.profile_steps:
        # the real-integral leg runs first, on device memory that nothing has
        # used yet: allocating again the ~250 GB that the synthetic B held
        # takes the runtime 5-6 s (measured; allocations up to 9 GB after a
        # free take ~2 ms), which would land in its B formation time
        scf = scf_leg(args.waters, args.mixed, log=log)
        log(f"scf leg: |E_mixed - E_fp64| = {scf['final_energy_error_vs_fp64_Eh']:.2e} Eh")
    N, Q, o = water_cluster_shape(args.waters)
    B = synthetic_B(Q, N, seed=2024, device=dev)
    C = orthonormal_C(N, o, seed=2024, device=dev)
    fb = FockBuilder(B=B, device=dev)
    fb.B_fp32()
    for kv in args.set:
        k, v = kv.split("=", 1)
        fb.set_tuning(k, int(v))
    torch.cuda.synchronize()
    print(f"[bench] (H2O)_{args.waters}: N={N} Q={Q} o={o} inputs resident", file=sys.stderr, flush=True)

    if args.profile_steps:
        timed_builds(fb, C, args.profile_precision or args.mixed, args.steps, args.warmup)
        return

    t_mix, Jm, Km = timed_builds(fb, C, args.mixed, args.steps, args.warmup)
    print(f"[bench] {args.mixed}: {t_mix*1e3:.2f} ms/build", file=sys.stderr, flush=True)
    t64, J64, K64 = timed_builds(fb, C, "fp64", max(2, args.steps // 2), 1)
    print(f"[bench] fp64: {t64*1e3:.2f} ms/build", file=sys.stderr, flush=True)
    err_j = float((Jm - J64).abs().max())
    err_k = float((Km - K64).abs().max())
    rel_j = err_j / max(float(J64.abs().max()), 1e-300)
    rel_k = err_k / max(float(K64.abs().max()), 1e-300)

    ph_mix, la_mix = kernel_times(fb, C, args.mixed)
    ph64, la64 = kernel_times(fb, C, "fp64")
    jb_mix, kf = fb.work([o], args.mixed)
    jb64, _ = fb.work([o], "fp64")
    kx_flops = 2.0 * Q * N * N * o          # algorithmic half-transform flops of one build
    kx_ms_launch = ph_mix["kx"] / max(1, la_mix["kx"])
    kx_flops_launch = kx_flops / max(1, la_mix["kx"])
    # B read once plus X written once, per launch (FP32 B, hi/lo bf16 X)
    kx_bytes_launch = (4.0 * Q * N * N + 4.0 * Q * N * 32 * ((o + 31) // 32)) / max(1, la_mix["kx"])
    achieved = kx_flops_launch / (kx_ms_launch * 1e-3) / 1e12
    peak = PEAK_TFLOPS[args.mixed]
    k_ms = ph_mix["kx"] + ph_mix["ks"] + ph_mix["k_reduce"]
    k64_ms = ph64["kx"] + ph64["ks"] + ph64["k_reduce"]
    j_ms = ph_mix["j_gamma"] + ph_mix["j_acc"]
    j64_ms = ph64["j_gamma"] + ph64["j_acc"]

    peaks = measured_peaks(read_buf=fb.B_fp32())
    mpeak = peaks["mfma_bf16_TFLOPs" if args.mixed == "bf16x3" else "mfma_fp32_TFLOPs"]
    print(f"[bench] measured peaks: {json.dumps(peaks)}", file=sys.stderr, flush=True)

    cpu = None
    if args.cpu_sample_aux > 0:
        cpu = cpu_baseline(B, C, min(args.cpu_sample_aux, Q))
        print(f"[bench] cpu baseline: {cpu['value']:.4f} builds/s on {cpu['cores']} cores",
              file=sys.stderr, flush=True)
    # the BF16x3 half-transform runs in 256-row tiles (kx_bf16x3_tall_kernel)
    # unless the launch shape says 128 (the library's rule: 256 for nbf >= 768)
    rows = fb.knobs.get("kx_rows", 0) or (256 if N >= 768 and N % 4 == 0 else 128)
    kx_name = "kx_bf16x3_tall_kernel" if rows == 256 else "kx_bf16x3_kernel"
    if args.mixed != "bf16x3":
        # FP32: 320-column tiles above 128 occupied columns (the library's rule)
        wide = fb.knobs.get("kx_wide", -1)
        kx_name = "kx_wide_kernel" if (wide == 1 or (wide < 0 and o > 128)) else "kx_kernel"
    kx_prefix = f"f2jk::{kx_name}" if args.mixed == "bf16x3" else f"f2jk::{kx_name}<float"
    traffic, traffic_note = measured_traffic(args.waters, args.mixed, kx_prefix)

    # the synthetic B leaves HBM before the real-integral legs
    fb.release()
    del fb, B, C, Jm, Km, J64, K64
    torch.cuda.empty_cache()
    serve = None
    if args.serve_leg_requests > 0:
        last = [0.0]

        def slog(m):
            if time.perf_counter() - last[0] >= 10.0:
                last[0] = time.perf_counter()
                log(m)

        serve = serve_leg(args.serve_leg_requests, args.serve_leg_rate, log=slog)
        log(f"serving leg: p50 {serve['latency_p50_s']:.2f} s p99 {serve['latency_p99_s']:.2f} s "
            f"{serve['jobs_per_s']:.2f} jobs/s")

    line = {
        "metric": METRIC,
        "value": 1.0 / t_mix,
        "unit": "builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mix * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.mixed == "fp32" else "bf16x3/fp32",
        "data": "synthetic",
        "config": {
            "workload": f"(H2O)_{args.waters} RHF J/K build, cc-pVDZ / cc-pVTZ-JKFIT shapes "
                        f"(nbf {N}, naux {Q}, nocc {o}), B full-square (P,m,n) in HBM",
            "precision_form": args.mixed,
            "nbf": N, "naux": Q, "nocc": o,
        },
        "fp64": {"builds_per_s": 1.0 / t64, "ms_per_build": t64 * 1e3},
        "speedup_vs_fp64": t64 / t_mix,
        "max_abs_diff_vs_fp64": {"J": err_j, "K": err_k},
        # relative to the largest FP64 element; the SCF-level error of the
        # precision form is measured on real integrals by --scf
        "max_rel_diff_vs_fp64": {"J": rel_j, "K": rel_k},
        "phases_ms": {args.mixed: ph_mix, "fp64": ph64},
        "J": {
            "GBps": jb_mix / (j_ms * 1e-3) / 1e9, "fp64_GBps": jb64 / (j64_ms * 1e-3) / 1e9,
            "frac_of_hbm_peak": jb_mix / (j_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "frac_of_measured_read": jb_mix / (j_ms * 1e-3) / 1e9 / peaks["hbm_read_GBps"],
            "bytes_algorithmic": jb_mix, "layout": "lower triangle read from full-square B",
        },
        "K": {
            "TFLOPs": kf / (k_ms * 1e-3) / 1e12, "mfma_util": kf / (k_ms * 1e-3) / 1e12 / peak,
            "fp64_TFLOPs": kf / (k64_ms * 1e-3) / 1e12,
            "fp64_mfma_util": kf / (k64_ms * 1e-3) / 1e12 / PEAK_TFLOPS["fp64"],
            "flops_algorithmic": kf, "flops_full_gemm": 4.0 * Q * N * N * o,
        },
        "roofline": {
            "kernel": f"{kx_name} (K half-transform X = B C, {rows}-row tiles)" if args.mixed == "bf16x3"
                      else f"{kx_name} (K half-transform X = B C)",
            "bound": "mfma",
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": traffic,
            "traffic_note": traffic_note + "; algorithmic bytes per launch %.4g" % kx_bytes_launch,
            "launches_per_build": la_mix["kx"],
            "avg_launch_ms": kx_ms_launch,
            # BF16x3 runs three bf16 MFMA products per algorithmic product
            "executed_flops_per_algorithmic": 3 if args.mixed == "bf16x3" else 1,
            "executed_frac": achieved * (3 if args.mixed == "bf16x3" else 1) / peak,
            # the same executed rate against what register-only chains of the
            # kernel's MFMA instruction reach on this device (measured_peaks)
            "measured_peak": mpeak,
            "executed_frac_of_measured": achieved * (3 if args.mixed == "bf16x3" else 1) / mpeak,
        },
        "measured_peaks": peaks,
        "cpu_baseline": cpu,
        # the accuracy leg of the metric: the real (H2O)_n RHF in the precision
        # form against FP64 (final-energy error; the bar is 1e-8 Eh)
        "final_energy_error_vs_fp64_Eh": scf["final_energy_error_vs_fp64_Eh"] if scf else None,
        "scf": scf,
        # the latency leg: p50/p99 of a bounded slice of the serving stream
        "latency_p50_s": serve["latency_p50_s"] if serve else None,
        "latency_p99_s": serve["latency_p99_s"] if serve else None,
        "serving": serve,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
