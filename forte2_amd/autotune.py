"""Launch-shape autotuner for the J/K kernels, single-job and grouped.

The tile edges of the kernels are fixed by their register/LDS layout (128x128
FP64/FP32 tiles, 128 x 32*NBW BF16x3 tiles).  What varies with the job size is
the launch shape:
  * ``target_blocks``: how many workgroups a split launch aims for.  It fixes
    the P chunks of J pass 2 and the P splits of the K contraction, i.e. the
    parallelism / partial-buffer trade-off;
  * ``xchunk_bytes``: the size of the half-transformed chunk X kept in the
    workspace (single-job builds; a grouped build holds X in one chunk).  It
    trades L2/MALL reuse between KX and KS against launch count;
  * ``bf16_nbw_max`` (BF16x3): the widest half-transform tile, 32*NBW
    columns.  Narrower tiles use fewer registers and LDS but re-read the B
    rows once per column tile;
  * ``kx_waves`` (BF16x3): 4 or 8 waves per half-transform workgroup of
    the 128-row kernel;
  * ``ks_dma`` is not tuned: the contraction with X tiles loaded straight
    into LDS is bitwise the register-staged one and is used with the
    256-wide tiles, where it measured faster;
  * ``kx_rows`` (BF16x3, single-job): 128 or 256, the rows of a
    half-transform tile (256: the tall kernel, 2 x 4 waves, even widths);
  * ``ks_tile`` (BF16x3, single-job): 128 or 256, the lower-tile edge of the
    contraction.

``tune(shapes)`` times every candidate on synthetic B of each shape with HIP
events and keeps the fastest; ``tune_grouped(shapes)`` does the same for a
grouped build of several jobs of one shape, timed through the serving layer's
group path (the shapes the server batches).  ``save``/``load`` keep the table
as JSON (``forte2_amd/data/tuning_gfx950.json`` ships the table measured on an
MI355X).  ``apply(fb, precision, grouped=...)`` sets the entry of the nearest
tuned size class on a FockBuilder.  Candidates change only the decomposition:
every candidate gives the same J/K up to the reduction order (parity is tested
separately), and a given table gives bitwise-stable results.

Table keys: "{precision}:{nbf}:{naux}" (single-job) and
"group:{precision}:{nbf}:{naux}" (grouped).
"""

import itertools
import json
import math
import os
import time

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tuning_gfx950.json")

CANDIDATES = {
    "target_blocks": [256, 512, 1024, 2048, 4096, 8192, 16384],
    "xchunk_bytes": [1 << 29, 1 << 30, 1 << 31, 1 << 32],
    "bf16_nbw_max": [4, 5, 6, 8, 10],
    "kx_waves": [4, 8],
    "ks_tile": [128, 256],
    "kx_rows": [128, 256],
}
# the BF16x3 grid is larger (tile cap, waves): fewer chunk sizes there
XCHUNK_BF16 = [1 << 31, 1 << 32]
KEYS = ("target_blocks", "xchunk_bytes", "bf16_nbw_max", "kx_waves", "ks_tile", "kx_rows")


def _key(nbf, naux, precision, grouped=False):
    return ("group:" if grouped else "") + f"{precision}:{nbf}:{naux}"


def _noccs(o):
    return list(o) if isinstance(o, (list, tuple)) else [int(o)]


KS_TILE_MIN_NBF = 512  # below this a 256-wide contraction tile is mostly empty: not offered


def _candidates(precision, grouped, nbf=None):
    names = ["target_blocks"] + ([] if grouped else ["xchunk_bytes"])
    if precision == "bf16x3":
        names.append("bf16_nbw_max")
        if not grouped:  # grouped builds use 4-wave half-transform and 128-wide contraction tiles
            names.append("kx_waves")
            if nbf is None or nbf >= KS_TILE_MIN_NBF:
                names += ["ks_tile", "kx_rows"]
    pools = {**CANDIDATES, "xchunk_bytes": XCHUNK_BF16 if precision == "bf16x3" else CANDIDATES["xchunk_bytes"]}
    for vals in itertools.product(*[pools[n] for n in names]):
        cand = dict(zip(names, vals))
        if cand.get("kx_rows") == 256 and cand.get("kx_waves") == 4:
            continue  # the tall kernel has one wave count: the same launch as kx_waves 8
        yield cand


def _set(fb, cand):
    for k, v in cand.items():
        fb.set_tuning(k, v)


def tune(shapes, precisions=("fp64",), reps=3, device=None, log=print):
    """shapes: [(nbf, naux, nocc or [nocc per density])]; returns {key: entry}."""
    import torch

    from .jkbuilder import FockBuilder
    from .workloads import orthonormal_C, synthetic_B

    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    table = {}
    for N, Q, o in shapes:
        B = synthetic_B(Q, N, seed=N + Q, device=dev)
        Cs = [orthonormal_C(N, x, seed=N + k, device=dev) for k, x in enumerate(_noccs(o))]
        fb = FockBuilder(B=B, device=dev)
        for prec in precisions:
            best = None
            for cand in _candidates(prec, grouped=False, nbf=N):
                _set(fb, cand)
                fb.timing = False
                fb.build_JK(Cs, precision=prec)  # warm-up (workspace of this plan)
                fb.timing = True
                ms = []
                for _ in range(reps):
                    fb.build_JK(Cs, precision=prec)
                    ms.append(fb.last_timings["total"])
                t = sorted(ms)[len(ms) // 2]
                if best is None or t < best["ms"]:
                    best = dict(cand, ms=t)
            fb.timing = False
            table[_key(N, Q, prec)] = best
            if log:
                log(f"[autotune] N={N} Q={Q} o={o} {prec}: {best}")
        del fb, B
        torch.cuda.empty_cache()
    return table


def tune_grouped(shapes, precisions=("fp64",), group=4, reps=3, device=None, log=print,
                 hbm_fraction=0.6):
    """Grouped builds of `group` jobs of each shape (fewer when their B would
    not fit `hbm_fraction` of the free HBM); returns {"group:...": entry}."""
    import torch

    from .jkbuilder import FockBuilder
    from .serving import SCFServer
    from .workloads import orthonormal_C, synthetic_B

    class _Pending:
        def __init__(self, fb, C, prec):
            self.fb, self.pending = fb, ("JK", C, prec)

    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    srv = SCFServer(device=dev)
    table = {}
    for N, Q, o in shapes:
        if N % 2:
            continue  # the grouped path needs an even nbf
        free, _ = torch.cuda.mem_get_info(dev)
        per_job = 12 * Q * N * N + 8 * Q * N * 64 * len(_noccs(o))
        g = max(1, min(group, int(hbm_fraction * free // per_job)))
        fbs, Cs = [], []
        for j in range(g):
            fbs.append(FockBuilder(B=synthetic_B(Q, N, seed=N + Q + j, device=dev), device=dev))
            Cs.append([orthonormal_C(N, x, seed=N + 7 * j + k, device=dev) for k, x in enumerate(_noccs(o))])
        for prec in precisions:
            if not srv._groupable(_Pending(fbs[0], Cs[0], prec), "JK", prec, len(Cs[0])):
                if log:
                    log(f"[autotune] N={N} Q={Q} o={o} {prec}: not taken by the grouped path (runs alone)")
                continue
            best = None
            for cand in _candidates(prec, grouped=True):
                for fb in fbs:
                    _set(fb, cand)
                jobs = [_Pending(fb, C, prec) for fb, C in zip(fbs, Cs)]
                srv._build_group(jobs)  # warm-up
                torch.cuda.synchronize(dev)
                ms = []
                for _ in range(reps):
                    t0 = time.perf_counter()
                    srv._build_group(jobs)
                    torch.cuda.synchronize(dev)
                    ms.append(1e3 * (time.perf_counter() - t0))
                t = sorted(ms)[len(ms) // 2]
                if best is None or t < best["ms"]:
                    best = dict(cand, ms=t, group=g)
            table[_key(N, Q, prec, grouped=True)] = best
            if log:
                log(f"[autotune] group of {g}: N={N} Q={Q} o={o} {prec}: {best}")
        del fbs, Cs
        torch.cuda.empty_cache()
    return table


def save(table, path=TABLE):
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def load(path=TABLE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def lookup(nbf, naux, precision="fp64", table=None, grouped=False):
    """Entry of the tuned size class nearest to (nbf, naux) (log-distance), or None."""
    table = load() if table is None else table
    best, bd = None, None
    for k, v in table.items():
        parts = k.split(":")
        if (parts[0] == "group") != grouped:
            continue
        prec, n, q = parts[-3:]
        if prec != precision:
            continue
        d = abs(math.log(int(n) / nbf)) + abs(math.log(int(q) / naux))
        if bd is None or d < bd:
            best, bd = v, d
    return best


def apply(fb, precision="fp64", table=None, grouped=False):
    """Set the tuned launch shape of the nearest size class on FockBuilder `fb`."""
    e = lookup(fb.nbf, fb.naux, precision, table, grouped)
    if e is not None:
        for k in KEYS:
            if k in e:
                fb.set_tuning(k, e[k])
    return e


def water_shapes(ns):
    """(nbf, naux, nocc) of (H2O)_n in cc-pVDZ / cc-pVTZ-JKFIT."""
    return [(24 * n, 139 * n, 5 * n) for n in ns]


def alkane_shapes(ns):
    """(nbf, naux, [nocc_alpha, nocc_beta]) of the C_nH_(2n+2) cations in
    cc-pVTZ / cc-pVTZ-JKFIT (two densities per build)."""
    return [(58 * n + 28, 139 * n + 60, [4 * n + 1, 4 * n]) for n in ns]


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="tune J/K launch shapes on this GPU")
    ap.add_argument("--waters", default="1,2,4,8,12,16,24,32,64")
    ap.add_argument("--alkanes", default="2,5,10,15,20")
    ap.add_argument("--precisions", default="fp64,bf16x3")
    ap.add_argument("--grouped-waters", default="2,4,8,12,16,24")
    ap.add_argument("--group", type=int, default=4)
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--merge", action="store_true",
                    help="update the entries of the existing table at --out instead of replacing it")
    ap.add_argument("--skip-grouped", action="store_true", help="single-job entries only")
    ap.add_argument("--skip-single", action="store_true", help="grouped entries only")
    a = ap.parse_args()
    log = lambda m: print(m, flush=True)
    precs = tuple(a.precisions.split(","))
    ints = lambda s: [int(x) for x in s.split(",") if x]
    table = load(a.out) if a.merge else {}
    if not a.skip_single:
        table.update(tune(water_shapes(ints(a.waters)) + alkane_shapes(ints(a.alkanes)), precs, log=log))
    if not a.skip_grouped:
        table.update(tune_grouped(water_shapes(ints(a.grouped_waters)) + alkane_shapes(ints(a.alkanes)), precs,
                                  group=a.group, log=log))
    save(table, a.out)
    log(f"[autotune] wrote {len(table)} entries to {a.out}")
