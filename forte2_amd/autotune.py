"""Launch-shape autotuner for the J/K kernels.

The tile edges of the kernels are fixed by their register/LDS layout (128x128
FP64/FP32 tiles, 128 x 32*NBW BF16x3 tiles with NBW chosen from nocc).  What
varies with the job size is the launch shape:
  * ``target_blocks``: how many workgroups a split launch aims for.  This
    fixes the P chunks of J pass 2 and the P splits of the K contraction,
    i.e. the parallelism / partial-buffer trade-off;
  * ``xchunk_bytes``: the size of the half-transformed chunk X kept in the
    workspace.  It trades L2/MALL reuse between KX and KS against launch
    count.

``tune(shapes)`` times every candidate on synthetic B of each shape with HIP
events and keeps the fastest; ``save``/``load`` keep the table as JSON
(``forte2_amd/data/tuning_gfx950.json`` ships the table measured on an
MI355X).  ``apply(fb)`` sets the entry of the nearest tuned size class on a
FockBuilder.  Candidates change only the decomposition: every candidate gives
the same J/K up to the reduction order (parity is tested separately), and a
given table gives bitwise-stable results.
"""

import json
import os

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tuning_gfx950.json")

CANDIDATES = {
    "target_blocks": [256, 512, 1024, 2048, 4096],
    "xchunk_bytes": [1 << 29, 1 << 30, 1 << 31, 1 << 32],
}


def _key(nbf, naux, precision):
    return f"{precision}:{nbf}:{naux}"


def tune(shapes, precisions=("fp64",), reps=3, device=None, log=print):
    """shapes: [(nbf, naux, nocc)]; returns {key: {"target_blocks", "xchunk_bytes", "ms"}}."""
    import torch

    from .jkbuilder import FockBuilder
    from .workloads import orthonormal_C, synthetic_B

    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    table = {}
    for N, Q, o in shapes:
        B = synthetic_B(Q, N, seed=N + Q, device=dev)
        C = orthonormal_C(N, o, seed=N, device=dev)
        for prec in precisions:
            best = None
            for tb in CANDIDATES["target_blocks"]:
                for xb in CANDIDATES["xchunk_bytes"]:
                    fb = FockBuilder(B=B, device=dev)
                    fb.set_tuning("target_blocks", tb)
                    fb.set_tuning("xchunk_bytes", xb)
                    fb.build_JK([C], precision=prec)  # warm-up (allocates the workspace)
                    fb.timing = True
                    ms = []
                    for _ in range(reps):
                        fb.build_JK([C], precision=prec)
                        ms.append(fb.last_timings["total"])
                    t = sorted(ms)[len(ms) // 2]
                    if best is None or t < best["ms"]:
                        best = {"target_blocks": tb, "xchunk_bytes": xb, "ms": t}
                    del fb
            table[_key(N, Q, prec)] = best
            if log:
                log(f"[autotune] N={N} Q={Q} o={o} {prec}: {best}")
        del B
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


def lookup(nbf, naux, precision="fp64", table=None):
    """Entry of the tuned size class nearest to (nbf, naux) (log-distance), or None."""
    import math

    table = load() if table is None else table
    best, bd = None, None
    for k, v in table.items():
        prec, n, q = k.split(":")
        if prec != precision:
            continue
        d = abs(math.log(int(n) / nbf)) + abs(math.log(int(q) / naux))
        if bd is None or d < bd:
            best, bd = v, d
    return best


def apply(fb, precision="fp64", table=None):
    """Set the tuned launch shape of the nearest size class on FockBuilder `fb`."""
    e = lookup(fb.nbf, fb.naux, precision, table)
    if e is not None:
        fb.set_tuning("target_blocks", e["target_blocks"])
        fb.set_tuning("xchunk_bytes", e["xchunk_bytes"])
    return e


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="tune J/K launch shapes on this GPU")
    ap.add_argument("--waters", default="1,2,4,8,12,16,24,32,64")
    ap.add_argument("--precisions", default="fp64,bf16x3")
    ap.add_argument("--out", default=TABLE)
    a = ap.parse_args()
    shapes = [(24 * n, 139 * n, 5 * n) for n in map(int, a.waters.split(","))]
    save(tune(shapes, tuple(a.precisions.split(","))), a.out)
