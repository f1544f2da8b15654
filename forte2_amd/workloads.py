"""Synthetic inputs of the J/K workloads (shapes of the benchmarked systems).

Sizes with the cc-pVDZ orbital basis and cc-pVTZ-JKFIT auxiliary basis:
water N=24, Q=139, o=5 per molecule, so (H2O)_n has N=24n, Q=139n, o=5n.

``synthetic_B`` fills a (Q, N, N) tensor that is symmetric in (m, n) directly on
the device, one P block at a time, from a seeded generator; the values are
O(0.05) like a real B.  ``orthonormal_C`` returns SCF-like occupied
coefficients (orthonormal columns).
"""

import torch

WATER_N, WATER_Q, WATER_O = 24, 139, 5


def water_cluster_shape(n):
    return WATER_N * n, WATER_Q * n, WATER_O * n


def synthetic_B(Q, N, seed=0, device="cuda", dtype=torch.float64, scale=0.05, block=None):
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    B = torch.empty((Q, N, N), dtype=dtype, device=device)
    if block is None:
        block = max(1, min(Q, (256 << 20) // max(1, 8 * N * N)))
    for p0 in range(0, Q, block):
        p1 = min(Q, p0 + block)
        x = torch.randn((p1 - p0, N, N), generator=g, device=device, dtype=torch.float64)
        x = 0.5 * scale * (x + x.transpose(1, 2))
        B[p0:p1].copy_(x)
        del x
    return B


def orthonormal_C(N, o, seed=0, device="cuda"):
    g = torch.Generator(device="cpu")
    g.manual_seed(int(seed) + 7919)
    A = torch.randn((N, max(o, 1)), generator=g, dtype=torch.float64)
    q, _ = torch.linalg.qr(A)
    return q[:, :o].contiguous().to(device)
