"""Serving layer: grouped J/K builds and continuous batching on the GPU.

The stated tolerance for "the same alone and in the mix" is bitwise equality:
a job's grouped J/K reuse its single-job decomposition, so nothing about the
mix enters its arithmetic.
"""

import ctypes

import numpy as np
import pytest
import torch

from forte2_amd import geometry as geo
from forte2_amd.jkbuilder import FockBuilder
from forte2_amd.serving import SCFRequest, SCFServer, serving_stream
from forte2_amd.workloads import orthonormal_C, synthetic_B

pytestmark = pytest.mark.gpu


def test_grouped_build_is_bitwise_single_build(dev):
    shapes = [(139, 24, [5]), (278, 48, [10]), (700, 200, [33, 32]), (417, 72, [15]), (1668, 288, [60])]
    fbs, Cs = [], []
    for i, (Q, N, noccs) in enumerate(shapes):
        fbs.append(FockBuilder(B=synthetic_B(Q, N, seed=i, device=dev), device=dev))
        fbs[-1].set_tuning("target_blocks", 2048)  # one plan for grouped and single builds
        fbs[-1].set_tuning("ks_tile", 128)  # the grouped contraction tile
        Cs.append([orthonormal_C(N, o, seed=10 * i + k, device=dev) for k, o in enumerate(noccs)])
    singles = [fb.build_JK(C) for fb, C in zip(fbs, Cs)]
    srv = SCFServer(device=dev)

    class _Fake:
        def __init__(self, fb, C):
            self.fb, self.pending = fb, ("JK", C, "fp64")

    # mixed group (one density and two densities are grouped separately by the server)
    one = [_Fake(fb, C) for fb, C in zip(fbs, Cs) if len(C) == 1]
    outs = srv._build_group(one)
    k = 0
    for (fb, C), (Js, Ks) in zip([(f, c) for f, c in zip(fbs, Cs)], singles):
        if len(C) != 1:
            continue
        J, K = outs[k]
        k += 1
        assert torch.equal(J[0], Js[0]) and torch.equal(K[0], Ks[0])
    two = [_Fake(fb, C) for fb, C in zip(fbs, Cs) if len(C) == 2]
    outs = srv._build_group(two)
    (Js, Ks) = singles[2]
    assert all(torch.equal(a, b) for a, b in zip(outs[0][0], Js))
    assert all(torch.equal(a, b) for a, b in zip(outs[0][1], Ks))


class _Pending:
    """A stand-in job with one pending build request."""

    def __init__(self, fb, C, kind="JK", prec="fp64"):
        self.fb, self.pending = fb, (kind, C, prec)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("kind", ["JK", "K"])
def test_grouped_precision_forms(dev, prec, kind):
    # grouped builds of the precision forms: per job bitwise the same alone and
    # in the mix; K as the single-job build (same kernels, same k order, same
    # rounding of the coefficients); J within 1e-6 relative of it (the group
    # reads B with another vector width).  (700, 200, [33, 32]) once caught
    # the two pads rounding one coefficient differently (double -> bf16 in one
    # step vs through float).
    shapes = [(139, 24, [5]), (278, 48, [10]), (417, 72, [15]), (1668, 288, [60]), (700, 200, [33, 32]),
              (556, 96, [20, 19])]
    fbs, Cs = [], []
    for i, (Q, N, noccs) in enumerate(shapes):
        fbs.append(FockBuilder(B=synthetic_B(Q, N, seed=i, device=dev), device=dev))
        fbs[-1].set_tuning("target_blocks", 2048)  # one plan for grouped and single builds
        fbs[-1].set_tuning("ks_tile", 128)  # the grouped contraction tile
        Cs.append([orthonormal_C(N, o, seed=10 * i + k, device=dev) for k, o in enumerate(noccs)])
    srv = SCFServer(device=dev)
    for nd in (1, 2):
        idx = [i for i, C in enumerate(Cs) if len(C) == nd]
        mix = srv._build_group([_Pending(fbs[i], Cs[i], kind, prec) for i in idx])
        for k, i in enumerate(idx):
            alone = srv._build_group([_Pending(fbs[i], Cs[i], kind, prec)])[0]
            got = mix[k]
            if kind == "JK":
                ref_J, ref_K = fbs[i].build_JK(Cs[i], precision=prec)
                assert all(torch.equal(a, b) for a, b in zip(alone[0], got[0]))
                assert all(torch.equal(a, b) for a, b in zip(alone[1], got[1]))
                for a, b in zip(got[0], ref_J):
                    assert float((a - b).abs().max()) <= 1e-6 * float(b.abs().max())
                Kg = got[1]
            else:
                ref_K = fbs[i].build_K(Cs[i], precision=prec)
                assert all(torch.equal(a, b) for a, b in zip(alone, got))
                Kg = got
            for a, b in zip(Kg, ref_K):
                assert float((a - b).abs().max()) <= 1e-12 * float(b.abs().max())


def _requests():
    xyzs = [geo.water_cluster(n, seed=100 + n) for n in (1, 2, 3, 5)]
    reqs = [SCFRequest(i, x, arrival=0.0) for i, x in enumerate(xyzs)]
    reqs.append(SCFRequest(len(reqs), geo.alkane(2), charge=1, ms=0.5, arrival=0.0))
    return reqs


def test_results_same_alone_and_in_the_mix(dev):
    reqs = _requests()
    alone = {}
    for r in reqs:
        srv = SCFServer(device=dev, max_active=1, prep_threads=1, e_tol=1e-10)
        alone[r.id] = srv.run([SCFRequest(r.id, r.xyz, r.basis, r.aux, r.charge, r.ms, 0.0)])[r.id]
    srv = SCFServer(device=dev, max_active=12, prep_threads=4, e_tol=1e-10)
    mix = srv.run(reqs)
    print(f"[serving] mix: {srv.summary()}", flush=True)
    assert srv.summary()["mean_group_size"] > 1.0
    assert srv.group_failures == 0 and srv.build_failures == 0
    for r in reqs:
        a, m = alone[r.id], mix[r.id]
        print(f"   req {r.id}: alone {a.E:.12f} mix {m.E:.12f} iters {a.iterations}/{m.iterations}", flush=True)
        assert a.converged and m.converged
        assert a.E == m.E  # bitwise


def test_serving_matches_plain_scf(dev):
    from forte2_amd.scf import RHF
    from forte2_amd.system import System

    xyz = geo.water_cluster(3, seed=103)
    srv = SCFServer(device=dev, e_tol=1e-10)
    res = srv.run([SCFRequest(0, xyz)])[0]
    s = System(xyz, "cc-pvdz", "cc-pvtz-jkfit", device=dev)
    ref = RHF(0, e_tol=1e-10)(s).run()
    assert abs(res.E - ref.E) <= 1e-10


def test_stream_with_arrivals_and_tight_memory(dev):
    reqs = serving_stream(nreq=24, rate=20.0, nmax=8, seed=3)
    # a budget of ~2 of the largest jobs forces queueing on memory, not slots
    from forte2_amd.serving import job_footprint

    budget = 2 * job_footprint(24 * 8, 139 * 8, 80)
    srv = SCFServer(device=dev, hbm_budget=budget, max_active=12, e_tol=1e-8)
    res = srv.run(reqs)
    summ = srv.summary()
    print(f"[serving] stream: {summ}", flush=True)
    assert summ["completed"] == len(reqs)
    assert summ["group_failures"] == 0 and summ["build_failures"] == 0
    assert all(r.latency >= 0 for r in res.values())
    assert summ["latency_p99_s"] >= summ["latency_p50_s"]


def test_precision_form_in_the_mix(dev):
    reqs = _requests()[:3]
    srv = SCFServer(device=dev, e_tol=1e-10, precision="bf16x3")
    mix = srv.run(reqs)
    ref = SCFServer(device=dev, e_tol=1e-10).run([SCFRequest(r.id, r.xyz, arrival=0.0) for r in reqs])
    for r in reqs:
        assert mix[r.id].converged
        assert abs(mix[r.id].E - ref[r.id].E) <= 1e-8


def test_failures_are_isolated(dev):
    # a request with an unknown basis, one with a negative electron count and
    # one too large for the HBM budget fail on their own; the others finish
    from forte2_amd.serving import job_footprint

    w1, w2, w8 = (geo.water_cluster(n, seed=200 + n) for n in (1, 2, 8))
    reqs = [SCFRequest(0, w1, arrival=0.0),
            SCFRequest(1, w1, basis="no-such-basis", arrival=0.0),
            SCFRequest(2, w1, charge=20, arrival=0.0),
            SCFRequest(3, w2, arrival=0.0),
            SCFRequest(4, w8, arrival=0.0)]
    srv = SCFServer(device=dev, hbm_budget=int(1.5 * job_footprint(48, 278, 20)), e_tol=1e-8)
    res = srv.run(reqs)
    assert res[0].converged and res[3].converged
    for i in (1, 2, 4):
        assert res[i].error, i
        print(f"   request {i}: {res[i].error[:80]}", flush=True)
    assert "HBM budget" in res[4].error
    alone = SCFServer(device=dev, e_tol=1e-8).run([SCFRequest(3, w2, arrival=0.0)])
    assert res[3].E == alone[3].E


def test_water_stream_full_size_range_matches_plain_scf(dev):
    # the serving workload over its whole size range: (H2O)_n with n from 1 to
    # 24 (nbf up to 576, naux up to 3336), Poisson arrivals, continuous
    # batching; every served energy equals the plain device RHF of the same
    # request to 1e-10 Eh (rhf.py:35-65 semantics, FP64)
    from forte2_amd.scf import RHF
    from forte2_amd.system import System

    reqs = serving_stream(nreq=12, rate=50.0, nmax=24, seed=27)
    sizes = [r.xyz.count("O ") for r in reqs]
    assert max(sizes) == 24 and min(sizes) == 1
    srv = SCFServer(device=dev, e_tol=1e-10)
    res = srv.run(reqs)
    summ = srv.summary()
    print(f"[serving] (H2O)_n n={sizes}: {summ}", flush=True)
    assert summ["completed"] == len(reqs)
    assert summ["group_failures"] == 0 and summ["build_failures"] == 0
    assert summ["mean_group_size"] > 1.0
    assert summ["latency_p99_s"] >= summ["latency_p50_s"] > 0.0
    for r, n in zip(reqs, sizes):
        s = System(r.xyz, r.basis, r.aux, device=dev)
        ref = RHF(0, e_tol=1e-10)(s).run()
        s.fock_builder.release()
        got = res[r.id]
        print(f"   req {r.id} (H2O)_{n}: served {got.E:.12f} plain {ref.E:.12f} diff {got.E - ref.E:.2e}", flush=True)
        assert got.converged and abs(got.E - ref.E) <= 1e-10
    torch.cuda.empty_cache()
