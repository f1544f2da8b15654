"""GPU parity of the MO-basis blocks (FockBuilder.B_tensor_gen_block,
two_electron_integrals_gen_block, two_electron_integrals_block) against the
numpy oracle: FP64, <= 1e-10 max-abs.  Covers several half-transform chunks
and P blocks (block boundaries), odd nbf, column-slice C with a row stride,
numpy in / numpy out, empty orbital sets, and the reference's DF-MP2 golden
energy of water in cc-pVQZ / cc-pVQZ-JKFIT (tests/mp2/test_mp2.py:8-54).
"""

import json
import os

import numpy as np
import pytest
import torch

from forte2_amd.jkbuilder import FockBuilder
from forte2_amd.workloads import orthonormal_C, synthetic_B
from oracle import jk as ojk

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_values.json")))
EREF = GOLD["energies"]

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _err(a, b):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else a
    return float(np.abs(a - b).max())


@pytest.mark.parametrize("Q,N,ncol", [(300, 131, 21), (139, 24, 21), (300, 131, 32), (200, 260, 200), (120, 340, 336)])
def test_half_transform_blocks(dev, Q, N, ncol):
    # 21 columns: a column slice (row stride 30), written through the
    # workspace; 32 columns: X rows as wide as the kernel's, written in place;
    # 200 and 336 columns: the 320-column tiles (one tile through the
    # workspace; two column tiles, the second partial, in place)
    print(f"[mo-gpu] half-transform Q={Q} N={N} ncol={ncol}", flush=True)
    B = synthetic_B(Q, N, seed=Q, device=dev)
    Bh = B.cpu().numpy()
    wide = orthonormal_C(N, max(32, ncol), seed=3, device=dev)
    C = wide[:, 4:25] if ncol == 21 else wide
    fb = FockBuilder(B=B, device=dev)
    fb.set_tuning("xchunk_bytes", 1 << 20)  # several kernel chunks per block
    Xr = np.einsum("Pmn,ni->Pmi", Bh, C.cpu().numpy())
    seen = 0
    for P0, P1, X in fb._x_blocks(C, block_bytes=8 * N * ncol * 37):  # 37 P per block
        assert P0 == seen
        seen = P1
        assert _err(X, Xr[P0:P1]) <= TOL
    assert seen == Q


def test_B_tensor_gen_block_matches_oracle(dev):
    Q, N = 300, 131
    B = synthetic_B(Q, N, seed=9, device=dev)
    Bh = B.cpu().numpy()
    fb = FockBuilder(B=B, device=dev)
    fb.set_tuning("xchunk_bytes", 1 << 21)
    C1 = orthonormal_C(N, 40, seed=1, device=dev)[:, 3:12]
    C2 = orthonormal_C(N, 17, seed=2, device=dev)
    out = fb.B_tensor_gen_block(C1, C2)
    assert isinstance(out, torch.Tensor) and out.shape == (Q, 9, 17)
    ref = ojk.B_tensor_gen_block(Bh, C1.cpu().numpy(), C2.cpu().numpy())
    e = _err(out, ref)
    print(f"   |dB_mo| = {e:.2e}", flush=True)
    assert e <= TOL
    # numpy in, numpy out (the reference's interface)
    out_np = fb.B_tensor_gen_block(C1.cpu().numpy(), C2.cpu().numpy())
    assert isinstance(out_np, np.ndarray) and _err(out_np, ref) <= TOL
    # an empty orbital set gives an empty block
    assert fb.B_tensor_gen_block(C1[:, :0], C2).shape == (Q, 0, 17)
    with pytest.raises(ValueError):
        fb.B_tensor_gen_block(C1[:-1], C2)


def test_two_electron_integrals_match_oracle(dev):
    Q, N = 200, 64
    B = synthetic_B(Q, N, seed=4, device=dev)
    Bh = B.cpu().numpy()
    fb = FockBuilder(B=B, device=dev)
    Cs = [orthonormal_C(N, k, seed=10 + k, device=dev) for k in (3, 5, 6, 7)]
    Ch = [c.cpu().numpy() for c in Cs]
    V = fb.two_electron_integrals_gen_block(*Ch)
    assert isinstance(V, np.ndarray) and V.shape == (3, 5, 6, 7)
    e = _err(V, ojk.two_electron_integrals_gen_block(Bh, *Ch))
    print(f"   |dV| gen = {e:.2e}", flush=True)
    assert e <= TOL
    C = orthonormal_C(N, 10, seed=20, device=dev)
    Vb = fb.two_electron_integrals_block(C)
    Cn = C.cpu().numpy()
    assert isinstance(Vb, torch.Tensor)
    assert _err(Vb, ojk.two_electron_integrals_gen_block(Bh, Cn, Cn, Cn, Cn)) <= TOL


def test_mp2_golden(dev):
    # tests/mp2/test_mp2.py:8-54 (Psi4 values): RHF, then DF-MP2 from <ij|ab>
    from forte2_amd.scf import RHF
    from forte2_amd.system import System

    water = ("O 0.000000000000 0.000000000000 -0.061664597388\n"
             "H 0.000000000000 -0.711620616369 0.489330954643\n"
             "H 0.000000000000 0.711620616369 0.489330954643")
    print("[mo-gpu] water cc-pVQZ MP2", flush=True)
    s = System(water, "cc-pvqz", "cc-pvqz-jkfit", device=dev)
    scf = RHF(charge=0, e_tol=1e-10)(s).run()
    assert scf.E == pytest.approx(EREF["water_qz_mp2"]["E_scf"], abs=5e-8)
    fb = s.fock_builder
    no = scf.na
    C = scf.C[0]
    eps = scf.eps[0]
    C = C.cpu().numpy() if isinstance(C, torch.Tensor) else np.asarray(C)
    eps = eps.cpu().numpy() if isinstance(eps, torch.Tensor) else np.asarray(eps)
    Co, Cv = C[:, :no], C[:, no:]
    V = fb.two_electron_integrals_gen_block(Co, Co, Cv, Cv)
    Vr = ojk.two_electron_integrals_gen_block(fb.B_Pmn.cpu().numpy(), Co, Co, Cv, Cv)
    e = _err(V, Vr)
    print(f"   |dV| = {e:.2e}", flush=True)
    assert e <= TOL
    eo, ev = eps[:no], eps[no:]
    den = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    emp2 = scf.E + np.sum(V * (2.0 * V - V.transpose(0, 1, 3, 2)) / den)
    print(f"   E(MP2) = {emp2:.12f}", flush=True)
    assert emp2 == pytest.approx(EREF["water_qz_mp2"]["E_total"], abs=5e-8)


def test_ao_jk_equal_mo_reconstruction(dev):
    # tests/jkbuilder/test_jkbuilder.py:56-85: J and K of the occupied density
    # rebuilt from the MO-basis <pq|rs> of the full orbital set agree with the
    # AO J/K builds (two independent device paths: MFMA half-transform and
    # contraction vs MO blocks by GEMMs)
    Q, N, no = 300, 40, 7
    B = synthetic_B(Q, N, seed=21, device=dev)
    fb = FockBuilder(B=B, device=dev)
    C = orthonormal_C(N, N, seed=22, device=dev)
    J, K = fb.build_JK([C[:, :no]])
    V = fb.two_electron_integrals_block(C)          # <pq|rs> = (pr|qs)
    occ = slice(0, no)
    J_mo = torch.einsum("piqi->pq", V[:, occ, :, occ].contiguous())
    K_mo = torch.einsum("piiq->pq", V[:, occ, occ, :].contiguous())
    ej = float((C.T @ J[0] @ C - J_mo).abs().max())
    ek = float((C.T @ K[0] @ C - K_mo).abs().max())
    print(f"   |dJ_mo|={ej:.2e} |dK_mo|={ek:.2e}", flush=True)
    assert ej <= TOL and ek <= TOL
