"""GPU parity of the HIP J/K kernels against the numpy oracle (oracle/jk.py).

Plain form: FP64 J and K must match the oracle to <= 1e-10 max-abs (the bar of
the plain FP64 form).  Shapes cover the water and (H2O)_16 workloads, odd nbf
(scalar loads), nbf not a multiple of the 128 tile, nocc not a multiple of 16,
several densities per call (UHF), column-slice C with a row stride, and the
FP32 precision form (relative tolerance stated per test).
"""

import numpy as np
import pytest
import torch

from forte2_amd.jkbuilder import FockBuilder
from forte2_amd.workloads import orthonormal_C, synthetic_B
from oracle import jk as ojk

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-10


def _case(dev, Q, N, noccs, seed):
    print(f"[jk-gpu] Q={Q} N={N} nocc={noccs} seed={seed}", flush=True)
    B = synthetic_B(Q, N, seed=seed, device=dev)
    Cs = [orthonormal_C(N, o, seed=seed + i, device=dev) for i, o in enumerate(noccs)]
    Bh = B.cpu().numpy()
    Ch = [c.cpu().numpy() for c in Cs]
    return B, Cs, Bh, Ch


@pytest.mark.parametrize(
    "Q,N,noccs",
    [
        (139, 24, [5]),           # water
        (100, 19, [5]),           # odd nbf: scalar J loads, scalar KX loads
        (300, 130, [17, 13]),     # two densities, nbf just over one tile
        (50, 64, [7, 3, 20]),     # three densities (J processed 2 + 1)
        (2224, 384, [80]),        # (H2O)_16
        (2224, 384, [80, 79]),    # (H2O)_16, UHF-like
    ],
)
def test_build_jk_fp64_matches_oracle(dev, Q, N, noccs):
    B, Cs, Bh, Ch = _case(dev, Q, N, noccs, seed=Q + N)
    fb = FockBuilder(B=B, device=dev)
    J, K = fb.build_JK(Cs)
    torch.cuda.synchronize()
    Jr, Kr = ojk.build_JK(Bh, Ch)
    for k in range(len(noccs)):
        ej = np.abs(J[k].cpu().numpy() - Jr[k]).max()
        ek = np.abs(K[k].cpu().numpy() - Kr[k]).max()
        print(f"   dens {k}: |dJ|={ej:.2e} |dK|={ek:.2e} max|J|={np.abs(Jr[k]).max():.2e}", flush=True)
        assert ej <= TOL_FP64
        assert ek <= TOL_FP64


def test_build_j_and_k_separately(dev):
    Q, N = 400, 96
    B, Cs, Bh, Ch = _case(dev, Q, N, [11, 6], seed=5)
    fb = FockBuilder(B=B, device=dev)
    D = [c @ c.T for c in Cs]
    J = fb.build_J(D)
    K = fb.build_K(Cs)
    Jr = ojk.build_J(Bh, [c @ c.T for c in Ch])
    Kr = ojk.build_K_nPm(Bh, Ch)
    for k in range(2):
        assert np.abs(J[k].cpu().numpy() - Jr[k]).max() <= TOL_FP64
        assert np.abs(K[k].cpu().numpy() - Kr[k]).max() <= TOL_FP64


def test_numpy_column_slice_inputs(dev):
    # the reference passes C[:, :nocc] slices of an (nbf, nmo) array (rhf.py:36)
    Q, N, nmo, na = 139, 24, 24, 5
    B = synthetic_B(Q, N, seed=3, device=dev)
    Cfull = orthonormal_C(N, nmo, seed=3, device="cpu").numpy()
    fb = FockBuilder(B=B.cpu().numpy(), device=dev)
    J, K = fb.build_JK([Cfull[:, :na]])
    assert isinstance(J[0], np.ndarray) and isinstance(K[0], np.ndarray)
    Jr, Kr = ojk.build_JK(B.cpu().numpy(), [Cfull[:, :na]])
    assert np.abs(J[0] - Jr[0]).max() <= TOL_FP64
    assert np.abs(K[0] - Kr[0]).max() <= TOL_FP64


def test_small_xchunk_and_split_paths(dev):
    # force several X chunks and a large split count: same result bit for bit
    Q, N = 700, 200
    B, Cs, Bh, Ch = _case(dev, Q, N, [33], seed=9)
    fb = FockBuilder(B=B, device=dev)
    J0, K0 = fb.build_JK(Cs)
    fb2 = FockBuilder(B=B, device=dev)
    fb2.set_tuning("xchunk_bytes", 1 << 20)
    J1, K1 = fb2.build_JK(Cs)
    Kr = ojk.build_K_nPm(Bh, Ch)[0]
    assert np.abs(K1[0].cpu().numpy() - Kr).max() <= TOL_FP64
    assert torch.equal(J0[0], J1[0])
    # repeated builds are bitwise reproducible
    J2, K2 = fb.build_JK(Cs)
    assert torch.equal(J0[0], J2[0]) and torch.equal(K0[0], K2[0])


def test_fp32_precision_form(dev):
    Q, N = 2224, 384
    B, Cs, Bh, Ch = _case(dev, Q, N, [80], seed=1)
    fb = FockBuilder(B=B, device=dev)
    J, K = fb.build_JK(Cs, precision="fp32")
    Jr, Kr = ojk.build_JK(Bh, Ch)
    rj = np.abs(J[0].cpu().numpy() - Jr[0]).max() / np.abs(Jr[0]).max()
    rk = np.abs(K[0].cpu().numpy() - Kr[0]).max() / np.abs(Kr[0]).max()
    print(f"   fp32: rel|dJ|={rj:.2e} rel|dK|={rk:.2e}", flush=True)
    # FP32 operands (unit roundoff 6e-8) with FP64 accumulation across P
    assert rj < 1e-6
    assert rk < 1e-5


@pytest.mark.parametrize(
    "Q,N,noccs",
    [
        (2224, 384, [80]),        # (H2O)_16
        (300, 130, [17, 13]),     # two densities, nbf over one tile, nocc not a multiple of 32
        (100, 19, [5]),           # odd nbf: scalar FP32 loads in the split path
        (500, 256, [330]),        # more than 320 occupied columns: two column tiles
    ],
)
@pytest.mark.parametrize("ks_tile", [128, 256])
def test_bf16x3_precision_form(dev, Q, N, noccs, ks_tile):
    # ks_tile: the contraction's lower-tile edge (256: one 8-wave workgroup
    # per CU, chosen by default for nbf >= 768; forced here at every size,
    # including a single partial tile and odd nbf)
    B, Cs, Bh, Ch = _case(dev, Q, N, noccs, seed=17)
    fb = FockBuilder(B=B, device=dev)
    fb.set_tuning("ks_tile", ks_tile)
    J, K = fb.build_JK(Cs, precision="bf16x3")
    Jr, Kr = ojk.build_JK(Bh, Ch)
    for k in range(len(noccs)):
        rj = np.abs(J[k].cpu().numpy() - Jr[k]).max() / np.abs(Jr[k]).max()
        rk = np.abs(K[k].cpu().numpy() - Kr[k]).max() / np.abs(Kr[k]).max()
        print(f"   bf16x3 tile {ks_tile} dens {k}: rel|dJ|={rj:.2e} rel|dK|={rk:.2e}", flush=True)
        # J streams the FP32 copy of B (FP64 accumulation).  K: each operand is
        # hi + lo with bf16 hi, lo (unit roundoff u = 2^-8), so the operand
        # residual and the dropped lo*lo product are each <= u^2 = 1.5e-5 of a
        # product; summed with uncorrelated rounding the error stays below
        # u^2 of max|K|.  A lost cross term (~u = 4e-3) fails the bound.
        assert rj < 1e-6
        assert rk < 2e-5
        assert torch.equal(K[k], K[k].T)


@pytest.mark.parametrize(
    "Q,N,noccs,xchunk",
    [
        (2224, 384, [80], 0),            # (H2O)_16
        (300, 132, [17, 13], 0),         # two densities, a partial last 256-row tile
        (500, 256, [330], 0),            # two column tiles of 192 (zero coefficient rows past 352)
        (1112, 192, [40], 64 << 20),     # several X chunks
    ],
)
def test_bf16x3_tall_half_transform(dev, Q, N, noccs, xchunk):
    # The 256-row half-transform (default for nbf >= 768, forced here) runs
    # the same MFMA sequence per output element as the 128-row kernel (k
    # steps in order, the three products in the same order), so K is bitwise
    # the same; both are checked against the oracle at the split bound.
    B, Cs, Bh, Ch = _case(dev, Q, N, noccs, seed=23)
    Jr, Kr = ojk.build_JK(Bh, Ch)
    out = {}
    for rows in (128, 256):
        fb = FockBuilder(B=B, device=dev)
        fb.set_tuning("kx_rows", rows)
        if xchunk:
            fb.set_tuning("xchunk_bytes", xchunk)
        J, K = fb.build_JK(Cs, precision="bf16x3")
        out[rows] = K
        for k in range(len(noccs)):
            rk = np.abs(K[k].cpu().numpy() - Kr[k]).max() / np.abs(Kr[k]).max()
            print(f"   bf16x3 {rows}-row half-transform dens {k}: rel|dK|={rk:.2e}", flush=True)
            assert rk < 2e-5
        fb.release()
    for k in range(len(noccs)):
        assert torch.equal(out[128][k], out[256][k])


@pytest.mark.parametrize(
    "Q,N,noccs",
    [
        (600, 200, [150]),          # one wide tile, 160 of 320 columns occupied
        (300, 132, [100, 90]),      # two densities side by side (208 columns)
        (300, 240, [200, 150]),     # two wide column tiles (368 columns), the second partial
        (200, 197, [160]),          # odd nbf (one element per A load), a k tail
    ],
)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fp64_fp32_wide_half_transform(dev, Q, N, noccs, prec):
    # The 320-column half-transform (default above 128 occupied columns) runs
    # the MFMA sequence of the 128-column kernel per output element (k steps
    # in order): K bitwise the same, and FP64 within 1e-10 of the oracle.
    B, Cs, Bh, Ch = _case(dev, Q, N, noccs, seed=31)
    Jr, Kr = ojk.build_JK(Bh, Ch)
    out = {}
    for wide in (0, 1):
        fb = FockBuilder(B=B, device=dev)
        fb.set_tuning("kx_wide", wide)
        J, K = fb.build_JK(Cs, precision=prec)
        out[wide] = K
        for k in range(len(noccs)):
            dk = np.abs(K[k].cpu().numpy() - Kr[k]).max()
            print(f"   {prec} kx_wide={wide} dens {k}: |dK|={dk:.2e}", flush=True)
            if prec == "fp64":
                assert dk <= 1e-10
            else:
                assert dk <= 1e-5 * np.abs(Kr[k]).max()
        fb.release()
    for k in range(len(noccs)):
        assert torch.equal(out[0][k], out[1][k])


@pytest.mark.parametrize(
    "Q,N,noccs",
    [
        (2224, 384, [80]),        # (H2O)_16
        (300, 130, [17, 13]),     # two densities, rows past nbf clamped in the last tile
        (100, 19, [5]),           # odd nbf, one partial tile
        (500, 256, [330]),        # two column tiles in the half-transform
    ],
)
@pytest.mark.parametrize("ks_tile", [128, 256])
def test_bf16x3_contraction_by_direct_to_lds_loads(dev, Q, N, noccs, ks_tile):
    # The contraction with its X tiles loaded straight into LDS runs the same
    # tiles, splits and MFMA sequence as the register-staged one: K bitwise
    # the same, and within the split bound of the oracle.
    B, Cs, Bh, Ch = _case(dev, Q, N, noccs, seed=29)
    Jr, Kr = ojk.build_JK(Bh, Ch)
    out = {}
    for dma in (0, 1):
        fb = FockBuilder(B=B, device=dev)
        fb.set_tuning("ks_tile", ks_tile)
        fb.set_tuning("ks_dma", dma)
        J, K = fb.build_JK(Cs, precision="bf16x3")
        out[dma] = K
        fb.release()
    for k in range(len(noccs)):
        rk = np.abs(out[1][k].cpu().numpy() - Kr[k]).max() / np.abs(Kr[k]).max()
        print(f"   bf16x3 contraction tile {ks_tile} by direct-to-LDS loads, dens {k}: rel|dK|={rk:.2e}", flush=True)
        assert rk < 2e-5
        assert torch.equal(out[0][k], out[1][k])


def test_error_paths(dev):
    B = synthetic_B(10, 8, seed=0, device=dev)
    fb = FockBuilder(B=B, device=dev)
    with pytest.raises(ValueError):
        fb.build_JK([torch.zeros((9, 2), dtype=torch.float64, device=dev)])
    with pytest.raises(ValueError):
        fb.build_JK([torch.zeros((8, 2), dtype=torch.float64, device=dev)] * 9)
    with pytest.raises(ValueError):
        fb.build_JK([torch.zeros((8, 2), dtype=torch.float64, device=dev)], precision="fp16")
    # launch-shape knobs outside their documented values are refused
    for key, value in (("kx_wide", 2), ("kx_rows", 64), ("ks_tile", 96), ("ks_dma", -2), ("no_such_knob", 1)):
        with pytest.raises(ValueError):
            fb.set_tuning(key, value)


def test_edge_shapes(dev):
    # an empty density (UHF with no electron of one spin), nocc = nbf, and a
    # single auxiliary function
    for Q, N, noccs in ((60, 40, [0, 7]), (60, 40, [40]), (1, 33, [5])):
        B, Cs, Bh, Ch = _case(dev, Q, N, noccs, seed=Q + N + len(noccs))
        if noccs[0] == 0:
            Cs[0] = torch.zeros((N, 0), dtype=torch.float64, device=dev)
            Ch[0] = np.zeros((N, 0))
        J, K = FockBuilder(B=B, device=dev).build_JK(Cs)
        Jr, Kr = ojk.build_JK(Bh, Ch)
        for k in range(len(noccs)):
            assert np.abs(J[k].cpu().numpy() - Jr[k]).max() <= TOL_FP64
            assert np.abs(K[k].cpu().numpy() - Kr[k]).max() <= TOL_FP64
        if noccs[0] == 0:
            assert float(J[0].abs().max()) == 0.0 and float(K[0].abs().max()) == 0.0


def _screened(B, N, group=8, seed=3):
    """B zeroed, for every P, on the (m, n) blocks of 'distant' function
    groups (a symmetric random pattern of group x group blocks, about 70% of
    them, diagonal blocks kept): the structure of a molecule's B on its
    screened shell pairs.  Returns the mask of kept (m, n)."""
    rng = np.random.default_rng(seed)
    ng = -(-N // group)
    keep = rng.random((ng, ng)) < 0.3
    keep = keep | keep.T | np.eye(ng, dtype=bool)
    mask = np.kron(keep, np.ones((group, group), dtype=bool))[:N, :N]
    B.mul_(torch.from_numpy(mask).to(B.device, B.dtype))
    return mask


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize(
    "Q,N,noccs",
    [
        (300, 132, [17, 13]),     # two densities, FP32 J in 4-wide elements
        (200, 96, [20]),
        (120, 31, [6]),           # odd nbf: 1-wide elements
    ],
)
def test_j_over_nonzero_elements(dev, prec, Q, N, noccs):
    # J through the list of lower-triangle elements where B is nonzero for some
    # P equals J over every element (FP64 to rounding of the gamma sums, the
    # precision forms to their J bound) and the oracle, and is exactly zero
    # where B is zero for every P.  K does not use the list: bitwise equal.
    B, Cs, _, Ch = _case(dev, Q, N, noccs, seed=41)
    mask = _screened(B, N)
    Bh = B.cpu().numpy()
    Jr, Kr = ojk.build_JK(Bh, Ch)
    out = {}
    for sparse in (False, True):
        fb = FockBuilder(B=B, device=dev, j_sparse=sparse)
        out[sparse] = fb.build_JK(Cs, precision=prec)
        # build_J from densities (f2jk_build_j) through the same list
        out[(sparse, "D")] = fb.build_J([c @ c.T for c in Cs], precision=prec)
        fb.release()
    for k in range(len(noccs)):
        Js, Jd = out[True][0][k], out[False][0][k]
        jr = Jr[k]
        err = float(np.abs(Js.cpu().numpy() - jr).max())
        rel = err / float(np.abs(jr).max())
        print(f"   {prec} J over nonzero elements, dens {k}: |dJ|={err:.2e} rel {rel:.2e}", flush=True)
        if prec == "fp64":
            assert err <= TOL_FP64
            assert float((Js - Jd).abs().max()) <= 1e-12 * float(Jd.abs().max())
        else:
            assert rel < 1e-6
        assert bool((Js.cpu().numpy()[~mask] == 0.0).all())
        assert torch.equal(out[True][1][k], out[False][1][k])
        Jds = out[(True, "D")][k]
        assert float((Jds - Js).abs().max()) <= 1e-12 * float(Js.abs().max()) + (0.0 if prec == "fp64" else 1e-7)


def test_j_over_nonzero_elements_real_h2o16(dev):
    # the (H2O)_16 cluster of the SCF tests (cc-pVDZ / cc-pVTZ-JKFIT, seed 16,
    # 9 A box): its B formed on the device is exactly zero on the screened
    # shell pairs, and J read through the nonzero-element list matches the
    # oracle on the same B to 1e-10
    from forte2_amd import geometry as geo
    from forte2_amd.system import System

    s = System(geo.water_cluster(16, 9.0, seed=16), "cc-pvdz", "cc-pvtz-jkfit", device=dev)
    fb = s.fock_builder
    B = fb.B_Pmn
    N = s.nbf
    C = orthonormal_C(N, 80, seed=16, device=dev)
    J, K = fb.build_JK([C])
    Bh = B.cpu().numpy()
    zero = ~(Bh != 0.0).any(axis=0)
    print(f"   (H2O)_16: B zero for every P on {zero.mean():.1%} of (m, n)", flush=True)
    Jr, Kr = ojk.build_JK(Bh, [C.cpu().numpy()])
    ej = float(np.abs(J[0].cpu().numpy() - Jr[0]).max())
    ek = float(np.abs(K[0].cpu().numpy() - Kr[0]).max())
    print(f"   |dJ|={ej:.2e} |dK|={ek:.2e}", flush=True)
    assert zero.mean() > 0.2
    assert ej <= TOL_FP64 and ek <= TOL_FP64
    assert bool((J[0].cpu().numpy()[zero] == 0.0).all())
    fb.release()
    s.fock_builder = None
