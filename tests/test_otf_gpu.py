"""GPU parity of the integral-direct builder (forte2_amd/otf.py) against the
integral-direct oracle (oracle/otf.py) and the in-core device builder, FP64
<= 1e-10 max-abs.  Cases follow the reference's own OTF tests
(tests/jkbuilder/test_jkbuilder.py:192-283): N2 in cc-pVQZ / cc-pVQZ-JKFIT at
a 4.5 MB budget (many aux blocks and occupied sweeps), J, K and the combined
build, two densities per call, the generalized Fock, the MO-basis blocks,
the truncated-eigen metric, the (P|mn) block cache, and an RHF run through
the System factory that ends on the reference's golden energy.
"""

import json
import os

import numpy as np
import pytest
import torch

from forte2_amd.jkbuilder import FockBuilder
from forte2_amd.otf import FockBuilderOTF
from forte2_amd.system import System
from oracle import otf as ootf

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_values.json")))
EREF = GOLD["energies"]

pytestmark = pytest.mark.gpu

TOL = 1e-10
N2 = "N 0.0 0.0 0.0\nN 0.0 0.0 2.0"


def _np(x):
    return x.cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


def _err(a, b):
    return float(np.abs(_np(a) - _np(b)).max())


def _orth(n, k, seed):
    return np.linalg.qr(np.random.default_rng(seed).standard_normal((n, k)))[0]


@pytest.fixture(scope="module")
def n2qz(dev):
    print("[otf-gpu] N2 cc-pVQZ / cc-pVQZ-JKFIT, 4.5 MB", flush=True)
    s = System(N2, "cc-pvqz", "cc-pvqz-jkfit", unit="bohr", device=dev)
    fb = FockBuilderOTF(s, jk_mem_thres_mb=4.5, device=dev)
    met = ootf.Metric(s.coulomb_2c())
    T = s.coulomb_3c()
    blocks = [(p0, p1) for _, _, p0, p1 in fb.blocks]
    return s, fb, met, T, blocks


def test_block_shapes(n2qz):
    s, fb, *_ = n2qz
    assert (s.nbf, s.naux) == (110, 212)
    assert (fb.iblksize, fb.pblksize) == (7, 21)
    assert len(fb.blocks) > 10


def test_jk_matches_oracle(n2qz):
    s, fb, met, T, blocks = n2qz
    C = _orth(s.nbf, 50, 12345)
    D = C @ C.T
    Jr = ootf.J_kernel(met, T, blocks, D)
    Kr = ootf.K_kernel(met, T, blocks, C, fb.iblksize)
    J = fb.build_J([D])[0]
    K = fb.build_K([C])[0]
    print(f"   |dJ|={_err(J, Jr):.2e} |dK|={_err(K, Kr):.2e}", flush=True)
    assert _err(J, Jr) <= TOL and _err(K, Kr) <= TOL
    J2, K2 = fb.build_JK([C])
    assert _err(J2[0], Jr) <= TOL and _err(K2[0], Kr) <= TOL
    # the in-core device builder agrees
    fi = FockBuilder(s, device=s.device)
    Ji, Ki = fi.build_JK([C])
    assert _err(J2[0], Ji[0]) <= TOL and _err(K2[0], Ki[0]) <= TOL
    fi.release()


def test_two_densities_share_sweeps(n2qz):
    s, fb, met, T, blocks = n2qz
    Ca, Cb = _orth(s.nbf, 11, 1), _orth(s.nbf, 9, 2)
    J, K = fb.build_JK([Ca, Cb[:, :9]])
    for k, C in enumerate((Ca, Cb)):
        Jr, Kr = ootf.JK_kernel(met, T, blocks, C, fb.iblksize)
        assert _err(J[k], Jr) <= TOL and _err(K[k], Kr) <= TOL
    # device tensors in, device tensors out; a column slice with a row stride
    wide = torch.from_numpy(_orth(s.nbf, 20, 3)).to(s.device)
    Kd = fb.build_K([wide[:, 2:15]])[0]
    assert isinstance(Kd, torch.Tensor)
    assert _err(Kd, ootf.K_kernel(met, T, blocks, _np(wide)[:, 2:15], fb.iblksize)) <= TOL


def test_generalized_fock_and_mo_blocks(n2qz):
    s, fb, met, T, blocks = n2qz
    rng = np.random.default_rng(12345)
    C = rng.standard_normal((s.nbf, s.nbf)) * 0.1
    Cact = C[:, 12:24]
    g = rng.standard_normal((12, 12))
    g1 = (g + g.T) @ (g + g.T).T
    fi = FockBuilder(s, device=s.device)
    J, K = fb.build_JK_generalized(Cact, g1)
    Ji, Ki = fi.build_JK_generalized(Cact, g1)
    assert _err(J, Ji) <= TOL * max(1.0, np.abs(Ji).max()) and _err(K, Ki) <= TOL * max(1.0, np.abs(Ki).max())
    C1, C2, C3, C4 = (_orth(s.nbf, k, 20 + k) for k in (2, 3, 4, 5))
    V = fb.two_electron_integrals_gen_block(C1, C2, C3, C4)
    Vi = fi.two_electron_integrals_gen_block(C1, C2, C3, C4)
    assert V.shape == (2, 3, 4, 5) and _err(V, Vi) <= TOL
    Bmo = fb.B_tensor_gen_block(C1, C2)
    assert _err(Bmo, ootf.B_tensor_gen_block(met, T, blocks, C1, C2)) <= TOL
    fi.release()


def test_block_cache_gives_same_results(dev):
    s = System(N2, "cc-pvdz", "cc-pvtz-jkfit", unit="bohr", device=dev)
    C = _orth(s.nbf, 7, 5)
    plain = FockBuilderOTF(s, jk_mem_thres_mb=0.4, device=dev)
    cached = FockBuilderOTF(s, jk_mem_thres_mb=0.4, cache_mb=0.4, device=dev)
    assert len(plain.blocks) > 2
    J0, K0 = plain.build_JK([C])
    J1, K1 = cached.build_JK([C])
    first = cached.host_blocks_computed
    J2, K2 = cached.build_JK([C])
    # part of (P|mn) stays resident: the second build evaluates fewer blocks on the host
    assert cached.host_blocks_computed - first < first
    for a, b in ((J0, J1), (K0, K1), (J1, J2), (K1, K2)):
        assert np.array_equal(a[0], b[0])


def test_single_block_is_evaluated_once_per_build(dev):
    # with one aux block the J sweep reuses the block the K/gamma sweep left
    # on the device; every build still evaluates it once
    s = System(N2, "cc-pvdz", "cc-pvtz-jkfit", unit="bohr", device=dev)
    fb = FockBuilderOTF(s, jk_mem_thres_mb=4000, device=dev)
    assert len(fb.blocks) == 1
    C = _orth(s.nbf, 7, 8)
    J1, K1 = fb.build_JK([C])
    assert fb.host_blocks_computed == 1
    J2, K2 = fb.build_JK([C])
    assert fb.host_blocks_computed == 2
    assert np.array_equal(J1[0], J2[0]) and np.array_equal(K1[0], K2[0])
    fi = FockBuilder(s, device=s.device)
    Ji, Ki = fi.build_JK([C])
    assert _err(J1[0], Ji[0]) <= TOL and _err(K1[0], Ki[0]) <= TOL
    Jd = fb.build_J([C @ C.T])[0]  # J alone: two sweeps, one evaluation
    assert fb.host_blocks_computed == 3 and _err(Jd, Ji[0]) <= TOL
    fi.release()


def test_eigen_metric_branch(dev):
    s = System(N2, "cc-pvdz", "cc-pvtz-jkfit", unit="bohr", df_ortho_rtol=1e-10, device=dev)
    fb = FockBuilderOTF(s, jk_mem_thres_mb=0.5, device=dev)
    met = ootf.Metric(s.coulomb_2c(), 1e-10)
    T = s.coulomb_3c()
    blocks = [(p0, p1) for _, _, p0, p1 in fb.blocks]
    C = _orth(s.nbf, 7, 6)
    J, K = fb.build_JK([C])
    Jr, Kr = ootf.JK_kernel(met, T, blocks, C, fb.iblksize)
    assert _err(J[0], Jr) <= TOL and _err(K[0], Kr) <= TOL


def test_rhf_through_factory(dev):
    # tests/orbitals/test_cube.py:15-23 golden energy, with B never stored
    from forte2_amd.scf import RHF

    water = ("O 0.000000000000 0.000000000000 -0.061664597388\n"
             "H 0.000000000000 -0.711620616369 0.489330954643\n"
             "H 0.000000000000 0.711620616369 0.489330954643")
    s = System(water, "cc-pvdz", "cc-pvtz-jkfit", device=dev, jk_mem_thres_mb=0.4)
    assert isinstance(s.fock_builder, FockBuilderOTF)
    scf = RHF(charge=0, e_tol=1e-10)(s).run()
    print(f"   E = {scf.E:.12f}", flush=True)
    assert scf.E == pytest.approx(EREF["water_dz_rhf"]["E"], abs=5e-8)


def test_cl2_golden_through_otf(dev):
    # tests/scf/test_rhf.py:97-117: Cl2 cc-pVQZ / cc-pVQZ-JKFIT with the
    # integral-direct builder assigned to the system at a 10 MB budget
    from forte2_amd.scf import RHF

    print("[otf-gpu] Cl2 cc-pVQZ / cc-pVQZ-JKFIT, 10 MB", flush=True)
    s = System("Cl 0 0 0\nCl 0 0 1.681", "cc-pvqz", "cc-pvqz-jkfit", device=dev)
    s.fock_builder = FockBuilderOTF(s, jk_mem_thres_mb=10, backend="libcint", device=dev)
    assert len(s.fock_builder.blocks) > 1
    scf = RHF(charge=0)(s).run()
    print(f"   E = {scf.E:.12f}", flush=True)
    assert scf.E == pytest.approx(EREF["cl2_qz_rhf"]["E"], abs=5e-8)


def test_errors(dev):
    s = System(N2, "cc-pvdz", "cc-pvtz-jkfit", unit="bohr", device=dev)
    with pytest.raises(ValueError, match="too low"):
        FockBuilderOTF(s, jk_mem_thres_mb=0.01, device=dev)
    fb = FockBuilderOTF(s, jk_mem_thres_mb=1, device=dev)
    with pytest.raises(ValueError, match="fp64"):
        fb.build_JK([_orth(s.nbf, 3, 1)], precision="bf16x3")
    with pytest.raises(ValueError):
        fb.build_K([np.zeros((s.nbf - 1, 2))])
