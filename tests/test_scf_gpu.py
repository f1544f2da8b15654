"""Device SCF (forte2_amd.scf) against the SCF oracle and the reference's golden values.

Plain form: the device RHF/UHF energy equals the oracle's to <= 1e-10 Eh on the
same integrals and guess (water; (H2O)_16; an alkane cation).  Precision form:
the final energy equals the device FP64 energy to <= 1e-8 Eh.
"""

import json
import os

import numpy as np
import pytest
import torch

from forte2_amd import geometry as geo
from forte2_amd.scf import RHF, UHF
from forte2_amd.system import System
from oracle import jk as ojk
from oracle import scf as oscf

from golden_systems import MORE_RHF

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_values.json")))
EREF = GOLD["energies"]

pytestmark = pytest.mark.gpu

WATER = """O            0.000000000000     0.000000000000    -0.061664597388
H            0.000000000000    -0.711620616369     0.489330954643
H            0.000000000000     0.711620616369     0.489330954643"""
E_WATER_DZ = EREF["water_dz_rhf"]["E"]       # tests/orbitals/test_cube.py:15
E_WATER_QZ_UHF = EREF["water_qz_uhf_triplet"]["E"]     # tests/scf/test_uhf.py:11
S2_WATER_QZ_UHF = EREF["water_qz_uhf_triplet"]["S2"]


def _oracle(s, ms=None, charge=0, e_tol=1e-10, d_tol=1e-6):
    B = ojk.build_B_density_fitting(s.coulomb_2c(), s.coulomb_3c())
    args = (s.ints_overlap(), s.ints_hcore(), s.nuclear_repulsion, oscf.DFJK(B), s.Zsum - charge)
    kw = dict(X=s.Xorth, V_sap=s.sap_potential(), e_tol=e_tol, d_tol=d_tol)
    if ms is None:
        return oscf.run_rhf(*args, **kw), B
    return oscf.run_uhf(*args, ms, **kw), B


def test_device_B_matches_oracle_B(dev):
    s = System(WATER, "cc-pvdz", "cc-pvtz-jkfit", device=dev)
    Bd = s.fock_builder.B_Pmn.cpu().numpy()
    Bo = ojk.build_B_density_fitting(s.coulomb_2c(), s.coulomb_3c())
    assert np.abs(Bd - Bo).max() < 1e-12


def test_device_B_pair_blocks_match_oracle_B(dev):
    # many pair blocks (1 MiB each): pipeline boundaries, exact m/n symmetry
    from forte2_amd.df import build_B_density_fitting

    s = System(geo.water_cluster(4, seed=4), "cc-pvdz", "cc-pvtz-jkfit", device=dev)
    st = {}
    Bd = build_B_density_fitting(s, dev, block_bytes=1 << 20, stats=st)
    assert st["blocks"] > 10
    Bo = ojk.build_B_density_fitting(s.coulomb_2c(), s.coulomb_3c())
    Bh = Bd.cpu().numpy()
    print(f"   |dB| = {np.abs(Bh - Bo).max():.2e} over {st['blocks']} blocks", flush=True)
    assert np.abs(Bh - Bo).max() < 1e-11
    assert np.array_equal(Bh, Bh.transpose(0, 2, 1))


def test_rhf_water_plain_form(dev):
    print("[scf-gpu] water cc-pVDZ/cc-pVTZ-JKFIT RHF", flush=True)
    s = System(WATER, "cc-pvdz", "cc-pvtz-jkfit", device=dev)
    scf = RHF(charge=0, e_tol=1e-10)(s).run()
    ref, _ = _oracle(s)
    print(f"   device {scf.E:.12f} oracle {ref.E:.12f} diff {scf.E - ref.E:.2e} iters {scf.iter}/{ref.iter}", flush=True)
    assert abs(scf.E - ref.E) <= 1e-10
    assert scf.E == pytest.approx(E_WATER_DZ, abs=5e-8)
    assert scf.converged


def test_rhf_h2o16_plain_form(dev):
    print("[scf-gpu] (H2O)_16 cc-pVDZ/cc-pVTZ-JKFIT RHF", flush=True)
    s = System(geo.water_cluster(16, 9.0, seed=16), "cc-pvdz", "cc-pvtz-jkfit", device=dev)
    assert (s.nbf, s.naux) == (384, 2224)
    scf = RHF(charge=0, e_tol=1e-10)(s).run()
    ref, _ = _oracle(s)
    print(f"   device {scf.E:.12f} oracle {ref.E:.12f} diff {scf.E - ref.E:.2e} iters {scf.iter}/{ref.iter} "
          f"time {scf.time:.2f}s", flush=True)
    assert abs(scf.E - ref.E) <= 1e-10


def test_rhf_naphthalene_tz_golden(dev):
    # tests/scf/test_rhf_long.py:8-36 (C10H8 cc-pVTZ/cc-pVTZ-JKFIT, nbf 412,
    # naux 1030; a slow test there): device FP64 against the golden energy
    # and the oracle, the precision form against FP64
    xyz = open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "naphthalene.xyz")).read()
    print("[scf-gpu] C10H8 cc-pVTZ/cc-pVTZ-JKFIT RHF", flush=True)
    s = System(xyz, "cc-pvtz", "cc-pvtz-jkfit", device=dev)
    assert (s.nbf, s.naux) == (412, 1030)
    scf = RHF(charge=0, e_tol=1e-10)(s).run()
    ref, _ = _oracle(s)
    mix = RHF(charge=0, e_tol=1e-10, precision="bf16x3")(s).run()
    print(f"   device {scf.E:.12f} oracle {ref.E:.12f} golden {EREF['naphthalene_tz_rhf']['E']:.12f} "
          f"bf16x3 {mix.E:.12f} iters {scf.iter}/{ref.iter}/{mix.iter}", flush=True)
    assert scf.converged and mix.converged
    assert scf.E == pytest.approx(EREF["naphthalene_tz_rhf"]["E"], abs=5e-8)
    assert abs(scf.E - ref.E) <= 1e-10
    assert abs(mix.E - scf.E) <= 1e-8
    s.fock_builder.release()


@pytest.mark.parametrize("xyz,basis,aux,unit,key,kw", MORE_RHF, ids=[c[4] for c in MORE_RHF])
def test_more_rhf_golden(dev, xyz, basis, aux, unit, key, kw):
    # RHF energies asserted by the reference's DSRG and CASSCF tests
    kw = dict(kw)
    charge = kw.pop("charge", 0)
    s = System(xyz, basis, aux, unit=unit, device=dev)
    scf = RHF(charge=charge, **kw)(s).run()
    print(f"[scf-gpu] {key}: device {scf.E:.12f} golden {EREF[key]['E']:.12f}", flush=True)
    assert scf.converged
    assert scf.E == pytest.approx(EREF[key]["E"], abs=5e-8)


def test_uhf_water_cation_golden(dev):
    # tests/scf/test_ghf.py:24-49 (its UHF part)
    g = EREF["water_cation_qz_uhf"]
    s = System(WATER, "cc-pvqz", "cc-pvqz-jkfit", device=dev)
    scf = UHF(charge=1, ms=0.5)(s).run()
    print(f"[scf-gpu] water cation cc-pVQZ UHF: device {scf.E:.12f} <S2> {scf.S2:.10f} golden {g['E']:.12f}",
          flush=True)
    assert scf.E == pytest.approx(g["E"], abs=5e-8)
    assert scf.S2 == pytest.approx(g["S2"], abs=5e-8)


def test_linearly_dependent_metric_raises_on_the_device(dev):
    # tests/jkbuilder/test_jkbuilder.py:392-415
    from forte2_amd.basis import BasisSet
    from forte2_amd.otf import FockBuilderOTF

    s = System("H 0 0 0\nH 0 0 1.0", "cc-pvdz", "cc-pvtz-jkfit", unit="bohr", device=dev)
    fake = BasisSet("two identical s functions")
    fake.add(0, [1.0], [1.0], [0.0, 0.0, 0.0])
    fake.add(0, [1.0], [1.0], [0.0, 0.0, 0.0])
    s.auxiliary_basis, s.naux = fake, fake.size
    with pytest.raises(ValueError, match="positive definite"):
        s.fock_builder.B_Pmn
    with pytest.raises(ValueError, match="positive definite"):
        s.fock_builder = FockBuilderOTF(s, jk_mem_thres_mb=10, device=dev)


def test_rhf_linear_dependence_golden(dev):
    # tests/scf/test_lindep.py:9-30: cond(S) > 1e14, 79 of 90 orbitals kept
    # (a rectangular orthogonaliser on the device), tight d_tol
    g = EREF["h10_augdz_rhf_lindep"]
    print("[scf-gpu] H10 chain aug-cc-pVDZ/cc-pVQZ-JKFIT RHF, linear dependence", flush=True)
    s = System("\n".join(f"H 0 0 {i}" for i in range(10)), "aug-cc-pvdz", "cc-pvqz-jkfit", unit="bohr",
               overlap_ortho_rtol=g["overlap_ortho_rtol"], device=dev)
    assert (s.nbf, s.nmo) == (g["nbf"], g["nmo"])
    scf = RHF(charge=0, e_tol=1e-10, d_tol=1e-8)(s).run()
    ref, _ = _oracle(s, d_tol=1e-8)
    print(f"   device {scf.E:.12f} oracle {ref.E:.12f} golden {g['E']:.12f}", flush=True)
    assert scf.converged
    assert scf.E == pytest.approx(g["E"], abs=5e-8)
    assert abs(scf.E - ref.E) <= 1e-10


def test_rhf_zero_virtuals_and_zero_electrons(dev):
    # tests/scf/test_rhf.py:25-44
    s = System("Ne 0 0 0", "sto-3g", "def2-universal-jkfit", device=dev)
    scf = RHF(charge=0)(s).run()
    assert scf.E == pytest.approx(EREF["ne_sto3g_rhf_zero_virtuals"]["E"], abs=5e-8)
    s = System("H 0 0 0\nH 0 0 1.0", "cc-pvqz", "cc-pvqz-jkfit", device=dev)
    scf = RHF(charge=2)(s).run()
    print(f"[scf-gpu] Ne STO-3G {EREF['ne_sto3g_rhf_zero_virtuals']['E']:.12f}; H2++ {scf.E:.12f} "
          f"(Vnn {s.nuclear_repulsion:.12f})", flush=True)
    assert scf.E == pytest.approx(s.nuclear_repulsion, abs=5e-8)


def test_rhf_level_shift_golden(dev):
    # tests/scf/test_rhf.py:62-79: formamide, level shift 0.5, no DIIS
    print("[scf-gpu] formamide cc-pVTZ/cc-pVTZ-JKFIT RHF, level shift 0.5", flush=True)
    xyz = ("C 0.000000 0.418626 0.000000\nH -0.460595 1.426053 0.000000\nO 1.196516 0.242075 0.000000\n"
           "N -0.936579 -0.568753 0.000000\nH -0.634414 -1.530889 0.000000\nH -1.921071 -0.362247 0.000000")
    s = System(xyz, "cc-pvtz", "cc-pvtz-jkfit", device=dev)
    scf = RHF(charge=0, level_shift=0.5, do_diis=False)(s).run()
    print(f"   E = {scf.E:.12f} iters {scf.iter}", flush=True)
    assert scf.E == pytest.approx(EREF["formamide_tz_rhf_level_shift"]["E"], abs=5e-8)
    assert scf.iter < 40


def test_uhf_guess_mix_golden(dev):
    # tests/scf/test_uhf.py:73-90 (Coulson-Fischer point of stretched H2)
    s = System("H 0 0 0\nH 0 0 2.7", "cc-pvtz", "cc-pvqz-jkfit", device=dev)
    u = UHF(charge=0, ms=0.0, guess_mix=True)(s).run()
    assert u.E == pytest.approx(EREF["h2_stretched_tz_uhf_guess_mix"]["E"], abs=5e-8)
    assert u.S2 == pytest.approx(EREF["h2_stretched_tz_uhf_guess_mix"]["S2"], abs=5e-8)


def test_uhf_water_singlet_golden(dev):
    # tests/scf/test_uhf.py:24-40: ms = 0 UHF of the cc-pVQZ water is its RHF
    g = EREF["water_qz_uhf_singlet"]
    s = System(WATER, "cc-pvqz", "cc-pvqz-jkfit", device=dev)
    scf = UHF(charge=0, ms=0.0)(s).run()
    print(f"   E = {scf.E:.12f} <S^2> = {scf.S2:.2e}", flush=True)
    assert scf.E == pytest.approx(g["E"], abs=5e-8)
    assert scf.S2 == pytest.approx(g["S2"], abs=5e-8)


def test_uhf_water_triplet_golden(dev):
    s = System(WATER, "cc-pvqz", "cc-pvqz-jkfit", device=dev)
    scf = UHF(charge=0, ms=1.0, e_tol=1e-10)(s).run()
    assert scf.E == pytest.approx(E_WATER_QZ_UHF, abs=5e-8)
    assert scf.S2 == pytest.approx(S2_WATER_QZ_UHF, abs=5e-8)


def test_uhf_alkane_cation_plain_form(dev):
    print("[scf-gpu] C3H8+ cc-pVDZ/cc-pVTZ-JKFIT UHF", flush=True)
    s = System(geo.alkane(3), "cc-pvdz", "cc-pvtz-jkfit", device=dev)
    scf = UHF(charge=1, ms=0.5, e_tol=1e-10)(s).run()
    ref, _ = _oracle(s, ms=0.5, charge=1)
    print(f"   device {scf.E:.12f} oracle {ref.E:.12f} diff {scf.E - ref.E:.2e} S2 {scf.S2:.8f}/{ref.S2:.8f}",
          flush=True)
    assert abs(scf.E - ref.E) <= 1e-10
    assert abs(scf.S2 - ref.S2) <= 1e-7


@pytest.mark.parametrize("tail", ["reref", "full"])
@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
@pytest.mark.parametrize("n", [1, 16])
def test_precision_form_final_energy(dev, precision, n, tail):
    xyz = WATER if n == 1 else geo.water_cluster(16, 9.0, seed=16)
    s = System(xyz, "cc-pvdz", "cc-pvtz-jkfit", device=dev)
    e64 = RHF(charge=0, e_tol=1e-10)(s).run()
    mix = RHF(charge=0, e_tol=1e-10, precision=precision, fp64_tail=tail, lp_error_check=True)(s).run()
    stages = [h[1] for h in mix.history]
    print(f"[scf-gpu] (H2O)_{n} {precision} tail {tail}: E {mix.E:.12f} fp64 {e64.E:.12f} "
          f"diff {mix.E - e64.E:.2e} stages {''.join(x[0] for x in stages)} "
          f"lp-vs-fp64 {mix.lp_error_vs_fp64}", flush=True)
    assert mix.converged and abs(mix.E - e64.E) <= 1e-8
    assert stages[0] == "lp" and "ref" in stages
    # the energy reported is that of the FP64 build after the last cycle
    assert stages[-1] in (("fp64",) if tail == "full" else ("inc", "ref", "fp64"))
    assert mix.lp_error_vs_fp64 is not None


def test_generalized_fock_matches_oracle(dev):
    s = System(WATER, "cc-pvdz", "cc-pvtz-jkfit", device=dev)
    fb = s.fock_builder
    B = fb.B_Pmn.cpu().numpy()
    rng = np.random.default_rng(0)
    C = rng.standard_normal((s.nbf, 4)) * 0.3
    A = rng.standard_normal((4, 4))
    g1 = A @ A.T
    J, K = fb.build_JK_generalized(C, g1)
    Jr, Kr = ojk.build_JK_generalized(B, C, g1)
    assert np.abs(J - Jr).max() < 1e-10 and np.abs(K - Kr).max() < 1e-10
    H = s.ints_hcore()
    assert np.abs(fb.build_core_fock(C[:, :2], H) - ojk.build_core_fock(B, C[:, :2], H)).max() < 1e-10
    assert np.abs(fb.build_active_fock(C, g1) - ojk.build_active_fock(B, C, g1)).max() < 1e-10


def test_metric_inverse_sqrt_branch_gives_same_jk(dev):
    # df_ortho_rtol selects X = M^-1/2 (jkbuilder.py:137-139); J and K do not
    # depend on the choice of metric factor when nothing is discarded
    from forte2_amd.df import build_B_density_fitting
    from forte2_amd.jkbuilder import FockBuilder

    s = System(WATER, "cc-pvdz", "cc-pvtz-jkfit", device=dev)
    Bc = build_B_density_fitting(s, dev)
    s.df_ortho_rtol = 1e-14
    Bs = build_B_density_fitting(s, dev)
    C = torch.linalg.qr(torch.randn(s.nbf, 5, dtype=torch.float64, generator=torch.Generator().manual_seed(1)))[0].to(dev)
    Jc, Kc = FockBuilder(B=Bc, device=dev).build_JK([C])
    Js, Ks = FockBuilder(B=Bs, device=dev).build_JK([C])
    assert float((Jc[0] - Js[0]).abs().max()) < 1e-10
    assert float((Kc[0] - Ks[0]).abs().max()) < 1e-10


def test_uhf_one_electron_golden(dev):
    # tests/scf/test_uhf.py:44-55: H2+ with ms = -1/2, so the alpha density is
    # empty (a zero-column C in every build)
    s = System("H 0 0 0\nH 0 0 1.0", "cc-pvqz", "cc-pvqz-jkfit", device=dev)
    scf = UHF(charge=1, ms=-0.5, e_tol=1e-10)(s).run()
    print(f"   E = {scf.E:.12f} <S^2> = {scf.S2:.6f}", flush=True)
    assert scf.E == pytest.approx(EREF["h2plus_qz_uhf"]["E"], abs=5e-8)
    assert scf.S2 == pytest.approx(0.75, abs=5e-8)
