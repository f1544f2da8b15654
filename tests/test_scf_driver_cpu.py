"""The device SCF driver (forte2_amd.scf) run on CPU tensors with a numpy-backed
builder: checks the loop, the precision schedule and the build-request
protocol without a GPU.  The builder's "low precision" rounds B to FP32."""

import json
import os
import re

import numpy as np
import pytest
import torch

from forte2_amd.scf import RHF, UHF
from forte2_amd.system import System
from oracle import jk as ojk
from oracle import scf as oscf

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_values.json")))
EREF = GOLD["energies"]

WATER = """O 0.000000000000 0.000000000000 -0.061664597388
H 0.000000000000 -0.711620616369 0.489330954643
H 0.000000000000 0.711620616369 0.489330954643"""


def _hilo(x):
    """x rounded as the BF16x3 kernels see it: FP32, then split into a bf16
    high part and a bf16 remainder (about 16 significant bits)."""
    x32 = torch.as_tensor(x).to(torch.float32)
    hi = x32.to(torch.bfloat16).to(torch.float32)
    lo = (x32 - hi).to(torch.bfloat16).to(torch.float32)
    return (hi + lo).to(torch.float64)


class HostBuilder:
    """build_JK / build_J / build_K over the oracle, on CPU torch tensors.
    "fp32" rounds B to FP32; "bf16x3" rounds B and the orbital factors to
    hi + lo bf16 pairs, the operand rounding of the BF16x3 products."""

    def __init__(self, B):
        self.B = B
        self.B32 = B.astype(np.float32).astype(np.float64)
        self.B16 = _hilo(B).numpy()
        self.device = torch.device("cpu")
        self.calls = []

    def _B(self, precision):
        return {"fp64": self.B, "fp32": self.B32, "bf16x3": self.B16}[precision]

    def _C(self, C, precision):
        return [(_hilo(c) if precision == "bf16x3" else c).numpy() for c in C]

    def build_JK(self, C, precision="fp64"):
        self.calls.append(("JK", precision))
        J, K = ojk.build_JK(self._B(precision), self._C(C, precision))
        return [torch.from_numpy(j) for j in J], [torch.from_numpy(k) for k in K]

    def build_J(self, D, precision="fp64"):
        self.calls.append(("J", precision))
        return [torch.from_numpy(j) for j in ojk.build_J(self._B(precision), [d.numpy() for d in D])]

    def build_K(self, C, precision="fp64"):
        self.calls.append(("K", precision))
        return [torch.from_numpy(k) for k in ojk.build_K_nPm(self._B(precision), self._C(C, precision))]


@pytest.fixture(scope="module")
def water():
    s = System(WATER, "cc-pvdz", "cc-pvtz-jkfit")
    B = ojk.build_B_density_fitting(s.coulomb_2c(), s.coulomb_3c())
    return s, B


def test_plain_form_matches_oracle(water):
    s, B = water
    s.fock_builder = HostBuilder(B)
    r = RHF(charge=0, e_tol=1e-10)(s).run()
    ref = oscf.run_rhf(s.ints_overlap(), s.ints_hcore(), s.nuclear_repulsion, oscf.DFJK(B), s.Zsum,
                       X=s.Xorth, V_sap=s.sap_potential(), e_tol=1e-10)
    assert abs(r.E - ref.E) < 1e-10
    assert r.iter == ref.iter
    # the final iteration diagonalises the new Fock matrix, which moves the
    # density once more (scf_base.py:214-221); at a loose threshold the move
    # shows in the energy (a repeat of the last diagonalisation would leave
    # it bit for bit)
    loose = RHF(charge=0, e_tol=1e-5, d_tol=1e-2, do_diis=False)(s).run()
    assert 1e-12 < abs(loose.E - loose.history[-1][2]) < 1e-4


@pytest.mark.parametrize("method", ["rhf", "uhf"])
def test_level_shift_matches_oracle(water, method):
    # rhf.py:90-95, uhf.py:150-159, switched off once |dE| < 1e-5
    # (scf_base.py:203-204): same energy and iteration count as the oracle
    s, B = water
    s.fock_builder = HostBuilder(B)
    common = dict(X=s.Xorth, V_sap=s.sap_potential(), e_tol=1e-10, do_diis=False)
    if method == "rhf":
        r = RHF(charge=0, e_tol=1e-10, do_diis=False, level_shift=0.3)(s).run()
        ref = oscf.run_rhf(s.ints_overlap(), s.ints_hcore(), s.nuclear_repulsion, oscf.DFJK(B), s.Zsum,
                           level_shift=0.3, **common)
    else:
        r = UHF(charge=1, ms=0.5, e_tol=1e-10, do_diis=False, level_shift=(0.2, 0.4))(s).run()
        ref = oscf.run_uhf(s.ints_overlap(), s.ints_hcore(), s.nuclear_repulsion, oscf.DFJK(B), s.Zsum - 1,
                           0.5, level_shift=(0.2, 0.4), **common)
    assert abs(r.E - ref.E) < 1e-10
    assert r.iter == ref.iter


def test_guess_mix_matches_oracle():
    # the device driver's mixed ms = 0 guess (scf_utils.py:76-104) on the
    # reference's stretched H2 (tests/scf/test_uhf.py:73-90)
    s = System("H 0 0 0\nH 0 0 2.7", "cc-pvtz", "cc-pvqz-jkfit")
    B = ojk.build_B_density_fitting(s.coulomb_2c(), s.coulomb_3c())
    s.fock_builder = HostBuilder(B)
    u = UHF(charge=0, ms=0.0, guess_mix=True, e_tol=1e-10)(s).run()
    ref = oscf.run_uhf(s.ints_overlap(), s.ints_hcore(), s.nuclear_repulsion, oscf.DFJK(B), s.Zsum, 0.0,
                       X=s.Xorth, V_sap=s.sap_potential(), e_tol=1e-10, guess_mix=True)
    assert abs(u.E - ref.E) < 1e-10 and u.iter == ref.iter
    assert u.E == pytest.approx(EREF["h2_stretched_tz_uhf_guess_mix"]["E"], abs=5e-8)
    assert u.S2 == pytest.approx(EREF["h2_stretched_tz_uhf_guess_mix"]["S2"], abs=5e-8)


def test_level_shift_validation(water):
    s, _ = water
    with pytest.raises(ValueError, match="non-negative"):
        RHF(charge=0, level_shift=-0.1)(s)
    with pytest.raises(ValueError, match="only valid for UHF"):
        RHF(charge=0, level_shift=(0.1, 0.2))(s)
    with pytest.raises(ValueError, match="length 2"):
        UHF(charge=0, ms=0.0, level_shift=(0.1, 0.2, 0.3))(s)
    u = UHF(charge=0, ms=0.0, level_shift=0.25)(s)
    assert u.level_shift == (0.25, 0.25)


def test_precision_schedule_and_requests(water):
    s, B = water
    hb = HostBuilder(B)
    s.fock_builder = hb
    r = RHF(charge=0, e_tol=1e-10, precision="fp32", lp_error_check=True)(s).run()
    e64 = RHF(charge=0, e_tol=1e-10)(s).run().E
    stages = "".join(h[1][0] for h in r.history)
    # low precision, the first FP64 reference, incremental cycles, one FP64
    # re-reference, incremental cycles to convergence (then the FP64 final build)
    assert stages.startswith("l") and stages.count("r") == 2 and stages.endswith("i")
    assert stages.index("r") < stages.index("i") < stages.rindex("r")
    assert abs(r.E - e64) < 1e-9
    assert r.lp_error_vs_fp64 is not None and r.lp_error_vs_fp64 > 0
    kinds = {c for c in hb.calls}
    assert ("JK", "fp32") in kinds and ("J", "fp32") in kinds and ("K", "fp32") in kinds
    assert ("JK", "fp64") in kinds


@pytest.mark.parametrize("tail", ["reref", "full"])
def test_bf16x3_like_tails_end_at_the_fp64_energy(water, tail):
    """Both tails of the precision form, on a builder with the BF16x3 operand
    rounding: the final energy is within 1e-9 Eh of the FP64 SCF's, and the
    re-reference tail makes fewer FP64 builds than the FP64-cycle tail."""
    s, B = water
    hb = HostBuilder(B)
    s.fock_builder = hb
    e64 = RHF(charge=0, e_tol=1e-10)(s).run().E
    hb.calls.clear()
    r = RHF(charge=0, e_tol=1e-10, precision="bf16x3", fp64_tail=tail, lp_error_check=True)(s).run()
    stages = "".join(h[1][0] for h in r.history)
    n64 = sum(1 for c in hb.calls if c[1] == "fp64")
    print(f"tail {tail}: stages {stages} FP64 builds {n64} |dE| {abs(r.E - e64):.1e}")
    assert r.converged and abs(r.E - e64) < 1e-9
    assert r.lp_error_vs_fp64 > 1e-7  # the emulated rounding is BF16x3-sized
    if tail == "reref":
        assert stages.count("r") == 2 and "f" not in stages and n64 == 3
    else:
        assert stages.count("r") == 1 and stages.endswith("f") and n64 == 1 + stages.count("f") + 1
    # without the check no extra low-precision build is made
    hb.calls.clear()
    r2 = RHF(charge=0, e_tol=1e-10, precision="bf16x3", fp64_tail=tail)(s).run()
    assert r2.lp_error_vs_fp64 is None and r2.E == r.E
    # the first build, one JK per lp/ref/fp64 cycle, J + K per incremental
    # cycle, the final FP64 build
    assert len(hb.calls) == 1 + sum(2 if h[1] == "inc" else 1 for h in r2.history[1:]) + 1


def test_tail_takes_another_reference_before_fp64_cycles(water):
    """With tail_refs=2 a stalled tail takes one more FP64 reference, then
    FP64 cycles."""
    s, B = water
    s.fock_builder = HostBuilder(B)
    e64 = RHF(charge=0, e_tol=1e-10)(s).run().E
    r = RHF(charge=0, e_tol=1e-10, precision="bf16x3", inc_max_tail=1, tail_refs=2)(s).run()
    stages = "".join(h[1][0] for h in r.history)
    # low precision, the first reference, incremental cycles to |dE| < 1e-6,
    # then the tail: reference, one cycle, reference, one cycle, FP64 cycles
    assert re.fullmatch(r"l+ri+ririf+", stages), stages
    assert r.converged and abs(r.E - e64) < 1e-9


def test_tail_falls_back_to_fp64_cycles(water):
    """An incremental tail that has not converged within inc_max_tail cycles
    hands over to FP64 cycles."""
    s, B = water
    s.fock_builder = HostBuilder(B)
    e64 = RHF(charge=0, e_tol=1e-10)(s).run().E
    r = RHF(charge=0, e_tol=1e-10, precision="bf16x3", inc_max_tail=0)(s).run()
    stages = "".join(h[1][0] for h in r.history)
    assert stages.count("r") == 2 and stages.endswith("f")
    assert stages[stages.rindex("r") + 1] == "f"  # no incremental cycle after the re-reference
    assert r.converged and abs(r.E - e64) < 1e-9
    with pytest.raises(ValueError, match="fp64_tail"):
        RHF(charge=0, fp64_tail="fp32")


def test_generator_protocol_yields_build_requests(water):
    s, B = water
    hb = HostBuilder(B)
    s.fock_builder = hb
    scf = RHF(charge=0, e_tol=1e-8)(s)
    gen = scf.iterate()
    req = next(gen)
    assert req[0] == "JK" and req[2] == "fp64" and req[1][0].shape == (s.nbf, 5)
    n = 0
    try:
        while True:
            req = gen.send(scf._execute(req))
            n += 1
    except StopIteration:
        pass
    assert scf.converged and n == scf.iter + 1


def test_uhf_cation_on_host_builder():
    s = System("H 0 0 0\nH 0 0 0.74\nH 0 0.9 0.37", "cc-pvdz", "cc-pvtz-jkfit")
    B = ojk.build_B_density_fitting(s.coulomb_2c(), s.coulomb_3c())
    s.fock_builder = HostBuilder(B)
    r = UHF(charge=0, ms=0.5, e_tol=1e-10)(s).run()
    ref = oscf.run_uhf(s.ints_overlap(), s.ints_hcore(), s.nuclear_repulsion, oscf.DFJK(B), s.Zsum, 0.5,
                       X=s.Xorth, V_sap=s.sap_potential(), e_tol=1e-10)
    assert abs(r.E - ref.E) < 1e-10
    assert abs(r.S2 - ref.S2) < 1e-8
