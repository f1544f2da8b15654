"""MO-basis integrals on the device builder against the oracle
(tests/jkbuilder/test_mointegrals.py:8-38): V against the einsum over B with
random C, its antisymmetrized form, and the core-orbital energy and one-body
terms against the oracle's J/K."""

import numpy as np
import pytest

from forte2_amd.mointegrals import RestrictedMOIntegrals
from forte2_amd.system import System
from oracle import jk as ojk

pytestmark = pytest.mark.gpu


def test_restricted_mo_integrals(dev):
    s = System("H 0.0 0.0 0.0\nF 0.0 0.0 2.0", "cc-pvdz", "cc-pvtz-jkfit", unit="bohr", device=dev)
    B = s.fock_builder.B_Pmn.cpu().numpy()
    C = np.random.default_rng(5).random((s.nbf, s.nbf))
    ints = RestrictedMOIntegrals(s, C, orbitals=range(C.shape[1]))
    B_mo = np.einsum("Bpq,pi,qj->Bij", B, C, C, optimize=True)
    V_ref = np.einsum("Bij,Bkl->ikjl", B_mo, B_mo, optimize=True)
    scale = np.abs(V_ref).max()
    print(f"   |dV|/max = {np.abs(ints.V - V_ref).max() / scale:.2e}", flush=True)
    assert np.abs(ints.V - V_ref).max() <= 1e-12 * scale
    anti = RestrictedMOIntegrals(s, C, orbitals=list(range(C.shape[1])), antisymmetrize=True)
    assert np.abs(anti.V - (V_ref - V_ref.swapaxes(2, 3))).max() <= 1e-12 * scale


def test_core_orbitals_fold_into_E_and_H(dev):
    s = System("H 0.0 0.0 0.0\nF 0.0 0.0 2.0", "cc-pvdz", "cc-pvtz-jkfit", unit="bohr", device=dev)
    B = s.fock_builder.B_Pmn.cpu().numpy()
    C = np.linalg.qr(np.random.default_rng(6).standard_normal((s.nbf, s.nbf)))[0]
    core, act = [0, 1], [2, 3, 4, 5]
    ints = RestrictedMOIntegrals(s, C, orbitals=act, core_orbitals=core)
    h = s.ints_hcore()
    Cc, Ca = C[:, core], C[:, act]
    J, K = ojk.build_JK(B, [Cc])
    G = 2 * J[0] - K[0]
    E_ref = s.nuclear_repulsion + 2 * np.einsum("mi,mn,ni->", Cc, h, Cc) + np.einsum("mi,mn,ni->", Cc, G, Cc)
    H_ref = Ca.T @ (h + G) @ Ca
    assert abs(ints.E - E_ref) <= 1e-10
    assert np.abs(ints.H - H_ref).max() <= 1e-10
    assert ints.V.shape == (4, 4, 4, 4)
