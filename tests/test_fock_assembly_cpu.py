"""Generalized-Fock assembly on a builder whose J/K methods are replaced, as
forte2's own test does (tests/jkbuilder/test_jkbuilder.py:9-53): the
assembly depends only on the duck-typed build_JK / build_JK_generalized
surface, for the in-core and the integral-direct builder alike.  CPU only
(the builders are made with __new__, no device is touched)."""

import numpy as np
import pytest

from forte2_amd.jkbuilder import FockBuilder
from forte2_amd.otf import FockBuilderOTF


@pytest.mark.parametrize("builder_cls", [FockBuilder, FockBuilderOTF])
@pytest.mark.parametrize("two_component", [False, True])
def test_generalized_fock_assembly(builder_cls, two_component):
    hcore = np.array([[1.0, 0.1], [0.1, 0.8]])
    Jcore = np.array([[0.4, 0.2], [0.2, 0.3]])
    Kcore = np.array([[0.1, 0.05], [0.05, 0.2]])
    Jact = np.array([[0.3, 0.04], [0.04, 0.2]])
    Kact = np.array([[0.08, 0.03], [0.03, 0.06]])

    class DummySystem:
        def __init__(self):
            self.two_component = two_component

        def ints_hcore(self):
            return hcore

    fb = builder_cls.__new__(builder_cls)
    fb.system = DummySystem()
    fb.build_JK = lambda C: ([Jcore], [Kcore])
    fb.build_JK_generalized = lambda C, g1: (Jact, Kact)
    if two_component:
        core_ref, active_ref = hcore + Jcore - Kcore, Jact - Kact
    else:
        core_ref, active_ref = hcore + 2.0 * Jcore - Kcore, Jact - 0.5 * Kact
    eye = np.eye(2)
    assert np.array_equal(fb.build_core_fock(eye), core_ref)
    assert np.array_equal(fb.build_active_fock(eye, eye), active_ref)
    assert np.array_equal(fb.build_generalized_fock(eye, eye, eye), core_ref + active_ref)
    assert np.allclose(fb.build_core_fock(eye, hcore=hcore + eye), core_ref + eye, rtol=0, atol=1e-15)
