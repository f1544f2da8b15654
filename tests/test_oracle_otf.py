"""The integral-direct oracle (oracle/otf.py) against the in-core oracle, and
the host logic of forte2_amd/otf.py (block sizes, aux-shell blocks, column
sweeps, the System factory choice).  CPU only."""

import math

import numpy as np
import pytest

from forte2_amd import otf
from forte2_amd.system import System
from oracle import jk as ojk
from oracle import otf as ootf

N2 = "N 0.0 0.0 0.0\nN 0.0 0.0 2.0"


@pytest.fixture(scope="module")
def n2():
    s = System(N2, "cc-pvdz", "cc-pvtz-jkfit", unit="bohr")
    return s, s.coulomb_2c(), s.coulomb_3c()


def _blocks(s, pblk):
    return [(p0, p1) for _, _, p0, p1 in otf.aux_shell_blocks(s.auxiliary_basis.shell_offsets(), pblk)]


@pytest.mark.parametrize("pblk,iblk", [(10_000, 100), (25, 3), (9, 1)])
@pytest.mark.parametrize("rtol", [None, 1e-10])
def test_otf_oracle_equals_incore(n2, pblk, iblk, rtol):
    s, M, T = n2
    rng = np.random.default_rng(3)
    C = np.linalg.qr(rng.standard_normal((s.nbf, 7)))[0]
    met = ootf.Metric(M, rtol)
    B = np.einsum("PQ,Qmn->Pmn", met.Mm12, T) if rtol is not None else ojk.build_B_density_fitting(M, T)
    Jr, Kr = ojk.build_JK(B, [C])
    blocks = _blocks(s, pblk)
    J = ootf.J_kernel(met, T, blocks, C @ C.T)
    K = ootf.K_kernel(met, T, blocks, C, iblk)
    J2, K2 = ootf.JK_kernel(met, T, blocks, C, iblk)
    for a, b in ((J, Jr[0]), (K, Kr[0]), (J2, Jr[0]), (K2, Kr[0])):
        assert np.abs(a - b).max() < 1e-10
    C1, C2 = rng.standard_normal((s.nbf, 2)), rng.standard_normal((s.nbf, 3))
    Bmo = ootf.B_tensor_gen_block(met, T, blocks, C1, C2)
    assert np.abs(Bmo - ojk.B_tensor_gen_block(B, C1, C2)).max() < 1e-10


def test_otf_oracle_metric_factor_matters(n2):
    # dropping M^-1 from the J sweep changes J: the checks above can fail
    s, M, T = n2
    met = ootf.Metric(M)
    C = np.linalg.qr(np.random.default_rng(4).standard_normal((s.nbf, 5)))[0]
    D = C @ C.T
    blocks = _blocks(s, 10_000)
    good = ootf.J_kernel(met, T, blocks, D)
    met.apply_Mm1 = lambda y: y
    assert np.abs(ootf.J_kernel(met, T, blocks, D) - good).max() > 1e-3


def test_block_sizes_rule():
    # jkbuilder.py:700-716: nbf 110, naux 212, 7 occupied, 4.5 MB
    iblk, pblk = otf.block_sizes(4.5, 110, 212, 7, 9)
    assert iblk == min(7, math.ceil(4.5 * 2**20 / (24 * 110 * 212))) == 7
    assert pblk == 21
    # a budget that holds everything
    assert otf.block_sizes(4000, 28, 158, 7, 9) == (7, 158)
    with pytest.raises(ValueError, match="too low"):
        otf.block_sizes(0.2, 110, 212, 7, 9)


def test_aux_shell_blocks_cover_shells():
    offs = [0, 1, 4, 9, 10, 17, 26, 27]
    for pblk in (9, 10, 20, 100):
        bl = otf.aux_shell_blocks(offs, pblk)
        assert bl[0][2] == 0 and bl[-1][3] == offs[-1]
        for (s0, s1, p0, p1), nxt in zip(bl, bl[1:] + [None]):
            assert p0 == offs[s0] and p1 == offs[s1] and 0 < p1 - p0 <= pblk
            if nxt is not None:
                assert nxt[0] == s1
                # greedy: the next shell would not have fitted
                assert offs[s1 + 1] - p0 > pblk


def test_column_sweeps_partition():
    for noccs, ib in (([5], 2), ([5, 4], 3), ([7, 0, 3], 7), ([1, 1, 1], 2), ([50, 49], 7)):
        sw = otf.column_sweeps(noccs, ib)
        seen = {k: [] for k in range(len(noccs))}
        for s in sw:
            assert sum(i1 - i0 for _, i0, i1 in s) <= ib
            for k, i0, i1 in s:
                assert 0 <= i0 < i1 <= noccs[k]
                seen[k].extend(range(i0, i1))
        for k, o in enumerate(noccs):
            assert sorted(seen[k]) == list(range(o))


def test_factory_choice(monkeypatch):
    made = []

    class Stub:
        def __init__(self, system, jk_mem_thres_mb, backend, device):
            made.append((jk_mem_thres_mb, backend))

    class InCore:
        def __init__(self, system, store_B_nPm=True, device=None):
            self.store_B_nPm = store_B_nPm

    from forte2_amd import jkbuilder

    monkeypatch.setattr(otf, "FockBuilderOTF", Stub)
    monkeypatch.setattr(jkbuilder, "FockBuilder", InCore)
    s = System(N2, "cc-pvdz", "cc-pvtz-jkfit", unit="bohr", jk_mem_thres_mb=0.5)
    assert isinstance(s.fock_builder, Stub) and made == [(0.5, "auto")]
    # B is 0.945 MB and the metric 0.19 MB (system.py:412-434)
    s = System(N2, "cc-pvdz", "cc-pvtz-jkfit", unit="bohr", jk_mem_thres_mb=1.5)
    assert isinstance(s.fock_builder, InCore) and not s.fock_builder.store_B_nPm
    s = System(N2, "cc-pvdz", "cc-pvtz-jkfit", unit="bohr", jk_mem_thres_mb=4000)
    assert isinstance(s.fock_builder, InCore) and s.fock_builder.store_B_nPm
    with pytest.raises(ValueError, match="auxiliary_basis_set"):
        System(N2, "cc-pvdz", None, unit="bohr", jk_mem_thres_mb=1).fock_builder


def test_invalid_backend_rejected(n2):
    with pytest.raises(ValueError, match="Invalid backend"):
        otf.FockBuilderOTF(n2[0], backend="cuda")
