"""B formation algebra of forte2_amd/df.py on torch CPU tensors against the
oracle's B = L^-1 (Q|mn) (jkbuilder.py:135-159): the pair-block pipeline
and the packed host/device split of the serving layer, both metric branches.
The GPU runs of the same code are in tests/test_scf_gpu.py."""

import numpy as np
import pytest

from forte2_amd import df
from forte2_amd import geometry as geo
from forte2_amd.system import System
from oracle import jk as ojk
from oracle import otf as ootf


@pytest.fixture(scope="module")
def w4():
    s = System(geo.water_cluster(4, seed=4), "cc-pvdz", "cc-pvtz-jkfit")
    return s, s.coulomb_2c(), s.coulomb_3c()


def test_pair_block_pipeline_matches_oracle(w4):
    s, M, T = w4
    st = {}
    B = df.build_B_density_fitting(s, "cpu", block_bytes=1 << 20, stats=st).numpy()
    assert st["blocks"] > 10
    assert np.abs(B - ojk.build_B_density_fitting(M, T)).max() < 1e-11
    assert np.array_equal(B, B.transpose(0, 2, 1))


def test_packed_host_device_split_matches_oracle(w4):
    s, M, T = w4
    Mh, packed = df.host_df_integrals(s)
    assert np.array_equal(Mh, M)
    B = df.device_B_from_host(Mh, packed, "cpu", s.nbf, block_bytes=1 << 20).numpy()
    assert np.abs(B - ojk.build_B_density_fitting(M, T)).max() < 1e-11
    # the truncated-eigen metric branch (jkbuilder.py:137-139)
    Be = df.device_B_from_host(Mh, packed, "cpu", s.nbf, df_ortho_rtol=1e-10).numpy()
    Xe = ootf.Metric(M, 1e-10).Mm12
    assert np.abs(Be - np.einsum("PQ,Qmn->Pmn", Xe, T)).max() < 1e-10


def test_linearly_dependent_metric_raises():
    # tests/jkbuilder/test_jkbuilder.py:392-415: two identical fitting
    # functions make (P|Q) exactly singular; the B formation and the
    # integral-direct metric refuse it (jkbuilder.py:141-148), whatever the
    # rounding order of the library's Cholesky leaves as the last pivot
    import torch

    from forte2_amd import integrals as ints
    from forte2_amd.basis import BasisSet

    fake = BasisSet("two identical s functions")
    fake.add(0, [1.0], [1.0], [0.0, 0.0, 0.0])
    fake.add(0, [1.0], [1.0], [0.0, 0.0, 0.0])
    M = torch.from_numpy(ints.coulomb_2c(fake))
    with pytest.raises(ValueError, match="positive definite"):
        df.metric_factor(M)
    with pytest.raises(ValueError, match="positive definite"):
        ojk.build_B_density_fitting(M.numpy(), np.zeros((2, 1, 1)))
    # a well-conditioned metric still factors
    kind, L = df.metric_factor(M + torch.eye(2, dtype=M.dtype))
    assert kind == "chol"
