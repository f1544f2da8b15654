"""Host logic of the device SCF, exercised on CPU tensors (no GPU needed)."""

import numpy as np
import pytest
import torch

from forte2_amd.scf import DeviceDIIS, _procrustes_split, canonical_orth
from oracle import scf as oscf


def _orthonormal(rng, N, o, S):
    A = rng.standard_normal((N, o))
    # S-orthonormalise: C = A (A^T S A)^-1/2
    w, U = np.linalg.eigh(A.T @ S @ A)
    return A @ U @ np.diag(w ** -0.5) @ U.T


def test_difference_density_factorisation_is_exact():
    rng = np.random.default_rng(4)
    N, o = 12, 4
    R = rng.standard_normal((N, N))
    S = R @ R.T / N + np.eye(N)
    Co = _orthonormal(rng, N, o, S)
    # new orbitals: a small rotation of the old occupied space, S-orthonormalised
    Cn = Co + 0.05 * rng.standard_normal((N, o))
    w, U = np.linalg.eigh(Cn.T @ S @ Cn)
    Cn = Cn @ U @ np.diag(w ** -0.5) @ U.T
    t = torch.from_numpy
    Up, Um = _procrustes_split(t(Cn), t(Co), t(S))
    lhs = 0.5 * (Up @ Up.T - Um @ Um.T)
    assert float((lhs - t(Cn @ Cn.T - Co @ Co.T)).abs().max()) < 1e-13
    # balanced factors: both carry a norm ~ sqrt(|P| |A|)
    assert 0.2 < float(torch.linalg.norm(Up) / torch.linalg.norm(Um)) < 5
    # a rotation within the occupied space is no change of density
    Q, _ = np.linalg.qr(rng.standard_normal((o, o)))
    split = _procrustes_split(t(Co @ Q), t(Co), t(S))
    if split is not None:
        assert float((0.5 * (split[0] @ split[0].T - split[1] @ split[1].T)).abs().max()) < 1e-12


def test_device_diis_matches_oracle_diis():
    rng = np.random.default_rng(1)
    d1, d2 = DeviceDIIS(1, 8, 2), oscf.DIIS(1, 8, 2)
    for k in range(12):
        p = rng.standard_normal((5, 5))
        e = rng.standard_normal((5, 5)) * 0.5 ** k
        a = d1.update(torch.from_numpy(p), torch.from_numpy(e)).numpy()
        b = d2.update(p, e)
        assert np.abs(a - b).max() < 1e-12
        assert d1.status == d2.status


def test_device_diis_skips_singular_system():
    d = DeviceDIIS(1, 8, 2)
    p = torch.eye(3, dtype=torch.float64)
    e = torch.zeros((3, 3), dtype=torch.float64)
    d.update(p, e)
    out = d.update(2 * p, e)  # all-zero errors: singular B matrix
    assert torch.equal(out, 2 * p)


def test_canonical_orth_matches_oracle():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((10, 6))
    S = A @ A.T + 1e-10 * np.eye(10)  # rank-deficient metric: 4 vectors discarded
    Xd = canonical_orth(torch.from_numpy(S)).numpy()
    Xo = oscf.canonical_orth(S.copy())
    assert Xd.shape == Xo.shape == (10, 6)
    assert np.abs(Xd.T @ S @ Xd - np.eye(6)).max() < 1e-10
    assert np.abs(Xd @ Xd.T - Xo @ Xo.T).max() < 1e-8


def _affine_min(E):
    """argmin ||E c|| subject to sum(c) = 1, by brute force: c = c0 + Z z with
    c0 = 1/n and Z an orthonormal basis of {v : sum(v) = 0}, then a plain
    least-squares solve for z (no Lagrange system, no normalisation)."""
    n = E.shape[1]
    c0 = np.full(n, 1.0 / n)
    Z = np.linalg.svd(np.ones((1, n)))[2][1:].T
    z = np.linalg.lstsq(E @ Z, -E @ c0, rcond=None)[0]
    return c0 + Z @ z


@pytest.mark.parametrize("impl", ["oracle", "device"])
def test_diis_coefficients_minimise_the_error_on_the_affine_hull(impl):
    # diis.py:102-125: the CDIIS coefficients sum to one and minimise the norm
    # of the combined error vector over the stored window (nvec = 8).  The
    # stored "Fock matrices" are one-hot vectors, so the extrapolated output is
    # the coefficient vector itself.
    rng = np.random.default_rng(12)
    nit, nvec = 13, 8
    d = oscf.DIIS(1, nvec, 2) if impl == "oracle" else DeviceDIIS(1, nvec, 2)
    errs = []
    for k in range(nit):
        p = np.zeros(nit)
        p[k] = 1.0
        e = rng.standard_normal((4, 5)) * (0.7 ** k)
        errs.append(e.ravel())
        if impl == "oracle":
            c = d.update(p, e)
        else:
            c = d.update(torch.from_numpy(p), torch.from_numpy(e)).numpy()
        if k == 0:
            assert np.array_equal(c, p)  # fewer than diis_min vectors: no extrapolation
            continue
        lo = max(0, k + 1 - nvec)
        assert np.all(c[:lo] == 0.0) and np.all(c[k + 1:] == 0.0)  # only the window contributes
        w = c[lo:k + 1]
        assert abs(w.sum() - 1.0) < 1e-12
        ref = _affine_min(np.array(errs[lo:k + 1]).T)
        assert np.abs(w - ref).max() < 1e-8 * max(1.0, np.abs(ref).max())


def test_polar_factor_matches_the_svd_and_falls_back():
    """The Newton-Schulz polar factor of the difference-density split equals
    U Vh of the SVD (the orthogonal Procrustes solution) for a near-orthogonal
    M, is orthogonal to roundoff, and a singular M takes the SVD path."""
    from forte2_amd.scf import _polar, _procrustes_split

    g = torch.Generator().manual_seed(5)
    n = 40
    Q, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, generator=g))
    for eps in (1e-6, 1e-3, 5e-2):
        M = Q + eps * torch.randn(n, n, dtype=torch.float64, generator=g)
        U, _, Vh = torch.linalg.svd(M)
        R = _polar(M)
        assert torch.allclose(R, U @ Vh, atol=1e-12, rtol=0)
        assert float((R.T @ R - torch.eye(n, dtype=torch.float64)).abs().max()) < 5e-14
    Msing = Q.clone()
    Msing[:, 0] = 0.0  # a 90-degree rotation: singular, Newton-Schulz cannot converge
    R = _polar(Msing)
    assert float((R.T @ R - torch.eye(n, dtype=torch.float64)).abs().max()) < 1e-12
    # the split's identity with the polar factor: Cn Cn^T - Co Co^T = (U+U+^T - U-U-^T) / 2
    N, o = 60, 12
    S = torch.eye(N, dtype=torch.float64)
    Co, _ = torch.linalg.qr(torch.randn(N, o, dtype=torch.float64, generator=g))
    Cn, _ = torch.linalg.qr(Co + 1e-2 * torch.randn(N, o, dtype=torch.float64, generator=g))
    Up, Um = _procrustes_split(Cn, Co, S)
    lhs = Cn @ Cn.T - Co @ Co.T
    assert torch.allclose(lhs, 0.5 * (Up @ Up.T - Um @ Um.T), atol=1e-14, rtol=0)
