"""Host integral engine (libf2ints.so) against the reference's known answers,
closed forms and the AO convention of forte2/integrals/real_spherical_harmonics.hpp.

Golden values: tests/integrals/test_two_electron.py:10-46 (25 STO-3G water
(mn|rs)), tests/integrals/test_libcint.py:219-241,276-287 (Frobenius norms of
(P|Q), (m|P), (P|mn)), all at the reference's tolerances (abs 5e-8, rel 1e-6).
"""

import json
import math
import os

import mpmath
import numpy as np
import pytest

from forte2_amd import basis as bs
from forte2_amd import geometry as geo
from forte2_amd import integrals as ints

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_values.json")))
NORMS = GOLD["norms"]

WATER_STO3G = """O 0.000000000000     0.000000000000    -0.061664597388
H 0.000000000000    -0.711620616370     0.489330954643
H 0.000000000000     0.711620616370     0.489330954643"""

STO3G_ERI = {tuple(k): v for k, v in GOLD["sto3g_water_eri"]["values"]}  # tests/integrals/test_two_electron.py:10-46


def test_sto3g_water_eri_golden():
    b = bs.build_basis("sto-3g", geo.parse_geometry(WATER_STO3G))
    V = ints.coulomb_4c(b)
    for k, v in STO3G_ERI.items():
        assert V[k] == pytest.approx(v, abs=5e-8), k
    # 8-fold permutational symmetry
    assert np.abs(V - V.transpose(1, 0, 2, 3)).max() < 1e-14
    assert np.abs(V - V.transpose(2, 3, 0, 1)).max() < 1e-14


def test_three_center_norm_golden():
    atoms = geo.parse_geometry("O\nH 1 1.1\nH 1 1.1 2 104.5")
    P = ints.coulomb_3c(bs.build_basis("cc-pvtz-jkfit", atoms), bs.build_basis("cc-pvdz", atoms))
    assert np.linalg.norm(P) == pytest.approx(NORMS["water_dz_tzjkfit_3c"]["value"], rel=1e-6)
    assert np.abs(P - P.transpose(0, 2, 1)).max() == 0.0


def test_two_center_norms_golden():
    atoms = geo.parse_geometry("N 0 0 0\nN 0 0 1.1")
    b, a = bs.build_basis("cc-pvdz", atoms), bs.build_basis("cc-pvtz-jkfit", atoms)
    M = ints.coulomb_2c(a)
    assert np.linalg.norm(M) == pytest.approx(NORMS["n2_2c_metric"]["value"], rel=1e-6)
    assert np.linalg.norm(ints.coulomb_2c(b, a)) == pytest.approx(NORMS["n2_2c_mixed"]["value"], rel=1e-6)
    assert np.linalg.eigvalsh(M).min() > 0  # Coulomb metric is positive definite


@pytest.mark.parametrize("T", [0.0, 1e-9, 0.013, 0.7, 4.4, 17.3, 29.99, 35.99, 36.01, 80.0, 500.0])
def test_boys_function_against_mpmath(T):
    mpmath.mp.dps = 40
    F = ints.boys(24, T)
    for n in range(25):
        if T == 0.0:
            ref = 1.0 / (2 * n + 1)
        else:
            a = mpmath.mpf(n) + mpmath.mpf(1) / 2
            ref = float(mpmath.gammainc(a, 0, T) / (2 * mpmath.mpf(T) ** a))
        assert abs(F[n] - ref) <= 2e-15 * ref, (n, T)


def test_s_type_closed_forms():
    # two normalised s primitives: S = (4ab/(a+b)^2)^{3/4} e^{-ab/(a+b) R^2};
    # (a|b) Coulomb of unit-normalised s Gaussians via F_0
    a_, b_, R = 0.8, 1.7, 1.3
    basis = bs.BasisSet()
    basis.add(0, [a_], [1.0], [0, 0, 0])
    basis.add(0, [b_], [1.0], [0, 0, R])
    S = ints.overlap(basis)
    assert S[0, 1] == pytest.approx((4 * a_ * b_ / (a_ + b_) ** 2) ** 0.75 * math.exp(-a_ * b_ / (a_ + b_) * R * R), rel=1e-14)
    M = ints.coulomb_2c(basis)
    Na, Nb = (2 * a_ / math.pi) ** 0.75, (2 * b_ / math.pi) ** 0.75
    al = a_ * b_ / (a_ + b_)
    F0 = 0.5 * math.sqrt(math.pi / (al * R * R)) * math.erf(math.sqrt(al * R * R))
    ref = Na * Nb * 2 * math.pi ** 2.5 / (a_ * b_ * math.sqrt(a_ + b_)) * F0
    assert M[0, 1] == pytest.approx(ref, rel=1e-14)
    # kinetic energy of a normalised s Gaussian: 3a/2
    assert ints.kinetic(basis)[0, 0] == pytest.approx(1.5 * a_, rel=1e-14)
    # nuclear attraction of a unit charge at the centre: -2 sqrt(2a/pi)
    V = ints.nuclear(basis, [(1, np.zeros(3))])
    assert V[0, 0] == pytest.approx(-2.0 * math.sqrt(2 * a_ / math.pi), rel=1e-14)


def _poly(T, l, x, y, z):
    cart = [(a, b, l - a - b) for a in range(l, -1, -1) for b in range(l - a, -1, -1)]
    mono = np.array([x ** a * y ** b * z ** c for a, b, c in cart])
    return T @ mono


def test_solid_harmonic_convention():
    # AO order and phases of forte2/integrals/real_spherical_harmonics.hpp:33-60
    rng = np.random.default_rng(0)
    s3 = math.sqrt(3.0)
    for x, y, z in rng.standard_normal((5, 3)):
        assert np.allclose(_poly(ints.cart2sph(1), 1, x, y, z), [y, z, x], atol=1e-14)
        assert np.allclose(_poly(ints.cart2sph(2), 2, x, y, z),
                           [s3 * x * y, s3 * y * z, z * z - 0.5 * (x * x + y * y), s3 * x * z,
                            0.5 * s3 * (x * x - y * y)], atol=1e-13)
    # all l: harmonic (Laplacian zero), mutually orthogonal on the unit sphere
    # with the norm of z^l, and the m = 0 component leads with z^l
    def sphere(a, b, c):  # integral of x^a y^b z^c over the unit sphere
        if a % 2 or b % 2 or c % 2:
            return 0.0
        return 2 * math.gamma((a + 1) / 2) * math.gamma((b + 1) / 2) * math.gamma((c + 1) / 2) / \
            math.gamma((a + b + c + 3) / 2)

    for l in range(1, 7):
        T = ints.cart2sph(l)
        cart = [(a, b, l - a - b) for a in range(l, -1, -1) for b in range(l - a, -1, -1)]
        Mc = np.array([[sphere(p[0] + q[0], p[1] + q[1], p[2] + q[2]) for q in cart] for p in cart])
        G = T @ Mc @ T.T
        zl = sphere(0, 0, 2 * l)
        assert np.abs(G - zl * np.eye(2 * l + 1)).max() < 1e-12 * zl
        assert T[l, cart.index((0, 0, l))] == pytest.approx(1.0)
        # exact Laplacian in the monomial basis of degree l-2
        if l >= 2:
            low = [(a, b, l - 2 - a - b) for a in range(l - 2, -1, -1) for b in range(l - 2 - a, -1, -1)]
            Lap = np.zeros((len(low), len(cart)))
            for j, (a, b, c) in enumerate(cart):
                for k, (da, db, dc) in enumerate(((2, 0, 0), (0, 2, 0), (0, 0, 2))):
                    e = (a, b, c)[k]
                    if e >= 2:
                        Lap[low.index((a - da, b - db, c - dc)), j] += e * (e - 1)
            assert np.abs(T @ Lap.T).max() < 1e-12


def test_contraction_normalisation():
    atoms = geo.parse_geometry(WATER_STO3G)
    for name in ("sto-3g", "cc-pvdz", "cc-pvqz"):
        S = ints.overlap(bs.build_basis(name, atoms))
        assert np.abs(np.diag(S) - 1.0).max() < 1e-13


def test_three_center_pair_blocks_match_square():
    # the packed unique half (f2ints_coulomb_3c_pairs) holds exactly the
    # (P|mn), m >= n, of the square routine, block by block
    from forte2_amd.df import pair_columns

    atoms = geo.parse_geometry("O\nH 1 1.1\nH 1 1.1 2 104.5")
    aux, b = bs.build_basis("cc-pvtz-jkfit", atoms), bs.build_basis("cc-pvdz", atoms)
    full = ints.coulomb_3c(aux, b)
    col0, m, n = pair_columns(b)
    N = b.size
    assert np.all(m >= n) and len(col0) - 1 == len(ints.ket_pairs(b)[0])
    # every unique (m, n) appears once
    assert len(set(zip(m.tolist(), n.tolist()))) == len(m) == N * (N + 1) // 2
    npairs = len(col0) - 1
    for q0, q1 in ((0, npairs), (0, 5), (5, 17), (17, npairs)):
        c0, c1 = col0[q0], col0[q1]
        out = np.full((aux.size, c1 - c0 + 3), np.nan)
        ints.coulomb_3c_pairs(aux, b, q0, q1, out)
        assert np.array_equal(out[:, : c1 - c0], full[:, m[c0:c1], n[c0:c1]])
        assert np.isnan(out[:, c1 - c0:]).all()  # nothing past the block's columns
    with pytest.raises(RuntimeError):
        ints.coulomb_3c_pairs(aux, b, 0, npairs, np.empty((aux.size, 4)))
    # the column-major layout (f2ints_coulomb_3c_pairs_t): the same values,
    # one row per packed column, over the same pair ranges
    for q0, q1 in ((0, npairs), (5, 17)):
        c0, c1 = col0[q0], col0[q1]
        out_t = np.full((c1 - c0 + 2, aux.size), np.nan)
        ints.coulomb_3c_pairs_t(aux, b, q0, q1, out_t)
        assert np.array_equal(out_t[: c1 - c0].T, full[:, m[c0:c1], n[c0:c1]])
        assert np.isnan(out_t[c1 - c0:]).all()
        assert ints.packed_columns(b, q0, q1) == c1 - c0
    with pytest.raises(ValueError):  # fewer rows than packed columns: refused before the engine runs
        ints.coulomb_3c_pairs_t(aux, b, 0, npairs, np.empty((int(col0[-1]) - 1, aux.size)))


@pytest.mark.parametrize("geom,orb,fit", [
    ("O\nH 1 1.1\nH 1 1.1 2 104.5\nO 1 3.0 2 100.0 3 180.0", "cc-pvdz", "cc-pvtz-jkfit"),
    ("Cu 0 0 0\nH 0 0 1.5", "cc-pvdz", "def2-universal-jkfit"),   # f orbital shells, aux shells up to l = 6
], ids=["water_dimer", "CuH"])
def test_three_center_pairs_by_aux_shells(geom, orb, fit):
    # f2ints_coulomb_3c_pairs_aux: the rows of the aux shells [s0, s1) of the
    # packed unique half, bit for bit; the scatter to (m, n) and (n, m)
    # rebuilds the square block (zero off the significant pairs in both)
    from forte2_amd.df import pair_columns

    atoms = geo.parse_geometry(geom)
    aux, b = bs.build_basis(fit, atoms), bs.build_basis(orb, atoms)
    col0, m, n = pair_columns(b)
    npairs, nc = len(col0) - 1, int(col0[-1])
    whole = np.empty((aux.size, nc))
    ints.coulomb_3c_pairs(aux, b, 0, npairs, whole)
    offs = aux.shell_offsets()
    nsh = len(offs) - 1
    for s0, s1 in ((0, 1), (0, 7), (7, 20), (20, nsh)):
        out = np.empty((offs[s1] - offs[s0], nc))
        ints.coulomb_3c_pairs_aux(aux, b, s0, s1, 0, npairs, out)
        assert np.array_equal(out, whole[offs[s0]:offs[s1]])
        # column-major, rows padded past the block (the padding untouched)
        out_t = np.full((nc, offs[s1] - offs[s0] + 3), np.nan)
        ints.coulomb_3c_pairs_aux_t(aux, b, s0, s1, 0, npairs, out_t)
        assert np.array_equal(out_t[:, : offs[s1] - offs[s0]], out.T)
        assert np.isnan(out_t[:, offs[s1] - offs[s0]:]).all()
        sq = ints.coulomb_3c(aux, b, s0, s1)
        rebuilt = np.zeros_like(sq)
        rebuilt[:, m, n] = out
        rebuilt[:, n, m] = out
        assert np.array_equal(rebuilt, sq)
    with pytest.raises(ValueError):
        ints.coulomb_3c_pairs_aux(aux, b, 3, 3, 0, npairs, np.empty((0, nc)))
    with pytest.raises(ValueError):
        ints.coulomb_3c_pairs_aux(aux, b, 0, 2, 0, npairs, np.empty((1, nc)))


def test_three_center_sum_matches_blocks():
    # f2ints_coulomb_3c_sum = sum over P of the (P|mn) blocks (the SAP guess potential)
    atoms = geo.parse_geometry("O\nH 1 1.1\nH 1 1.1 2 104.5")
    b = bs.build_basis("cc-pvdz", atoms)
    dens = bs.build_sap_density_basis(atoms)
    ref = ints.coulomb_3c(dens, b).sum(axis=0)
    V = ints.coulomb_3c_sum(dens, b)
    assert np.abs(V - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.array_equal(V, V.T)


def test_engine_handle_is_made_once_across_threads():
    # threads asking a fresh basis for its engine handle at once get the same
    # one (a second handle would free the first while it is in use)
    import threading

    atoms = geo.parse_geometry("O\nH 1 1.1\nH 1 1.1 2 104.5")
    for _ in range(5):
        b = bs.build_basis("cc-pvtz-jkfit", atoms)
        got, start = [], threading.Barrier(8)

        def ask():
            start.wait()
            got.append(b.handle.value)

        ts = [threading.Thread(target=ask) for _ in range(8)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        assert len(set(got)) == 1
