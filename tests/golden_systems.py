"""Inputs of the reference's RHF golden energies used by both the oracle
tests and the device tests (the energies themselves are in
golden/reference_values.json)."""

# RHF reference energies of other reference tests (their CASSCF/DSRG parts
# are not on this path): geometry, basis, fit set, unit, fixture key, SCF options
# (and the charge, 0 unless given)
MORE_RHF = [
    ("N 0 0 0\nN 0 0 2.0", "cc-pvdz", "cc-pvtz-jkfit", "bohr", "n2_dz_r2bohr_rhf", {}),
    ("N 0 0 0\nN 0 0 1.4", "cc-pvdz", "cc-pvtz-jkfit", "angstrom", "n2_dz_rhf", dict(e_tol=1e-12)),
    ("O 0 0 -0.069592187400\nH 0 -0.783151105291 0.552239257834\nH 0 0.783151105291 0.552239257834",
     "cc-pvdz", "def2-universal-jkfit", "angstrom", "water_dz_universal_rhf", dict(e_tol=1e-12, d_tol=1e-12)),
    ("H 0 0 -0.949624435830\nF 0 0 0.050375564170", "cc-pvdz", "cc-pvtz-jkfit", "angstrom", "hf_dz_rhf_eq",
     dict(e_tol=1e-12)),
    ("F 0 0 -0.075563346255\nH 0 0 1.424436653745", "cc-pvdz", "def2-universal-jkfit", "angstrom",
     "hf_dz_universal_rhf", dict(e_tol=1e-12)),
    # a closed-shell anion of a 3d metal: f orbital shells, aux shells up to i
    # (l = 6) and the SAP guess of Z = 29
    ("Cu 0 0 0", "cc-pvdz", "def2-universal-jkfit", "angstrom", "cu_anion_dz_rhf", dict(charge=-1)),
    # the cc-pVQZ water of the golden energy, rotated and shifted in space
    ("O 0.000000000000 0.030832298694 -0.053403107852\nH 0.000000000000 -0.860947008954 0.067962729394\n"
     "H 0.000000000000 0.371616054311 0.779583345763", "cc-pvqz", "cc-pvqz-jkfit", "angstrom",
     "water_qz_rhf_rotated", {}),
    # the RHF energies of the reference's GAS-CI tests: STO-6G (six-primitive
    # contractions), and a water of C_s symmetry in cc-pVTZ
    ("H 0 0 -0.375\nH 0 0 0.375", "sto-6g", "def2-universal-jkfit", "angstrom", "h2_sto6g_rhf", dict(e_tol=1e-12)),
    ("H 0.0 0.0 0.0\nF 0.0 0.0 1.0", "cc-pvdz", "def2-universal-jkfit", "angstrom", "hf_1a_dz_universal_rhf",
     dict(e_tol=1e-12, d_tol=1e-8)),
    ("O 0.000000000000 0.000000000000 -0.069592187400\nH 0.000000000000 -0.783151105291 0.552239257834\n"
     "H 0.000000000000 0.783151105291 0.552239257834", "sto-6g", "def2-universal-jkfit", "angstrom",
     "water_sto6g_rhf", dict(e_tol=1e-12, d_tol=1e-8)),
    ("O 0.0000000000 -0.0000000000 -0.0662628033\nH 0.0000000000 -0.7540256101 0.5259060578\n"
     "H -0.0000000000 0.7530256101 0.5260060578", "cc-pvtz", "def2-universal-jkfit", "angstrom",
     "water_cs_tz_universal_rhf", dict(e_tol=1e-12, d_tol=1e-12)),
]

# RHF with the Cholesky-TEI B (cc-pVDZ, tol 1e-10) of the reference's
# point-group tests (their orbital labels are not on this path): geometry,
# unit, fixture key
CHOLESKY_RHF = [
    ("N 0 0 0\nN 1.128 0 0", "angstrom", "n2_dz_1128_rhf_cholesky"),
    ("N 0 0 0\nN 0 0 1.128", "angstrom", "n2_dz_1128_rhf_cholesky"),
    ("C -1.2916277126 -1.4862694893 0.0\nC 1.2916277126 -1.4862694893 0.0\n"
     "C -1.2916277126 1.4862694893 0.0\nC 1.2916277126 1.4862694893 0.0\n"
     "H -2.7546827497 -2.9442264047 0.0\nH 2.7546827497 -2.9442264047 0.0\n"
     "H -2.7546827497 2.9442264047 0.0\nH 2.7546827497 2.9442264047 0.0", "bohr", "cbd_dz_rhf_cholesky"),
]
CHOLESKY_IDS = ["n2_along_x", "n2_along_z", "cyclobutadiene"]
