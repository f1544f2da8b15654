"""Inputs of the reference's RHF golden energies used by both the oracle
tests and the device tests (the energies themselves are in
golden/reference_values.json)."""

# RHF reference energies of other reference tests (their CASSCF/DSRG parts
# are not on this path): geometry, basis, fit set, unit, fixture key, SCF options
MORE_RHF = [
    ("N 0 0 0\nN 0 0 2.0", "cc-pvdz", "cc-pvtz-jkfit", "bohr", "n2_dz_r2bohr_rhf", {}),
    ("N 0 0 0\nN 0 0 1.4", "cc-pvdz", "cc-pvtz-jkfit", "angstrom", "n2_dz_rhf", dict(e_tol=1e-12)),
    ("O 0 0 -0.069592187400\nH 0 -0.783151105291 0.552239257834\nH 0 0.783151105291 0.552239257834",
     "cc-pvdz", "def2-universal-jkfit", "angstrom", "water_dz_universal_rhf", dict(e_tol=1e-12, d_tol=1e-12)),
    ("H 0 0 -0.949624435830\nF 0 0 0.050375564170", "cc-pvdz", "cc-pvtz-jkfit", "angstrom", "hf_dz_rhf_eq",
     dict(e_tol=1e-12)),
    ("F 0 0 -0.075563346255\nH 0 0 1.424436653745", "cc-pvdz", "def2-universal-jkfit", "angstrom",
     "hf_dz_universal_rhf", dict(e_tol=1e-12)),
]
