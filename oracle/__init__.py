"""CPU oracle for the density-fitted J/K path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy FP64, the algorithms of the reference that the
device path replaces (forte2/jkbuilder/jkbuilder.py, forte2/scf/*.py,
forte2/helpers/diis.py, forte2/helpers/matrix_functions.py).  It is the
checker: only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it.  The product package ``forte2_amd`` never
imports or calls anything here.

Every function takes plain arrays (integrals, coefficients) and returns plain
arrays, so the oracle shares no code with the device path.  Each function's
docstring cites the reference lines it restates and names the tests that pin
it (golden values from the reference's own test-suite, closed forms,
brute-force contractions).
"""
