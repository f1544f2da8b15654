"""The golden-value fixture (tests/golden/reference_values.json): every value
names the reference test that holds it, and every value is finite."""

import json
import math
import os

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_values.json")


def test_every_value_has_a_source():
    g = json.load(open(PATH))
    for group in ("energies", "norms"):
        for name, entry in g[group].items():
            assert entry["source"].startswith("tests/"), name
            nums = [v for k, v in entry.items() if k not in ("source", "rel")]
            assert nums and all(math.isfinite(v) for v in nums), name
    eri = g["sto3g_water_eri"]
    assert eri["source"].startswith("tests/") and len(eri["values"]) == 25
    assert all(len(k) == 4 and math.isfinite(v) for k, v in eri["values"])
