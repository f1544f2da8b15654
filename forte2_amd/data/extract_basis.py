"""Extract the basis-set subset this package ships from a forte2 checkout.

The basis sets are the public Basis Set Exchange data that forte2 keeps under
forte2/data/basis/*.json (BSE JSON schema).  Only the elements and fields the
J/K workloads and the parity tests use are kept:

    python forte2_amd/data/extract_basis.py /path/to/forte2/forte2/data/basis
"""

import json
import os
import sys

SETS = ["sto-3g", "sto-6g", "cc-pvdz", "cc-pvtz", "cc-pvqz", "cc-pvtz-jkfit", "cc-pvqz-jkfit",
        "def2-universal-jkfit", "sap_helfem_large", "aug-cc-pvdz"]
ELEMENTS = ["1", "3", "6", "7", "8", "9", "10", "17", "29"]  # H Li C N O F Ne Cl Cu


def main(src):
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis")
    os.makedirs(dst, exist_ok=True)
    for name in SETS:
        with open(os.path.join(src, name + ".json")) as f:
            data = json.load(f)
        out = {"name": data.get("name", name), "elements": {}}
        for Z in ELEMENTS:
            if Z in data["elements"]:
                shells = data["elements"][Z]["electron_shells"]
                out["elements"][Z] = {"electron_shells": [
                    {"angular_momentum": s["angular_momentum"], "exponents": s["exponents"],
                     "coefficients": s["coefficients"]} for s in shells]}
        with open(os.path.join(dst, name + ".json"), "w") as f:
            json.dump(out, f, separators=(",", ":"))
        print(name, sorted(out["elements"], key=int))


if __name__ == "__main__":
    main(sys.argv[1])
