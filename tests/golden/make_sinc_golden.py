"""Write tests/golden/sinc_tables.json: the 8-tap sinc interpolation tables
of base/sinc_bob.cc for jt = 1..8, as computed by that file compiled from the
reference sources into oracle/_ref (oracle/Makefile), in the reference's
default (-O2) build and its -Ofast build (base/CMakeLists.txt:10).

Run from the repository root after `make -C oracle`:
    python tests/golden/make_sinc_golden.py
Values are stored as the hexadecimal bit patterns of the float32 entries so
the comparison stays bit-exact."""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402


def bits(a):
    return [format(int(v), "08x") for v in np.asarray(a, np.float32).view(np.uint32).ravel()]


def main():
    out = {"source": "base/sinc_bob.cc compiled into oracle/_ref", "nsinc": 8, "tables": {}}
    for variant, key in (("", "O2"), ("_ofast", "Ofast")):
        for jt in range(1, 9):
            t = oracle.ref_sinc_table(jt, variant=variant)
            if t is None:
                sys.exit("oracle/_ref%s is not built" % variant)
            out["tables"].setdefault(key, {})[str(jt)] = bits(t)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sinc_tables.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
