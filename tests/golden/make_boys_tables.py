"""Write tests/golden/boys_reference_tables.npz from the reference's Boys data.

The reference holds its Boys function tables as data, in
src/intception/boys_data/boys_data_smallx.json (151 x 33 values of F_n(x),
x = k/10, n = 0..32) and boys_data_mediumx.json (201 x 7 values, x = 15 + k/10,
n = 0..6), read by src/intception/boys.py:27-97.  This script parses those two
JSON files as data (it imports nothing from the reference) and stores the
numbers, with their grid parameters, as a numpy fixture.  Run it where the
reference is mounted:

    python tests/golden/make_boys_tables.py /root/reference
"""
import json
import os
import sys

import numpy as np


def main(ref_root):
    d = os.path.join(ref_root, "src", "intception", "boys_data")
    out = {}
    for name in ("smallx", "mediumx"):
        with open(os.path.join(d, f"boys_data_{name}.json")) as fh:
            js = json.load(fh)
        tab = np.array([[float(v) for v in row] for row in js["boys data"]], dtype=np.float64)
        assert tab.shape == (js["xpoints"], js["nvals"]), tab.shape
        out[name] = tab
        out[name + "_grid"] = np.array([float(js["xmin"]), float(js["xmax"]), float(js["xstep"]),
                                        js["ntaylor"]], dtype=np.float64)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "boys_reference_tables.npz")
    np.savez(path, **out)
    print("wrote", path, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
