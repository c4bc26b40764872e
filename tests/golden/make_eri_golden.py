"""Write tests/golden/eri_mp_golden.npz: contracted ERIs from an independent
40-digit McMurchie-Davidson evaluation (oracle/mp_eri.py), used to pin the C
oracle without running the reference.

Inputs: for each class below, 3 seeded quartets of random shells (centres in
a 3-bohr cube so integrals are not negligible, K = 1..2, exponents 10^U[-1,1.5),
coefficients U[0.1,1] normalised), Cartesian and spherical.  The fixture holds
the inputs and the expected values.

    python tests/golden/make_eri_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from intception_amd import basis  # noqa: E402
from oracle import mp_eri  # noqa: E402

CLASSES = [(0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 0, 0), (1, 1, 1, 1),
           (0, 1, 0, 0), (0, 0, 1, 1), (2, 0, 0, 0), (2, 0, 1, 0), (0, 2, 1, 1),
           (2, 1, 1, 0), (2, 2, 0, 0), (1, 2, 2, 1), (3, 0, 1, 0), (0, 0, 3, 1)]


def main():
    rng = np.random.Generator(np.random.PCG64(20260))
    recs = {"l": [], "nprim": [], "exps": [], "coefs": [], "cen": [], "sph": [], "val": [], "off": [],
            "cls_id": []}
    vals = []
    off = 0
    for ci, cls in enumerate(CLASSES):
        for sph in (False, True):
            for rep in range(3):
                shells = []
                for l in cls:
                    k = int(rng.integers(1, 3)) if sum(cls) < 5 else 1
                    e = 10.0 ** rng.uniform(-1.0, 1.5, size=k)
                    c = basis.normalize_contraction(e, rng.uniform(0.1, 1.0, size=k), l)
                    shells.append({"l": l, "exps": e, "coefs": c, "cen": rng.uniform(0.0, 3.0, size=3)})
                v = [float(x) for x in mp_eri.contracted_eri(shells, sph, basis.SPH)]
                for s in shells:
                    recs["l"].append(s["l"])
                    recs["nprim"].append(len(s["exps"]))
                    recs["exps"].append(np.pad(s["exps"], (0, 2 - len(s["exps"]))))
                    recs["coefs"].append(np.pad(s["coefs"], (0, 2 - len(s["coefs"]))))
                    recs["cen"].append(s["cen"])
                recs["sph"].append(int(sph))
                recs["off"].append(off)
                recs["cls_id"].append(ci)
                vals += v
                off += len(v)
                print(cls, sph, rep, len(v), flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "eri_mp_golden.npz")
    np.savez(path, l=np.array(recs["l"], dtype=np.int32), nprim=np.array(recs["nprim"], dtype=np.int32),
             exps=np.array(recs["exps"]), coefs=np.array(recs["coefs"]), cen=np.array(recs["cen"]),
             sph=np.array(recs["sph"], dtype=np.int32), off=np.array(recs["off"] + [off], dtype=np.int64),
             cls_id=np.array(recs["cls_id"], dtype=np.int32), classes=np.array(CLASSES, dtype=np.int32),
             val=np.array(vals))
    print("wrote", path, len(vals), "values")


if __name__ == "__main__":
    main()
