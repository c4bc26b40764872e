"""Write tests/golden/eri3c_mp_golden.npz: three-centre quartets (ab|P s) with
a unit s-shell (exponent 0) from the 40-digit McMurchie-Davidson evaluation
(oracle/mp_eri.py), spherical output.

Classes (la lb | lP 0) for la, lb <= 2 and lP = 3, 4; K = 1..2 on orbital
shells, K = 1 on the auxiliary shell (exponent 10^U[log .05, log 200)),
centres in a 10-bohr cube.

    python tests/golden/make_eri3c_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from intception_amd import basis  # noqa: E402
from oracle import mp_eri  # noqa: E402

CLASSES = [(0, 0, 3, 0), (1, 0, 4, 0), (1, 1, 3, 0), (2, 0, 4, 0), (2, 1, 3, 0), (1, 2, 4, 0), (2, 2, 3, 0),
           (2, 2, 4, 0)]


def main():
    rng = np.random.Generator(np.random.PCG64(31))
    shells_all, vals, off = [], [], [0]
    for cls in CLASSES:
        for rep in range(2):
            sh = []
            for k, l in enumerate(cls[:2]):
                n = int(rng.integers(1, 3))
                e = 10.0 ** rng.uniform(-1.0, 2.0, size=n)
                sh.append({"l": l, "exps": e, "coefs": basis.normalize_contraction(e, rng.uniform(0.1, 1, n), l),
                           "cen": rng.uniform(0, 10, size=3)})
            e = 10.0 ** rng.uniform(np.log10(0.05), np.log10(200.0), size=1)
            sh.append({"l": cls[2], "exps": e, "coefs": basis.normalize_contraction(e, [1.0], cls[2]),
                       "cen": rng.uniform(0, 10, size=3)})
            sh.append({"l": 0, "exps": np.array([0.0]), "coefs": np.array([1.0]), "cen": np.zeros(3)})
            v = [float(x) for x in mp_eri.contracted_eri(sh, True, basis.SPH)]
            shells_all.append(sh)
            vals += v
            off.append(off[-1] + len(v))
            print(cls, rep, len(v), flush=True)
    L = np.array([[s["l"] for s in sh] for sh in shells_all], dtype=np.int32)
    K = np.array([[len(s["exps"]) for s in sh] for sh in shells_all], dtype=np.int32)
    E = np.array([[np.pad(s["exps"], (0, 2 - len(s["exps"]))) for s in sh] for sh in shells_all])
    C = np.array([[np.pad(s["coefs"], (0, 2 - len(s["coefs"]))) for s in sh] for sh in shells_all])
    X = np.array([[s["cen"] for s in sh] for sh in shells_all])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "eri3c_mp_golden.npz")
    np.savez(path, l=L, nprim=K, exps=E, coefs=C, cen=X, off=np.array(off, dtype=np.int64), val=np.array(vals))
    print("wrote", path, len(vals))


if __name__ == "__main__":
    main()
