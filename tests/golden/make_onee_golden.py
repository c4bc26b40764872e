"""Write tests/golden/onee_mp_golden.npz: primitive Cartesian overlap (a|b)
and nuclear attraction (a|1/r_C|b) blocks from the 40-digit
McMurchie-Davidson evaluation (oracle/mp_eri.py), for every (la, lb) with
la, lb <= 3, two pairs each: exponents 10^U[-1, 2), centres in a 4-bohr cube
(nuclear centre C in the same cube).

    python tests/golden/make_onee_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import mp_eri  # noqa: E402


def main():
    rng = np.random.Generator(np.random.PCG64(47))
    rows, ov, nu, off = [], [], [], [0]
    for la in range(4):
        for lb in range(4):
            for rep in range(2):
                xa, xb = 10.0 ** rng.uniform(-1.0, 2.0, size=2)
                A, B, C = (rng.uniform(0, 4, size=3) for _ in range(3))
                for ca in mp_eri.cart_components(la):
                    for cb in mp_eri.cart_components(lb):
                        ov.append(float(mp_eri.primitive_overlap(ca, cb, xa, xb, A, B)))
                        nu.append(float(mp_eri.primitive_nuclear(ca, cb, xa, xb, A, B, C)))
                rows.append([la, lb, xa, xb, *A, *B, *C])
                off.append(len(ov))
    np.savez(os.path.join(ROOT, "tests", "golden", "onee_mp_golden.npz"), pairs=np.array(rows),
             offsets=np.array(off, dtype=np.int64), overlap=np.array(ov), nuclear=np.array(nu))
    print("%d pairs, %d elements" % (len(rows), len(ov)))


if __name__ == "__main__":
    main()
