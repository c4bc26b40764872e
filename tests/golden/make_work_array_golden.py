"""Write tests/golden/work_array_size_golden.npz: the reference's own
<full>_work_array_size for the ERI input script, evaluated over a grid of
angular momenta and primitive/contraction counts.

The work-array size is the one function of the ERI path that nothing but the
reference itself pins (it is a memory-layout formula of the reference's
2-layer work array).  This script imports the reference generator in this
container only (/root/reference/src), generates the C99 for the ERI script
(Obara-Saika VRR on a, c, d with one auxiliary index, HGP HRR a -> b) into a
temporary directory, compiles it with gcc and calls the generated
work_array_size function through ctypes.  Only the resulting integers are
committed; no reference source or generated code is kept.

    python tests/golden/make_work_array_golden.py

The reference uses time.clock, which Python 3.10 lacks; it is aliased to
time.perf_counter before the import (the generator only times itself).
"""
import ctypes
import glob
import itertools
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = "/root/reference/src"


def eri_script(intc):
    """The ERI class in the reference's input-script vocabulary."""
    from math import pi
    ga = intc.dsl_cartesian_gaussian("a", constant=True)
    gb = intc.dsl_cartesian_gaussian("b", constant=True)
    gc = intc.dsl_cartesian_gaussian("c", constant=True)
    gd = intc.dsl_cartesian_gaussian("d", constant=True)
    m = intc.dsl_integer_index(name="m")
    xa, xb, xc, xd = ga.exp(), gb.exp(), gc.exp(), gd.exp()
    A, B, C, D = ga.cen(), gb.cen(), gc.cen(), gd.cen()
    S, P, V = intc.dsl_scalar, intc.dsl_position, intc.dsl_value
    two_pi52 = V(name="two_pi52", value=2.0 * pi ** 2.5, vartype="double")
    zeta_ab = S("zeta_ab", xa + xb)
    eta_cd = S("eta_cd", xc + xd)
    o_o_zpe = S("o_o_zpe", 1.0 / (zeta_ab + eta_cd))
    rho_zpe = S("rho_zpe", zeta_ab * eta_cd * o_o_zpe)
    PPab = P("PPab", (xa * A + xb * B) / zeta_ab)
    QQcd = P("QQcd", (xc * C + xd * D) / eta_cd)
    WWpq = P("WWpq", (zeta_ab * PPab + eta_cd * QQcd) * o_o_zpe)
    PA = P("PA", PPab - A)
    QC = P("QC", QQcd - C)
    QD = P("QD", QQcd - D)
    WP = P("WP", WWpq - PPab)
    WQ = P("WQ", WWpq - QQcd)
    RAB = P("RAB", A - B)
    RCD = P("RCD", C - D)
    RPQ = P("RPQ", PPab - QQcd)
    RAB2 = S("RAB2", RAB[0] * RAB[0] + RAB[1] * RAB[1] + RAB[2] * RAB[2])
    RCD2 = S("RCD2", RCD[0] * RCD[0] + RCD[1] * RCD[1] + RCD[2] * RCD[2])
    RPQ2 = S("RPQ2", RPQ[0] * RPQ[0] + RPQ[1] * RPQ[1] + RPQ[2] * RPQ[2])
    Tpq = S("Tpq", rho_zpe * RPQ2)
    o_o_2z = S("o_o_2z", 0.5 / zeta_ab)
    o_o_2e = S("o_o_2e", 0.5 / eta_cd)
    o_o_2ze = S("o_o_2ze", 0.5 * o_o_zpe)
    r_o_z = S("r_o_z", rho_zpe / zeta_ab)
    r_o_e = S("r_o_e", rho_zpe / eta_cd)
    Kab = S("Kab", intc.exp(-xa * xb / zeta_ab * RAB2))
    Kcd = S("Kcd", intc.exp(-xc * xd / eta_cd * RCD2))
    eri = intc.dsl_integral(ga, gb, gc, gd, m, name="eri")
    I = eri.int
    eri.set_base(two_pi52 / (zeta_ab * eta_cd * intc.sqrt(zeta_ab + eta_cd)) * Kab * Kcd * intc.boys_f(m, Tpq))
    eri.add_vrr("vrr_a", ga, PA * I(ga - 1, gb, gc, gd, m) + WP * I(ga - 1, gb, gc, gd, m + 1)
                + o_o_2z * (ga - 1) * (I(ga - 2, gb, gc, gd, m) - r_o_z * I(ga - 2, gb, gc, gd, m + 1))
                + o_o_2z * gb * (I(ga - 1, gb - 1, gc, gd, m) - r_o_z * I(ga - 1, gb - 1, gc, gd, m + 1))
                + o_o_2ze * gc * I(ga - 1, gb, gc - 1, gd, m + 1)
                + o_o_2ze * gd * I(ga - 1, gb, gc, gd - 1, m + 1))
    eri.add_vrr("vrr_c", gc, QC * I(ga, gb, gc - 1, gd, m) + WQ * I(ga, gb, gc - 1, gd, m + 1)
                + o_o_2e * (gc - 1) * (I(ga, gb, gc - 2, gd, m) - r_o_e * I(ga, gb, gc - 2, gd, m + 1))
                + o_o_2e * gd * (I(ga, gb, gc - 1, gd - 1, m) - r_o_e * I(ga, gb, gc - 1, gd - 1, m + 1))
                + o_o_2ze * ga * I(ga - 1, gb, gc - 1, gd, m + 1)
                + o_o_2ze * gb * I(ga, gb - 1, gc - 1, gd, m + 1))
    eri.add_vrr("vrr_d", gd, QD * I(ga, gb, gc, gd - 1, m) + WQ * I(ga, gb, gc, gd - 1, m + 1)
                + o_o_2e * (gd - 1) * (I(ga, gb, gc, gd - 2, m) - r_o_e * I(ga, gb, gc, gd - 2, m + 1))
                + o_o_2e * gc * (I(ga, gb, gc - 1, gd - 1, m) - r_o_e * I(ga, gb, gc - 1, gd - 1, m + 1))
                + o_o_2ze * ga * I(ga - 1, gb, gc, gd - 1, m + 1)
                + o_o_2ze * gb * I(ga, gb - 1, gc, gd - 1, m + 1))
    eri.add_hrr("hrr", ga, gb)
    return eri


def main():
    time.clock = time.perf_counter
    sys.path.insert(0, REF)
    import intception as intc
    grid = []
    results = {}
    with tempfile.TemporaryDirectory() as tmp:
        for sph in (False, True):
            out = os.path.join(tmp, "sph" if sph else "cart")
            os.makedirs(out)
            opt = intc.generator_options(contracted=True, spherical_transformed=sph, use_cblas=False,
                                         global_user_lmax=3, output_directory=out, silent=True)
            gen = intc.generator(eri_script(intc), opt=opt)
            gen.out()
            srcs = sorted(glob.glob(os.path.join(out, "*.c")))
            so = os.path.join(tmp, "ref_%d.so" % sph)
            subprocess.run(["gcc", "-std=c99", "-O1", "-fPIC", "-shared", "-fcommon", "-I", out, "-o", so] + srcs +
                           ["-lm"], check=True)
            L = ctypes.CDLL(so)
            name = [n for n in open(os.path.join(out, [f for f in os.listdir(out) if f.endswith(".h") and
                                                       "eri" in f][0])).read().split()
                    if n.endswith("_work_array_size(")]
            fn_name = name[0].rstrip("(") if name else None
            if fn_name is None:
                raise RuntimeError("no work_array_size prototype in the generated header")
            f = getattr(L, fn_name)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_int] * 12
            vals = []
            for ls in itertools.product(range(3), repeat=4):
                for ks in [(1, 1, 1, 1), (3, 3, 3, 3), (1, 2, 3, 4), (6, 1, 5, 2)]:
                    for kc in [(1, 1, 1, 1), (2, 1, 1, 2)]:
                        vals.append([int(sph), *ls, *ks, *kc, f(*ls, *ks, *kc)])
            results[sph] = vals
            print("spherical %s: %s, %d values" % (sph, fn_name, len(vals)))
    arr = np.array(results[False] + results[True], dtype=np.int64)
    np.savez(os.path.join(ROOT, "tests", "golden", "work_array_size_golden.npz"), rows=arr,
             columns=np.array(["spherical", "la", "lb", "lc", "ld", "na_prim", "nb_prim", "nc_prim", "nd_prim",
                               "na_cont", "nb_cont", "nc_cont", "nd_cont", "size"]))


if __name__ == "__main__":
    main()
