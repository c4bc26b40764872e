"""The contracted Coulomb ERI class as an Intception input script.

Obara-Saika VRRs on a, c and d with one auxiliary index m, and the
Head-Gordon-Pople HRR moving angular momentum from a to b, written in the
input-script vocabulary (the reference's example script,
examples/example_input.py:45-116, shows the style for overlap and nuclear
attraction; the reference ships no ERI script).

``make_eri()`` returns the ``dsl_integral``; ``main()`` drives the generator
with the HIP backend, as a user's script would.
"""
from math import pi

from .dsl import (boys_f, dsl_cartesian_gaussian, dsl_integer_index, dsl_integral, dsl_position, dsl_scalar,
                  dsl_value, exp, sqrt)


def make_eri(vrr_indices=("a", "c", "d"), hrr=True):
    """vrr_indices: the indices built by Obara-Saika VRRs; hrr: add the HGP HRR
    moving angular momentum from a to b.  The default is the VRR+HRR scheme
    (src/intception/generator/callables/main_source_algorithms/
    contracted_vrr_hrr_with_auxiliary.py); VRRs on a, b, c, d without the HRR
    is the reference's VRR-only variant (contracted_vrr_only_with_auxiliary.py),
    the same integrals by another recurrence path."""
    ga = dsl_cartesian_gaussian("a", constant=True)
    gb = dsl_cartesian_gaussian("b", constant=True)
    gc = dsl_cartesian_gaussian("c", constant=True)
    gd = dsl_cartesian_gaussian("d", constant=True)
    m = dsl_integer_index("m")
    xa, xb, xc, xd = ga.exp(), gb.exp(), gc.exp(), gd.exp()
    A, B, C, D = ga.cen(), gb.cen(), gc.cen(), gd.cen()

    two_pi52 = dsl_value("two_pi52", 2.0 * pi ** 2.5)
    zeta_ab = dsl_scalar("zeta_ab", xa + xb)
    eta_cd = dsl_scalar("eta_cd", xc + xd)
    o_o_zpe = dsl_scalar("o_o_zpe", 1.0 / (zeta_ab + eta_cd))
    rho_zpe = dsl_scalar("rho_zpe", zeta_ab * eta_cd * o_o_zpe)
    PPab = dsl_position("PPab", (xa * A + xb * B) / zeta_ab)
    QQcd = dsl_position("QQcd", (xc * C + xd * D) / eta_cd)
    WWpq = dsl_position("WWpq", (zeta_ab * PPab + eta_cd * QQcd) * o_o_zpe)
    PA = dsl_position("PA", PPab - A)
    PB = dsl_position("PB", PPab - B)
    QC = dsl_position("QC", QQcd - C)
    QD = dsl_position("QD", QQcd - D)
    WP = dsl_position("WP", WWpq - PPab)
    WQ = dsl_position("WQ", WWpq - QQcd)
    RAB = dsl_position("RAB", A - B)
    RCD = dsl_position("RCD", C - D)
    RPQ = dsl_position("RPQ", PPab - QQcd)
    RAB2 = dsl_scalar("RAB2", RAB[0] * RAB[0] + RAB[1] * RAB[1] + RAB[2] * RAB[2])
    RCD2 = dsl_scalar("RCD2", RCD[0] * RCD[0] + RCD[1] * RCD[1] + RCD[2] * RCD[2])
    RPQ2 = dsl_scalar("RPQ2", RPQ[0] * RPQ[0] + RPQ[1] * RPQ[1] + RPQ[2] * RPQ[2])
    Tpq = dsl_scalar("Tpq", rho_zpe * RPQ2)
    o_o_2z = dsl_scalar("o_o_2z", 0.5 / zeta_ab)
    o_o_2e = dsl_scalar("o_o_2e", 0.5 / eta_cd)
    o_o_2ze = dsl_scalar("o_o_2ze", 0.5 * o_o_zpe)
    r_o_z = dsl_scalar("r_o_z", rho_zpe / zeta_ab)
    r_o_e = dsl_scalar("r_o_e", rho_zpe / eta_cd)
    Kab = dsl_scalar("Kab", exp(-xa * xb / zeta_ab * RAB2))
    Kcd = dsl_scalar("Kcd", exp(-xc * xd / eta_cd * RCD2))

    eri = dsl_integral(ga, gb, gc, gd, m, name="eri")
    I = eri.int
    eri.set_base(two_pi52 / (zeta_ab * eta_cd * sqrt(zeta_ab + eta_cd)) * Kab * Kcd * boys_f(m, Tpq))
    vrr = {
        "a": (ga, PA * I(ga - 1, gb, gc, gd, m) + WP * I(ga - 1, gb, gc, gd, m + 1)
              + o_o_2z * (ga - 1) * (I(ga - 2, gb, gc, gd, m) - r_o_z * I(ga - 2, gb, gc, gd, m + 1))
              + o_o_2z * gb * (I(ga - 1, gb - 1, gc, gd, m) - r_o_z * I(ga - 1, gb - 1, gc, gd, m + 1))
              + o_o_2ze * gc * I(ga - 1, gb, gc - 1, gd, m + 1)
              + o_o_2ze * gd * I(ga - 1, gb, gc, gd - 1, m + 1)),
        "b": (gb, PB * I(ga, gb - 1, gc, gd, m) + WP * I(ga, gb - 1, gc, gd, m + 1)
              + o_o_2z * (gb - 1) * (I(ga, gb - 2, gc, gd, m) - r_o_z * I(ga, gb - 2, gc, gd, m + 1))
              + o_o_2z * ga * (I(ga - 1, gb - 1, gc, gd, m) - r_o_z * I(ga - 1, gb - 1, gc, gd, m + 1))
              + o_o_2ze * gc * I(ga, gb - 1, gc - 1, gd, m + 1)
              + o_o_2ze * gd * I(ga, gb - 1, gc, gd - 1, m + 1)),
        "c": (gc, QC * I(ga, gb, gc - 1, gd, m) + WQ * I(ga, gb, gc - 1, gd, m + 1)
              + o_o_2e * (gc - 1) * (I(ga, gb, gc - 2, gd, m) - r_o_e * I(ga, gb, gc - 2, gd, m + 1))
              + o_o_2e * gd * (I(ga, gb, gc - 1, gd - 1, m) - r_o_e * I(ga, gb, gc - 1, gd - 1, m + 1))
              + o_o_2ze * ga * I(ga - 1, gb, gc - 1, gd, m + 1)
              + o_o_2ze * gb * I(ga, gb - 1, gc - 1, gd, m + 1)),
        "d": (gd, QD * I(ga, gb, gc, gd - 1, m) + WQ * I(ga, gb, gc, gd - 1, m + 1)
              + o_o_2e * (gd - 1) * (I(ga, gb, gc, gd - 2, m) - r_o_e * I(ga, gb, gc, gd - 2, m + 1))
              + o_o_2e * gc * (I(ga, gb, gc - 1, gd - 1, m) - r_o_e * I(ga, gb, gc - 1, gd - 1, m + 1))
              + o_o_2ze * ga * I(ga - 1, gb, gc, gd - 1, m + 1)
              + o_o_2ze * gb * I(ga, gb - 1, gc, gd - 1, m + 1)),
    }
    for k in vrr_indices:
        idx, expr = vrr[k]
        eri.add_vrr("vrr_" + k, idx, expr)
    if hrr:
        eri.add_hrr("hrr", ga, gb)
    return eri


def main(output_directory="./output", spherical=True):
    from .generator import generator, generator_options
    opt = generator_options(output_directory=output_directory, contracted=True, spherical_transformed=spherical,
                            use_cblas=False, global_user_lmax=4, backend="hip")
    gen = generator(make_eri(), opt=opt)
    gen.out()
    return gen


if __name__ == "__main__":
    main()
