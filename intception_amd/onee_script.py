"""The 2-index overlap and nuclear attraction integrals as input scripts.

These are the integral kinds of the reference's example input script
(examples/example_input.py:79-110): the same definitions -- base expressions
and Obara-Saika VRRs, with the nuclear attraction's point charge at the centre
of a third Gaussian index -- written in this build's input-script vocabulary.
``generator(make_overlap(), make_nuclear(), opt=...)`` recognises them and binds
them to libintc's one-electron kernels.
"""
from math import pi

from .dsl import (boys_f, dsl_cartesian_gaussian, dsl_integer_index, dsl_integral, dsl_position, dsl_scalar,
                  dsl_value, exp, sqrt)


def _common():
    ga = dsl_cartesian_gaussian("a", constant=True)
    gb = dsl_cartesian_gaussian("b", constant=True)
    gc = dsl_cartesian_gaussian("c", constant=True)
    xa, xb = ga.exp(), gb.exp()
    A, B, C = ga.cen(), gb.cen(), gc.cen()
    v = {}
    v["zeta"] = dsl_scalar("zeta", xa + xb)
    v["o_o_zeta"] = dsl_scalar("o_o_zeta", 1.0 / v["zeta"])
    v["mu"] = dsl_scalar("mu", xa * xb * v["o_o_zeta"])
    v["o_o_2zeta"] = dsl_scalar("o_o_2zeta", 0.5 * v["o_o_zeta"])
    RAB = dsl_position("RAB", A - B)
    v["PA"] = dsl_position("PA", xb * v["o_o_zeta"] * -RAB)
    v["PB"] = dsl_position("PB", xa * v["o_o_zeta"] * RAB)
    v["PC"] = dsl_position("PC", xa * v["o_o_zeta"] * (A - C) + xb * v["o_o_zeta"] * (B - C))
    v["RAB2"] = dsl_scalar("RAB2", RAB[0] * RAB[0] + RAB[1] * RAB[1] + RAB[2] * RAB[2])
    PC = v["PC"]
    v["RPC2"] = dsl_scalar("RPC2", PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2])
    return ga, gb, gc, C, v


def make_overlap():
    ga, gb, gc, C, v = _common()
    pi3 = dsl_value("pi3", pi ** 3)
    ooz3 = dsl_scalar("ooz3", v["o_o_zeta"] * v["o_o_zeta"] * v["o_o_zeta"])
    s = dsl_integral(ga, gb, name="overlap_2idx")
    I = s.int
    s.set_base(sqrt(pi3 * ooz3) * exp(-v["mu"] * v["RAB2"]))
    h = v["o_o_2zeta"]
    s.add_vrr("vrr_a", ga, v["PA"] * I(ga - 1, gb) + h * (ga - 1) * I(ga - 2, gb) + h * gb * I(ga - 1, gb - 1))
    s.add_vrr("vrr_b", gb, v["PB"] * I(ga, gb - 1) + h * (gb - 1) * I(ga, gb - 2) + h * ga * I(ga - 1, gb - 1))
    return s


def make_nuclear(hrr=False):
    """VRRs on b then a (the example's order); hrr=True instead builds a by
    its VRR and moves angular momentum to b with the HGP HRR."""
    ga, gb, gc, C, v = _common()
    m = dsl_integer_index("m")
    n = dsl_integral(ga, gb, m, C, name="nuclear_2idx")
    I = n.int
    n.set_base(2.0 * (pi * v["o_o_zeta"]) * exp(-v["mu"] * v["RAB2"]) * boys_f(m, v["zeta"] * v["RPC2"]))
    h = v["o_o_2zeta"]
    vrr_b = (v["PB"] * I(ga, gb - 1, m) - v["PC"] * I(ga, gb - 1, m + 1)
             + h * (gb - 1) * (I(ga, gb - 2, m) - I(ga, gb - 2, m + 1))
             + h * ga * (I(ga - 1, gb - 1, m) - I(ga - 1, gb - 1, m + 1)))
    vrr_a = (v["PA"] * I(ga - 1, gb, m) - v["PC"] * I(ga - 1, gb, m + 1)
             + h * gb * (I(ga - 1, gb - 1, m) - I(ga - 1, gb - 1, m + 1))
             + h * (ga - 1) * (I(ga - 2, gb, m) - I(ga - 2, gb, m + 1)))
    if hrr:
        n.add_vrr("vrr_a", ga, vrr_a)
        n.add_hrr("hrr", ga, gb)
    else:
        n.add_vrr("vrr_b", gb, vrr_b)
        n.add_vrr("vrr_a", ga, vrr_a)
    return n


def main(output_directory="./output", contracted=False, spherical=False):
    """The example's generator call: primitive Cartesian by default."""
    from .generator import generator, generator_options
    opt = generator_options(output_directory=output_directory, contracted=contracted,
                            spherical_transformed=spherical, backend="hip")
    gen = generator(make_overlap(), make_nuclear(), opt=opt)
    gen.out()
    return gen


if __name__ == "__main__":
    main()
