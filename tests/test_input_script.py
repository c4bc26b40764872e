"""Input-script API: the Coulomb ERI script drives the HIP-backend generator.

CPU: the generator accepts the Obara-Saika/HGP ERI script, rejects scripts with
a dropped or wrong recurrence term, and writes the header that maps the
reference's function names onto the libintc C-ABI.
GPU: the bound per-quartet twin (reference argument layout, n_cont > 1,
caller strides) and the batched plan agree with the oracle.
"""
import os

import numpy as np
import pytest

from conftest import assert_parity
from intception_amd import basis, eri_script
from intception_amd.dsl import dsl_cartesian_gaussian
from intception_amd.generator import generator, generator_options, recognize_eri


def test_eri_script_recognised():
    eri = eri_script.make_eri()
    assert recognize_eri(eri)
    assert eri.full_name() == "intception_integral_eri_abcd_m"


def test_vrr_only_variant_recognised():
    """VRRs on a, b, c and d with no HRR (the reference's
    contracted_vrr_only_with_auxiliary algorithm): the same integral family,
    bound to the same kernels; a VRR-only script missing one index is
    rejected."""
    eri = eri_script.make_eri(("a", "b", "c", "d"), hrr=False)
    assert recognize_eri(eri)
    assert [r.rrtype() for r in eri.rr_list()] == ["vrr"] * 4
    gen = generator(eri, opt=generator_options(contracted=True, spherical_transformed=True, backend="hip"))
    assert gen.names() == ["intception_integral_eri_abcd_m_cont_sph"]
    with pytest.raises(NotImplementedError):
        recognize_eri(eri_script.make_eri(("a", "c", "d"), hrr=False))


def test_broken_scripts_rejected():
    import intception_amd.dsl as d
    eri = eri_script.make_eri()
    # drop the W-P term from the VRR on a
    r = eri.rr_list()[0]
    bad = d.dsl_integral(*eri.index_list(), *eri.aux_list(), name="eri")
    bad.set_base(eri.base())
    ga, gb, gc, gd = eri.index_list()
    m = eri.aux_list()[0]
    expr = r.expr()

    def find(node, name):
        if isinstance(node, d.dsl_position) and node.name() == name:
            return node
        for ch in (getattr(node, "_l", None), getattr(node, "_r", None), getattr(node, "_a", None)):
            if ch is not None:
                hit = find(ch, name)
                if hit is not None:
                    return hit
        return None

    WP = find(expr, "WP")
    assert WP is not None
    # halve the (W-P)_i (a-1_i)^(m+1) term
    bad.add_vrr("vrr_a", ga, expr - 0.5 * WP * bad.int(ga - 1, gb, gc, gd, m + 1))
    for rr in eri.rr_list()[1:]:
        if rr.rrtype() == "vrr":
            bad.add_vrr(rr.name(), rr.changing_index(), rr.expr())
    bad.add_hrr("hrr", ga, gb)
    with pytest.raises(NotImplementedError, match="not the Obara-Saika"):
        recognize_eri(bad)
    # wrong base (missing K_CD)
    bad2 = eri_script.make_eri()
    bad2.set_base(bad2.base() * 1.0001)
    with pytest.raises(NotImplementedError):
        recognize_eri(bad2)
    # incomplete: d never built
    with pytest.raises(NotImplementedError):
        recognize_eri(eri_script.make_eri(("a", "c")))


def test_generator_out_writes_header(tmp_path):
    opt = generator_options(output_directory=str(tmp_path), contracted=True, spherical_transformed=False,
                            use_cblas=False, global_user_lmax=4, backend="hip")
    gen = generator(eri_script.make_eri(), opt=opt)
    man = gen.out()
    name = "intception_integral_eri_abcd_m_cont_cart"
    assert list(man) == [name]
    h = (tmp_path / (name + ".h")).read_text()
    assert "#define %s_main intc_eri_cont_cart_main" % name in h
    assert gen.binding().work_array_size(1, 1, 1, 1, 3, 3, 3, 3) == 25930


def test_cartesian_requires_contracted_for_sph():
    with pytest.raises(ValueError):
        generator_options(contracted=False, spherical_transformed=True)


@pytest.mark.gpu
@pytest.mark.parametrize("sph", [False, True])
def test_bound_main_matches_oracle(gpu, sph):
    from oracle import oracle
    oracle.set_boys_mode(oracle.BOYS_EXACT)
    gen = generator(eri_script.make_eri(), opt=generator_options(spherical_transformed=sph, backend="hip"))
    b = gen.binding()
    rng = np.random.default_rng(17)
    for cls in [(0, 0, 0, 0), (1, 0, 2, 1), (2, 2, 1, 0), (1, 2, 2, 2)]:
        shells = []
        for l in cls:
            k = int(rng.integers(1, 4))
            e = 10 ** rng.uniform(-1, 2, size=k)
            c = np.stack([basis.normalize_contraction(e, rng.uniform(0.1, 1, size=k), l) for _ in range(2)], 1)
            shells.append({"l": l, "exps": e, "coefs": c, "cen": rng.uniform(0, 10, size=3), "sph": basis.SPH[l]})
        ns = [basis.nsph(l) if sph else basis.ncc(l) for l in cls]
        blk = int(np.prod(ns))
        strides = [ns[1] * ns[2] * ns[3], ns[2] * ns[3], ns[3], 1]
        cstr = [blk, 2 * blk, 4 * blk, 8 * blk]
        out = np.zeros(16 * blk)
        b.main(shells, out, strides=strides, cont_strides=cstr)
        ref = oracle.eri_quartet(shells, sph, basis.SPH, strides=strides, ncont_strides=cstr, out_size=16 * blk)
        assert_parity(out, ref, what="bound main %s sph=%s" % (cls, sph))


def test_onee_scripts_recognised(tmp_path):
    """The example script's integral kinds: the 2-index overlap and nuclear
    attraction (VRR-only, and the nuclear attraction with the HRR) are
    recognised, a dropped -(P-C) term and a wrong base are rejected, and the
    generator maps the reference names onto the one-electron C-ABI."""
    import json
    from intception_amd import onee_script
    from intception_amd import dsl as d
    from intception_amd.generator import recognize_onee
    assert recognize_onee(onee_script.make_overlap()) == "overlap"
    assert recognize_onee(onee_script.make_nuclear()) == "nuclear"
    assert recognize_onee(onee_script.make_nuclear(hrr=True)) == "nuclear"
    good = onee_script.make_nuclear()
    bad = d.dsl_integral(*good.index_list(), *good.aux_list(), *good.arg_list(), name="nuclear_2idx")
    bad.set_base(good.base())
    ga, gb = good.index_list()
    m = good.aux_list()[0]
    vb, va = good.rr_list()
    bad.add_vrr("vrr_b", gb, vb.expr())
    # drop -(P-C)_i (a-1|b)^(m+1): the overlap recurrence in the nuclear integral
    PC = onee_script._common()[4]["PC"]
    bad.add_vrr("vrr_a", ga, va.expr() + PC * bad.int(ga - 1, gb, m + 1))
    with pytest.raises(NotImplementedError, match="not the Obara-Saika nuclear"):
        recognize_onee(bad)
    wrong = onee_script.make_overlap()
    wrong.set_base(wrong.base() * 2.0)
    with pytest.raises(NotImplementedError):
        recognize_onee(wrong)
    gen = generator(onee_script.make_overlap(), onee_script.make_nuclear(),
                    opt=generator_options(output_directory=str(tmp_path), contracted=False,
                                          spherical_transformed=False))
    man = gen.out()
    assert set(man) == {"intception_integral_overlap_2idx_ab_prim_cart",
                        "intception_integral_nuclear_2idx_ab_m_cc_prim_cart"}
    hdr = open(tmp_path / "intception_integral_nuclear_2idx_ab_m_cc_prim_cart.h").read()
    assert "#define intception_integral_nuclear_2idx_ab_m_cc_prim_cart_main intc_nuclear_prim_cart_main" in hdr
    assert json.load(open(tmp_path / "intception_hip_manifest.json"))[
        "intception_integral_overlap_2idx_ab_prim_cart"]["kind"] == "overlap"


def test_generator_options_refuse_what_they_cannot_do():
    """use_cblas=True (broken upstream, no CBLAS path here) and unknown
    options raise instead of being ignored; the reference's other options are
    accepted."""
    with pytest.raises(NotImplementedError, match="CBLAS"):
        generator_options(use_cblas=True)
    with pytest.raises(TypeError, match="precision"):
        generator_options(precision="mixed")
    generator_options(stdout_width=100, add_suffix_to_global=True, silent=True)
