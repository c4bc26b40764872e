"""GPU: one-electron kinds of the input-script API (2-index overlap and
nuclear attraction of the reference's example script) against the oracle.

The device kernels build (e|0) by the VRR on a and move angular momentum to b
with the HRR on contracted values; the oracle (oracle/onee_oracle.c) follows
the script's VRR-only recursions, and is itself pinned to a 40-digit
McMurchie-Davidson evaluation.  Bar: |d| <= 1e-14 + 1e-12 |ref| relative to
each block's largest element (the HRR moves angular momentum across |AB| on
contracted values, so an element is as exact as its block's scale):
|d| <= 1e-12 |ref| + 1e-13 max|block| + 1e-300.

* batched plans, every (la, lb) with l <= 3, K 1..4, Cartesian and spherical,
  both kinds, nuclear attraction with both Boys modes;
* the per-call twins with the reference's argument layout: primitive
  Cartesian (the example's contracted=False functions) and contracted
  Cartesian/spherical with n_cont = 2 and caller strides;
* the generator's binding of the example-style scripts.
"""
import numpy as np
import pytest

from intception_amd import basis, runtime
from oracle import oracle

pytestmark = pytest.mark.gpu


class _Shells:
    def __init__(self, seed, n, lmax=3, kmax=4, box=5.0):
        rng = np.random.Generator(np.random.PCG64(seed))
        ls = rng.integers(0, lmax + 1, size=n)
        nps = rng.integers(1, kmax + 1, size=n)
        (self.shell_l, self.shell_nprim, self.shell_off, self.shell_cen, self.exps,
         self.coefs) = basis.random_shells(rng, ls, nps, box=box)


def _check_blocks(got, ref, sizes, offs, what):
    worst = (0.0, -1)
    for q in range(len(sizes)):
        g = got[offs[q]:offs[q] + sizes[q]]
        r = ref[offs[q]:offs[q] + sizes[q]]
        ratio = (np.abs(g - r) / (1e-12 * np.abs(r) + 1e-13 * np.abs(r).max() + 1e-300)).max()
        worst = max(worst, (float(ratio), q))
    print("  %s: worst |d| / bar = %.3g (item %d)" % (what, worst[0], worst[1]), flush=True)
    assert worst[0] <= 1.0, "%s: item %d outside the bar" % (what, worst[1])


@pytest.mark.parametrize("spherical", [False, True])
@pytest.mark.parametrize("kind", [0, 1])
def test_onee_plan_matches_oracle(gpu, kind, spherical):
    import torch
    sh = _Shells(17 + kind, 60)
    rng = np.random.default_rng(5)
    pairs = np.array([(a, b) for a in range(60) for b in range(60) if rng.random() < 0.3], dtype=np.int32)
    centres = rng.uniform(0, 5, size=(7, 3))
    items = pairs if kind == 0 else np.concatenate([pairs, rng.integers(0, 7, size=(len(pairs), 1))], axis=1)
    items = np.ascontiguousarray(items, dtype=np.int32)
    print("onee plan: kind %d spherical %s, %d items over l <= 3" % (kind, spherical, len(items)), flush=True)
    modes = [(0, oracle.BOYS_EXACT)] if kind == 0 else [(0, oracle.BOYS_EXACT), (1, oracle.BOYS_REFERENCE)]
    for dev_mode, ora_mode in modes:
        runtime.set_boys_mode(dev_mode)
        try:
            plan = runtime.OneEPlan(kind, sh, items, centres=centres if kind else None, spherical=spherical)
            got = plan.run().cpu().numpy()
            torch.cuda.synchronize()
            plan.close()
        finally:
            runtime.set_boys_mode(0)
        oracle.set_boys_mode(ora_mode)
        ref = oracle.onee_batch(kind, spherical, sh, items, plan.out_off, plan.out_size, centres, basis.SPH)
        _check_blocks(got, ref, plan.sizes, plan.out_off, "kind %d sph %s boys %d" % (kind, spherical, dev_mode))


def test_primitive_twins_match_oracle(gpu):
    """intc_overlap_prim_cart_main / intc_nuclear_prim_cart_main: the
    reference example's functions (primitive, Cartesian, caller strides)."""
    from intception_amd import lib
    import ctypes
    L = lib.load()
    dp = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(8)
    oracle.set_boys_mode(oracle.BOYS_EXACT)
    for la in range(4):
        for lb in range(4):
            xa, xb = 10 ** rng.uniform(-1, 1.5, size=2)
            A, B, C = (np.ascontiguousarray(rng.uniform(0, 3, size=3)) for _ in range(3))
            na, nb = basis.ncc(la), basis.ncc(lb)
            for kind in (0, 1):
                out = np.full(na * nb * 2, np.nan)
                # transposed layout with gaps: element (ia, ib) at ib * (2 na) + ia
                args = [xa, A.ctypes.data_as(dp), la, na, 1, xb, B.ctypes.data_as(dp), lb, nb, 2 * na]
                if kind == 0:
                    lib.check(L.intc_overlap_prim_cart_main(*args, None, out.ctypes.data_as(dp)))
                else:
                    lib.check(L.intc_nuclear_prim_cart_main(*args, C.ctypes.data_as(dp), None,
                                                            out.ctypes.data_as(dp)))
                got = out.reshape(nb, 2 * na)[:, :na].T.reshape(-1)
                ref = oracle.onee_prim(kind, la, lb, xa, A, xb, B, C)
                _check_blocks(got, ref, [na * nb], [0], "prim kind %d (%d %d)" % (kind, la, lb))


def test_bindings_contracted_and_primitive(gpu):
    """The generator's bindings: contracted spherical with n_cont = 2 and
    caller strides, and the primitive Cartesian functions of the example."""
    from intception_amd import onee_script
    from intception_amd.generator import generator, generator_options
    rng = np.random.default_rng(12)
    gen = generator(onee_script.make_overlap(), onee_script.make_nuclear(),
                    opt=generator_options(contracted=True, spherical_transformed=True))
    C = rng.uniform(0, 3, size=3)
    oracle.set_boys_mode(oracle.BOYS_EXACT)
    for la, lb in [(0, 0), (1, 2), (3, 1), (2, 3), (3, 3)]:
        shells = []
        for l in (la, lb):
            k = int(rng.integers(1, 4))
            e = 10 ** rng.uniform(-1, 1.5, size=k)
            c = np.stack([basis.normalize_contraction(e, rng.uniform(0.1, 1, k), l) for _ in range(2)], axis=1)
            shells.append({"l": l, "exps": e, "coefs": c, "cen": rng.uniform(0, 3, size=3), "sph": basis.SPH[l]})
        na, nb = basis.nsph(la), basis.nsph(lb)
        for name in gen.names():
            b = gen.binding(name)
            # out[ia * 1 + ka * na + ib * (2 na) + kb * (2 na nb)]
            out = np.zeros(4 * na * nb)
            b.main(shells, out, C=C, strides=[1, 2 * na], cont_strides=[na, 2 * na * nb])
            kind = 0 if b.kind == "overlap" else 1
            for ka in range(2):
                for kb in range(2):
                    s2 = [dict(s, coefs=s["coefs"][:, k]) for s, k in zip(shells, (ka, kb))]
                    ls = np.array([la, lb], dtype=np.int32)
                    nps = np.array([len(s2[0]["exps"]), len(s2[1]["exps"])], dtype=np.int32)

                    class Sh:
                        shell_l, shell_nprim = ls, nps
                        shell_off = np.array([0, nps[0]], dtype=np.int32)
                        shell_cen = np.stack([s2[0]["cen"], s2[1]["cen"]])
                        exps = np.concatenate([s2[0]["exps"], s2[1]["exps"]])
                        coefs = np.concatenate([s2[0]["coefs"], s2[1]["coefs"]])
                    it = np.array([[0, 1, 0]] if kind else [[0, 1]], dtype=np.int32)
                    ref = oracle.onee_batch(kind, True, Sh, it, [0], na * nb, C[None, :], basis.SPH)
                    got = np.array([out[ia + ka * na + ib * 2 * na + kb * 2 * na * nb]
                                    for ia in range(na) for ib in range(nb)])
                    _check_blocks(got, ref, [na * nb], [0], "%s (%d %d) cont %d %d" % (name, la, lb, ka, kb))
    gp = generator(onee_script.make_overlap(), onee_script.make_nuclear(),
                   opt=generator_options(contracted=False, spherical_transformed=False))
    A, B = rng.uniform(0, 3, size=(2, 3))
    for name in gp.names():
        b = gp.binding(name)
        kind = 0 if b.kind == "overlap" else 1
        out = np.zeros(basis.ncc(2) * basis.ncc(1))
        b.main([{"l": 2, "exps": 1.3, "cen": A}, {"l": 1, "exps": 0.4, "cen": B}], out, C=C)
        ref = oracle.onee_prim(kind, 2, 1, 1.3, A, 0.4, B, C)
        _check_blocks(out, ref, [len(ref)], [0], "primitive binding %s" % name)
