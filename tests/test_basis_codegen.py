"""CPU: caller-side data (normalisation, spherical matrices), the canonical
class mapping, the generated kernels' lane tables, the flop model, and the
C-ABI library's exported symbols."""
import itertools
import os
import re

import numpy as np

from intception_amd import basis, codegen, flops_model, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_contraction_normalised():
    rng = np.random.default_rng(0)
    for l in range(5):
        e = 10 ** rng.uniform(-1, 2, size=4)
        c = basis.normalize_contraction(e, rng.uniform(0.1, 1, size=4), l)
        p = e[:, None] + e[None, :]
        ov = basis._dfact(2 * l - 1) / (2 * p) ** l * (np.pi / p) ** 1.5
        assert abs(c @ ov @ c - 1.0) < 1e-13


def test_sph_orthonormal():
    for l in range(7):
        S = basis.SPH[l]
        G = S.T @ basis._cart_overlap(l) @ S
        assert np.abs(G - np.eye(2 * l + 1)).max() < 1e-13


def test_sph_d_known_values():
    S = basis.SPH[2]
    # m = 0 column: z^2 - (x^2 + y^2) / 2
    comps = basis.cart_components(2)
    col = dict(zip(comps, S[:, 2]))
    assert col[(0, 0, 2)] == 1.0 and col[(2, 0, 0)] == -0.5 and col[(0, 2, 0)] == -0.5


def test_canonical_classes():
    for cls in itertools.product(range(4), repeat=4):
        can, perm = codegen.canonical(*cls)
        assert sorted(perm) == [0, 1, 2, 3]
        assert tuple(cls[p] for p in perm) == can
        assert can[0] >= can[1] and can[2] >= can[3]
        assert (can[0] + can[1], can[0]) <= (can[2] + can[3], can[2])
        assert {perm[0], perm[1]} in ({0, 1}, {2, 3})
    assert len(codegen.canonical_classes(2)) == 21


def test_lane_chain_tables():
    # every lane's chain of directions rebuilds its own component
    for LX in range(7):
        for e in codegen.comps_upto(LX):
            t = [0, 0, 0]
            steps = []
            u = e
            while sum(u) > 0:
                i = codegen.first_nonzero(u)
                steps.append((i, u[i] - 1))
                u = codegen.dec(u, i)
            for i, k in reversed(steps):
                assert k == t[i]          # multiplier of the (t - 1_i) term
                t[i] += 1
            assert tuple(t) == e


def test_generated_kernel_is_well_formed():
    src, meta = codegen.emit_class_source((2, 1, 2, 2), True)
    # every register variable is defined before use
    defined = set()
    for line in src.splitlines():
        m = re.match(r"\s*const double (\w+) = (.*);", line)
        if m:
            for tok in re.findall(r"\b([vc]\d+_\d+|hk\d+_\d+|hb\d+_\d+)\b", m.group(2)):
                assert tok in defined, tok
            defined.add(m.group(1))
    assert meta["G"] == 20 and meta["NCD"] == 25


def test_slot_stride_avoids_bank_conflicts():
    # brute force over the wave's slot reads: the chosen stride is >= the
    # slot length, no worse than any stride up to 8 longer, and for 4-lane
    # groups never a multiple of 8 doubles (16 groups on the same banks)
    from intception_amd.codegen import _lds_conflict, slot_stride

    def worst(G, n, ipa):
        w = 0
        for st in range(G):
            for i in range(n):
                w = max(w, _lds_conflict([2 * ((g * G + st) * n + i) + h for g in range(64 // G) for h in (0, 1)]))
            w = max(w, _lds_conflict([2 * ((g * G + st) * n + ipa + k) + h
                                      for g in range(64 // G) for k in range(4) for h in (0, 1)]))
        return w
    for G in (4, 10, 20):
        for ns in range(18, 30):
            n = slot_stride(G, ns, 5)
            assert ns <= n <= ns + 8
            assert worst(G, n, 5) <= min(worst(G, m, 5) for m in range(ns, ns + 9))
            if G == 4:
                assert n % 8 != 0 and worst(G, n, 5) <= 2
    assert _lds_conflict([0, 64, 128, 1]) == 3


def test_float_slot_layout():
    # FP32 kernels: float slots of even length (the prefactor double at floats
    # 0-1 stays 8-byte aligned), no worse in bank conflicts than any even
    # length up to 16 longer; the generated step reads floats, converts none,
    # and the LDS size in doubles covers 64 float slots
    from intception_amd.codegen import _lds_conflict, slot_stride_f

    def worst(G, n, ns, ipa):
        w = 0
        for st in range(G):
            for i in range(ns):
                w = max(w, _lds_conflict([(g * G + st) * n + i for g in range(64 // G)]))
            w = max(w, _lds_conflict([(g * G + st) * n + ipa + k for g in range(64 // G) for k in range(4)]))
        return w
    for G in (4, 10, 20):
        for ns in range(18, 30):
            n = slot_stride_f(G, ns, 6)
            assert n % 2 == 0 and ns <= n <= ns + 17
            assert worst(G, n, ns, 6) <= min(worst(G, m, ns, 6) for m in range(ns + (ns & 1), ns + 18, 2))
    src, meta = codegen.emit_class_source((1, 1, 2, 1), True, "f")
    assert meta["NSF"] % 2 == 0 and meta["NS"] * 2 == meta["NSF"]
    step = src[src.index("const float *u = gslot + st * NSF;"):src.index("acc4 = fma(pref")]
    # the only FP32 slot reads with an index are the PA/WP rows (offsets 8 and 12) by chain direction
    assert "(float)u[" not in re.sub(r"\(float\)u\[(8|12) \+ (d\d|adir)\]", "", step)
    assert "const double pref = *(const double *)(u + 0);" in step
    src, meta = codegen.emit_class_source((1, 1, 2, 1), True, "d")
    assert meta["NSF"] == 0 and "const double *u = gslot + st * NS;" in src


def test_multi_source_depends_on_every_member():
    # the persistent kernel's source includes the class sources: a changed
    # body must rebuild it
    from intception_amd import build as bld
    metas = [dict(name="intc_eri_d0000_s", spherical=True), dict(name="intc_eri_f1121_s", spherical=True),
             dict(name="intc_eri_d0000_c", spherical=False)]
    files = [os.path.join(bld.GEN, n + ".hip") for n in ("intc_eri_d0000_s", "intc_eri_f1121_s", "intc_eri_d0000_c",
                                                          "intc_eri_multi")]
    deps = dict((os.path.basename(f), d) for f, d in bld.class_sources(files, metas))
    common = os.path.join(bld.CSRC, "intc_common.h")
    assert deps["intc_eri_d0000_s.hip"] == [common]
    assert deps["intc_eri_multi.hip"] == [common, files[0], files[1]]


def test_flops_model_small_classes():
    # (ps|ss): a-build of 3 components at m = 0, 2 terms each
    assert flops_model.vrr_terms(1, 0, 0, 0) == 3 * 2
    assert flops_model.vrr_terms(0, 0, 0, 0) == 0
    assert flops_model.quartet_flops((0, 0, 0, 0), (1, 1, 1, 1), False) == 40 + 22 + 1 + 2


def test_library_exports_every_declared_symbol():
    if not os.path.exists(lib.LIB_PATH):
        import pytest
        pytest.skip("libintc.so not built")
    hdr = open(os.path.join(ROOT, "include", "intc.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b(intc_\w+)\s*\(", hdr)) - {"intc_plan"}
    L = lib.load()
    for n in names:
        assert hasattr(L, n), n
    assert {s[0] for s in lib.SIGNATURES} == names
    # the network's entry points (include/intc_mlp.h)
    hdr = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "intc_mlp.h")).read(), flags=re.S)
    mlp_names = set(re.findall(r"\b(intc_mlp_\w+)\s*\(", hdr))
    assert len(mlp_names) == 7
    for n in mlp_names:
        assert hasattr(L, n), n
    # host-only entry points work without a GPU
    assert L.intc_eri_cont_cart_work_array_size(1, 1, 1, 1, 3, 3, 3, 3, 1, 1, 1, 1) == 25930
    assert L.intc_set_boys_mode(7) != 0
    assert b"boys mode" in L.intc_last_error()


def test_onee_plan_argument_checks():
    """intc_onee_plan_create refuses bad arguments before any device call
    (runs without a GPU): unknown kind, l beyond the built kernels, shell and
    centre indices out of range, negative sizes."""
    import ctypes
    import numpy as np
    if not os.path.exists(lib.LIB_PATH):
        import pytest
        pytest.skip("libintc.so not built")
    L = lib.load()
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    h = ctypes.c_void_p()
    sl = np.array([0, 1, 4], dtype=np.int32)
    snp = np.array([1, 2, 1], dtype=np.int32)
    soff = np.array([0, 1, 3], dtype=np.int32)
    oo = np.zeros(1, dtype=np.int64)
    llp = oo.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))

    def create(kind, nshell, items, ncentre=1):
        it = np.ascontiguousarray(items, dtype=np.int32)
        return L.intc_onee_plan_create(ctypes.byref(h), kind, 1, nshell, ip(sl), ip(snp), ip(soff), None, None, None,
                                       ncentre, 1, ip(it), llp)

    assert create(2, 2, [0, 1]) == -1
    assert b"kind" in L.intc_last_error()
    assert create(0, 3, [0, 2]) == -2          # shell 2 has l = 4
    assert create(0, 2, [0, 5]) == -1          # shell index out of range
    assert create(1, 2, [0, 1, 3], ncentre=2) == -1   # centre index out of range
    assert L.intc_onee_plan_create(ctypes.byref(h), 0, 1, -1, ip(sl), ip(snp), ip(soff), None, None, None, 0, 0,
                                   None, None) == -1


def test_pair_w_sums_matches_direct_loop():
    """runtime.pair_w_sums: the per-pair sum of the s-type Schwarz factor the
    device primitive screen cuts on, for shared-pair and per-quartet batches."""
    import math
    import numpy as np
    from intception_amd import basis, molecule, runtime
    m = molecule.Molecule(n_atoms=6, seed=3)
    m.pairs = m.prescreen_pairs()
    b = m.pair_batch(np.array([[0, 1], [2, 2], [1, 3]]))

    def direct(a, c):
        tot = 0.0
        for i in range(m.shell_nprim[a]):
            for j in range(m.shell_nprim[c]):
                al, be = m.exps[m.shell_off[a] + i], m.exps[m.shell_off[c] + j]
                z = al + be
                r2 = float(np.sum((m.shell_cen[a] - m.shell_cen[c]) ** 2))
                kap = math.sqrt(2 * math.pi ** 2.5) * math.exp(-al * be / z * r2) / z
                kap *= m.coefs[m.shell_off[a] + i] * m.coefs[m.shell_off[c] + j]
                tot += abs(kap) * (2 * z) ** -0.25
        return tot

    got = runtime.pair_w_sums(b)
    ref = np.array([direct(a, c) for a, c in m.pairs])
    assert np.allclose(got, ref, rtol=1e-13, atol=0)
    # per-quartet batch: the unique (sorted) bra and ket pairs
    q = np.array([[0, 1, 2, 3], [3, 2, 1, 0], [4, 4, 0, 1]], dtype=np.int32)
    bq = m.batch(q)
    pairs = sorted({(0, 1), (2, 3), (4, 4)})
    assert np.allclose(runtime.pair_w_sums(bq), [direct(a, c) for a, c in pairs], rtol=1e-13, atol=0)
