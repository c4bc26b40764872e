"""``generator`` / ``generator_options`` with a HIP backend.

Same entry points as the reference generator
(src/intception/generator/generator.py:102-524, 1352-1702):

    opt = generator_options(contracted=True, spherical_transformed=True,
                            use_cblas=False, global_user_lmax=4, backend="hip")
    gen = generator(eri, opt=opt)
    gen.out()

With ``backend="hip"`` the generator does not emit recurrence code from the
script's expression tree.  It checks, numerically at random exponents,
centres, target components and directions, that each integral is the
Obara-Saika / Head-Gordon-Pople Coulomb ERI family (base expression, every VRR
and the HRR) -- or, for 2-index integrals, the Obara-Saika overlap or nuclear
attraction of the reference's example script (``OneElectronBinding``) -- and
binds it to the hand-written gfx950 kernels of libintc:

* ``gen.binding(name)`` -> ``EriBinding`` with the per-quartet twin
  ``main(...)`` (same argument layout as the reference's ``<full>_main``),
  ``work_array_size(...)`` and ``plan(batch)`` for the batched path;
* ``gen.out()`` writes ``<full>.h`` to the output directory: the reference's
  function names (``<full>_main``, ``<full>_work_array_size``) mapped onto the
  libintc C-ABI (include/intc.h), plus a JSON manifest.

An integral that is not in the family raises ``NotImplementedError``.
"""
import json
import math
import os

import numpy as np

from . import dsl


class parallelism:
    def __init__(self, vectorized=False, vector_batch_size=4):
        self._vectorized, self._vbs = vectorized, vector_batch_size

    def vectorized(self):
        return self._vectorized


class generator_options:
    """Reference options (contracted, spherical_transformed, global_user_lmax,
    use_cblas, output_directory, suffixes, silent; generator.py:102-334) plus
    this build's: backend ("hip": bind to libintc's kernels; "c99": emit the
    C99 file set of the ERI class) and boys ("chebyshev" or "reference", the
    device Boys scheme).  use_cblas=True is refused: the reference's CBLAS
    contraction path does not run (an undefined name,
    generator_contract_function.py:264) and this build has no CBLAS path.
    Mixed precision and screening are properties of a run, not of the
    generated code: they come with the plan (EriPlan.run_routed with a
    screening.LearnedRouter, EriPlan.set_prim_screen)."""

    def __init__(self, contracted=None, spherical_transformed=None, global_user_lmax=None, use_cblas=None,
                 output_directory=None, auto_suffix=None, prim_suffix=None, cont_suffix=None, cart_suffix=None,
                 sph_suffix=None, silent=None, backend=None, boys=None,
                 **unused):
        self._contracted = True if contracted is None else bool(contracted)
        self._sph = True if spherical_transformed is None else bool(spherical_transformed)
        if self._sph and not self._contracted:
            raise ValueError("spherical transformation requires contracted integrals")
        self._lmax = 6 if global_user_lmax is None else int(global_user_lmax)
        if use_cblas:
            raise NotImplementedError("use_cblas=True: the reference's CBLAS contraction does not run "
                                      "(generator_contract_function.py:264) and this build has no CBLAS path")
        self._cblas = False
        self._outdir = output_directory or "./output"
        self._auto_suffix = True if auto_suffix is None else auto_suffix
        self._suffix = {"prim": prim_suffix or "prim", "cont": cont_suffix or "cont", "cart": cart_suffix or "cart",
                        "sph": sph_suffix or "sph"}
        self._silent = bool(silent)
        self._backend = backend or "hip"
        if self._backend not in ("hip", "c99"):
            raise ValueError("this build's generator has the 'hip' and 'c99' backends")
        self._boys = boys or "chebyshev"
        if self._boys not in ("chebyshev", "reference"):
            raise ValueError("boys must be 'chebyshev' or 'reference'")
        # the reference's remaining options are accepted; they do not change
        # this build's output (file names are fixed, no CBLAS, no
        # vectorisation placeholder); anything else is a mistake
        known = {"spherical_transform_conditionals", "pllm", "default_boys_f", "global_data_filename",
                 "global_functions_filename", "boys_function_filename", "intception_main_filename",
                 "cblas_header_filename", "cblas_dgemm_min_lengths", "add_suffix_to_global", "stdout_width"}
        bad = sorted(set(unused) - known)
        if bad:
            raise TypeError("generator_options: unknown option(s) %s" % ", ".join(bad))
        self._unused = unused

    def contracted(self): return self._contracted
    def spherical_transformed(self): return self._sph
    def global_user_lmax(self): return self._lmax
    def use_cblas(self): return self._cblas
    def output_directory(self): return self._outdir
    def backend(self): return self._backend
    def boys(self): return self._boys
    def parallelism(self): return parallelism()

    def suffix(self):
        if not self._auto_suffix:
            return ""
        s = "_" + self._suffix["cont" if self._contracted else "prim"]
        return s + "_" + self._suffix["sph" if self._sph else "cart"]


# ---------------------------------------------------------------------------
# numeric recognition of the OS/HGP Coulomb ERI family
# ---------------------------------------------------------------------------
def _boys_exact(n, x):
    if x < 30.0:
        term = 1.0 / (2 * n + 1)
        s = term
        for k in range(1, 400):
            term *= 2.0 * x / (2 * n + 2 * k + 1)
            s += term
            if term < 1e-18 * s:
                break
        return math.exp(-x) * s
    f = 0.5 * math.sqrt(math.pi / x) * math.erf(math.sqrt(x))
    for m in range(n):
        f = ((2 * m + 1) * f - math.exp(-x)) / (2 * x)
    return f


def _ctx(rng, names):
    ex = {n: 10 ** rng.uniform(-0.5, 0.8) for n in names}
    cen = {n: rng.uniform(-1.5, 1.5, size=3) for n in names}
    v = {}
    for n in names:
        v["x" + n] = ex[n]
        v["c" + n] = cen[n]
    return {"vars": v, "dir": 0, "index": {}, "aux": {}, "boys": _boys_exact}, ex, cen


def _expected_base(ex, cen, names, m):
    a, b, c, d = (ex[n] for n in names)
    A, B, C, D = (cen[n] for n in names)
    z, e = a + b, c + d
    P, Q = (a * A + b * B) / z, (c * C + d * D) / e
    T = z * e / (z + e) * float(np.sum((P - Q) ** 2))
    return (2 * math.pi ** 2.5 / (z * e * math.sqrt(z + e)) * math.exp(-a * b / z * float(np.sum((A - B) ** 2)))
            * math.exp(-c * d / e * float(np.sum((C - D) ** 2))) * _boys_exact(m, T))


def _expected_vrr(ex, cen, names, aux, X, target, i):
    """OS VRR for incrementing index X (position in names) to `target`
    (dict name -> (x,y,z)) in direction i: {shift key: coefficient}."""
    a, b, c, d = (ex[n] for n in names)
    A, B, C, D = (cen[n] for n in names)
    z, e = a + b, c + d
    P, Q = (a * A + b * B) / z, (c * C + d * D) / e
    W = (z * P + e * Q) / (z + e)
    rho = z * e / (z + e)
    bra = X in (0, 1)
    ctr = [A, B, C, D][X]
    own = [0, 1] if bra else [2, 3]
    other = [2, 3] if bra else [0, 1]
    partner = own[1] if X == own[0] else own[0]
    zz = z if bra else e
    center = P if bra else Q
    out = {}

    def key(dec):
        sh = [(n, -dec.get(k, 0)) for k, n in enumerate(names)]
        return tuple(sh) + ((aux, dec.get("m", 0)),)

    def add(dec, v):
        if v != 0.0:
            k = key(dec)
            out[k] = out.get(k, 0.0) + v

    add({X: 1}, (center - ctr)[i])
    add({X: 1, "m": 1}, (W - center)[i])
    nx = target[names[X]][i] - 1
    if nx > 0:
        add({X: 2}, nx / (2 * zz))
        add({X: 2, "m": 1}, -nx / (2 * zz) * rho / zz)
    npart = target[names[partner]][i]
    if npart > 0:
        add({X: 1, partner: 1}, npart / (2 * zz))
        add({X: 1, partner: 1, "m": 1}, -npart / (2 * zz) * rho / zz)
    for o in other:
        no = target[names[o]][i]
        if no > 0:
            add({X: 1, o: 1, "m": 1}, no / (2 * (z + e)))
    return out


def _lin_clean(d):
    return {k: v for k, v in d.items() if k is not None and v != 0.0}


def recognize_eri(integral, trials=24, seed=7):
    """Raise NotImplementedError unless `integral` is the OS/HGP Coulomb ERI
    (4 Cartesian indices a,b | c,d, one auxiliary index)."""
    idx, aux = integral.index_list(), integral.aux_list()
    if len(idx) != 4 or len(aux) != 1 or integral.arg_list():
        raise NotImplementedError("hip backend: integral %s is not a 4-index Coulomb ERI with one auxiliary index"
                                  % integral.name())
    names = [i.name() for i in idx]
    mname = aux[0].name()
    rng = np.random.default_rng(seed)
    base = integral.base()
    if base is None:
        raise NotImplementedError("no base expression")
    for _ in range(trials // 4):
        ctx, ex, cen = _ctx(rng, names)
        for m in range(4):
            ctx["aux"] = {mname: m}
            got = base.ev(ctx)
            ref = _expected_base(ex, cen, names, m)
            if not abs(got - ref) <= 1e-11 * abs(ref) + 1e-300:
                raise NotImplementedError("base expression of %s is not the Coulomb ERI base" % integral.name())
    vrrs = [r for r in integral.rr_list() if r.rrtype() == "vrr"]
    hrrs = [r for r in integral.rr_list() if r.rrtype() == "hrr"]
    if not vrrs:
        raise NotImplementedError("no VRR")
    for r in vrrs:
        X = names.index(r.changing_index().name())
        for _ in range(trials):
            ctx, ex, cen = _ctx(rng, names)
            target = {n: tuple(int(v) for v in rng.integers(0, 3, size=3)) for n in names}
            i = int(rng.integers(0, 3))
            if target[names[X]][i] == 0:
                t = list(target[names[X]]); t[i] = 1; target[names[X]] = tuple(t)
            ctx.update(dir=i, index=target, aux={mname: int(rng.integers(0, 3))})
            got = _lin_clean(r.expr().lin(ctx))
            ref = _expected_vrr(ex, cen, names, mname, X, target, i)
            keys = set(got) | set(ref)
            for k in keys:
                g, f = got.get(k, 0.0), ref.get(k, 0.0)
                if not abs(g - f) <= 1e-10 * max(abs(f), 1e-3):
                    raise NotImplementedError("VRR %s on %s is not the Obara-Saika ERI recurrence (term %s: %r vs %r)"
                                              % (r.name(), names[X], k, g, f))
    built = {r.changing_index().name() for r in vrrs}
    for r in hrrs:
        to, frm = r.changing_index().name(), r.move_from().name()
        if {to, frm} not in ({names[0], names[1]}, {names[2], names[3]}):
            raise NotImplementedError("HRR must move angular momentum within one electron's pair")
        built.add(to)
    if built != set(names):
        raise NotImplementedError("every index must be built by a VRR or the HRR (built: %s)" % sorted(built))
    return True


# ---------------------------------------------------------------------------
# numeric recognition of the 2-index overlap and nuclear attraction families
# (the reference's example script, examples/example_input.py:79-110)
# ---------------------------------------------------------------------------
def _expected_onee_base(kind, ex, cen, names, m, C):
    a, b = (ex[n] for n in names)
    A, B = (cen[n] for n in names)
    z = a + b
    P = (a * A + b * B) / z
    k = math.exp(-a * b / z * float(np.sum((A - B) ** 2)))
    if kind == "overlap":
        return (math.pi / z) ** 1.5 * k
    return 2 * math.pi / z * k * _boys_exact(m, z * float(np.sum((P - C) ** 2)))


def _expected_onee_vrr(kind, ex, cen, names, aux, X, target, i, C):
    """OS VRR of the overlap / nuclear attraction incrementing index X (0 or
    1) to `target` in direction i: {shift key: coefficient}."""
    a, b = (ex[n] for n in names)
    A, B = (cen[n] for n in names)
    z = a + b
    P = (a * A + b * B) / z
    partner = 1 - X
    nuc = kind == "nuclear"
    out = {}

    def key(dec):
        sh = tuple((n, -dec.get(k, 0)) for k, n in enumerate(names))
        return sh + (((aux, dec.get("m", 0)),) if nuc else ())

    def add(dec, v):
        if v != 0.0:
            k = key(dec)
            out[k] = out.get(k, 0.0) + v

    add({X: 1}, (P - [A, B][X])[i])
    if nuc:
        add({X: 1, "m": 1}, -(P - C)[i])
    nx = target[names[X]][i] - 1
    if nx > 0:
        add({X: 2}, nx / (2 * z))
        if nuc:
            add({X: 2, "m": 1}, -nx / (2 * z))
    npart = target[names[partner]][i]
    if npart > 0:
        add({X: 1, partner: 1}, npart / (2 * z))
        if nuc:
            add({X: 1, partner: 1, "m": 1}, -npart / (2 * z))
    return out


def recognize_onee(integral, trials=24, seed=11):
    """'overlap' or 'nuclear' when `integral` is the Obara-Saika 2-index
    overlap (two Cartesian indices) or nuclear attraction (two indices, one
    auxiliary index, one position argument: the point charge); raises
    NotImplementedError otherwise."""
    idx, aux, args = integral.index_list(), integral.aux_list(), integral.arg_list()
    if len(idx) != 2:
        raise NotImplementedError("hip backend: %s is not a 2-index integral" % integral.name())
    if not aux and not args:
        kind = "overlap"
    elif len(aux) == 1 and len(args) == 1 and isinstance(args[0], dsl.dsl_position):
        kind = "nuclear"
    else:
        raise NotImplementedError("hip backend: 2-index integral %s is neither the overlap nor the nuclear "
                                  "attraction" % integral.name())
    names = [i.name() for i in idx]
    mname = aux[0].name() if aux else None
    cname = args[0].name() if args else None
    rng = np.random.default_rng(seed)
    base = integral.base()
    if base is None:
        raise NotImplementedError("no base expression")

    def ctx_c():
        ctx, ex, cen = _ctx(rng, names)
        C = rng.uniform(-1.5, 1.5, size=3)
        if cname:
            ctx["vars"][cname] = C
        return ctx, ex, cen, C

    for _ in range(trials // 4):
        ctx, ex, cen, C = ctx_c()
        for m in range(4 if kind == "nuclear" else 1):
            ctx["aux"] = {mname: m} if mname else {}
            got = base.ev(ctx)
            ref = _expected_onee_base(kind, ex, cen, names, m, C)
            if not abs(got - ref) <= 1e-11 * abs(ref) + 1e-300:
                raise NotImplementedError("base expression of %s is not the %s base" % (integral.name(), kind))
    vrrs = [r for r in integral.rr_list() if r.rrtype() == "vrr"]
    hrrs = [r for r in integral.rr_list() if r.rrtype() == "hrr"]
    if not vrrs:
        raise NotImplementedError("no VRR")
    for r in vrrs:
        X = names.index(r.changing_index().name())
        for _ in range(trials):
            ctx, ex, cen, C = ctx_c()
            target = {n: tuple(int(v) for v in rng.integers(0, 3, size=3)) for n in names}
            i = int(rng.integers(0, 3))
            if target[names[X]][i] == 0:
                t = list(target[names[X]]); t[i] = 1; target[names[X]] = tuple(t)
            ctx.update(dir=i, index=target, aux={mname: int(rng.integers(0, 3))} if mname else {})
            got = _lin_clean(r.expr().lin(ctx))
            ref = _expected_onee_vrr(kind, ex, cen, names, mname, X, target, i, C)
            for k in set(got) | set(ref):
                g, f = got.get(k, 0.0), ref.get(k, 0.0)
                if not abs(g - f) <= 1e-10 * max(abs(f), 1e-3):
                    raise NotImplementedError("VRR %s on %s is not the Obara-Saika %s recurrence (term %s: %r vs %r)"
                                              % (r.name(), names[X], kind, k, g, f))
    built = {r.changing_index().name() for r in vrrs}
    for r in hrrs:
        built.add(r.changing_index().name())
    if built != set(names):
        raise NotImplementedError("every index must be built by a VRR or the HRR (built: %s)" % sorted(built))
    return kind


# ---------------------------------------------------------------------------
# bindings
# ---------------------------------------------------------------------------
class EriBinding:
    """One generated class bound to libintc."""

    def __init__(self, integral, options):
        self.integral = integral
        self.options = options
        self.full_name = integral.full_name() + options.suffix()
        self.spherical = options.spherical_transformed()

    def _lib(self):
        from . import lib
        L = lib.load()
        lib.check(L.intc_set_boys_mode(1 if self.options.boys() == "reference" else 0))
        return L

    def work_array_size(self, la, lb, lc, ld, na_prim, nb_prim, nc_prim, nd_prim, na_cont=1, nb_cont=1,
                        nc_cont=1, nd_cont=1):
        from . import lib
        f = lib.load().intc_eri_cont_sph_work_array_size if self.spherical else \
            lib.load().intc_eri_cont_cart_work_array_size
        return f(la, lb, lc, ld, na_prim, nb_prim, nc_prim, nd_prim, na_cont, nb_cont, nc_cont, nd_cont)

    def main(self, shells, output_array, work_array=None, strides=None, cont_strides=None):
        """Per-quartet twin of <full>_main.  shells: 4 dicts with exps,
        cen, l, coefs (n_prim x n_cont), and 'sph' (ncc x nsph) when spherical."""
        import ctypes
        from . import lib
        from .basis import ncc, nsph
        L = self._lib()
        dp = ctypes.POINTER(ctypes.c_double)
        ns = [nsph(s["l"]) if self.spherical else ncc(s["l"]) for s in shells]
        if strides is None:
            strides = [ns[1] * ns[2] * ns[3], ns[2] * ns[3], ns[3], 1]
        cont_strides = cont_strides or [0, 0, 0, 0]
        keep, args = [], []
        for k, s in enumerate(shells):
            x = np.ascontiguousarray(s["exps"], dtype=np.float64)
            c = np.asfortranarray(np.asarray(s["coefs"], dtype=np.float64).reshape(len(x), -1))
            cen = np.ascontiguousarray(s["cen"], dtype=np.float64)
            keep += [x, c, cen]
            args += [x.ctypes.data_as(dp), cen.ctypes.data_as(dp), s["l"], ns[k], int(strides[k]), len(x),
                     c.shape[1], int(cont_strides[k]), c.ctypes.data_as(dp)]
            if self.spherical:
                S = np.asfortranarray(s["sph"], dtype=np.float64)
                keep.append(S)
                args.append(S.ctypes.data_as(dp))
        work = np.zeros(1) if work_array is None else work_array
        out = output_array
        fn = L.intc_eri_cont_sph_main if self.spherical else L.intc_eri_cont_cart_main
        lib.check(fn(*args, work.ctypes.data_as(dp), out.ctypes.data_as(dp)))
        return out

    def plan(self, batch, device="cuda"):
        from .runtime import EriPlan
        assert batch.spherical == self.spherical
        self._lib()
        return EriPlan(batch, device)

    def header(self):
        n = self.full_name
        kind = "sph" if self.spherical else "cart"
        return "\n".join([
            "/* Generated by intception_amd.generator (hip backend) for integral '%s'. */" % self.integral.name(),
            "#ifndef %s_H" % n.upper(), "#define %s_H" % n.upper(), '#include "intc.h"',
            "/* Reference names mapped onto the libintc C-ABI; the functions return an",
            " * int status (0 = success) where the reference's return void. */",
            "#define %s_main intc_eri_cont_%s_main" % (n, kind),
            "#define %s_work_array_size intc_eri_cont_%s_work_array_size" % (n, kind),
            "#endif", ""])


class OneElectronBinding:
    """A 2-index overlap or nuclear-attraction integral bound to libintc's
    one-electron kernels (shells up to l = 3)."""

    def __init__(self, integral, kind, options):
        self.integral = integral
        self.kind = kind
        self.options = options
        self.full_name = integral.full_name() + options.suffix()
        self.spherical = options.spherical_transformed()
        self.contracted = options.contracted()

    def _fn(self):
        from . import lib
        L = lib.load()
        lib.check(L.intc_set_boys_mode(1 if self.options.boys() == "reference" else 0))
        pre = "intc_overlap_" if self.kind == "overlap" else "intc_nuclear_"
        if not self.contracted:
            return getattr(L, pre + "prim_cart_main")
        return getattr(L, pre + ("cont_sph_main" if self.spherical else "cont_cart_main"))

    def main(self, shells, output_array, C=None, strides=None, cont_strides=None):
        """Per-call twin of <full>_main.  Primitive (contracted=False): shells
        are 2 dicts with a scalar 'exps', 'cen', 'l'; contracted: as for the
        ERI binding (exps, coefs n_prim x n_cont, cen, l, sph when spherical).
        C: the point charge (nuclear attraction)."""
        import ctypes
        from . import lib
        from .basis import ncc, nsph
        fn = self._fn()
        dp = ctypes.POINTER(ctypes.c_double)
        ns = [nsph(s["l"]) if self.spherical else ncc(s["l"]) for s in shells]
        strides = strides or [ns[1], 1]
        cont_strides = cont_strides or [0, 0]
        keep, args = [], []
        for k, s in enumerate(shells):
            cen = np.ascontiguousarray(s["cen"], dtype=np.float64)
            keep.append(cen)
            if not self.contracted:
                args += [float(np.asarray(s["exps"]).reshape(-1)[0]), cen.ctypes.data_as(dp), s["l"], ns[k],
                         int(strides[k])]
                continue
            x = np.ascontiguousarray(s["exps"], dtype=np.float64)
            c = np.asfortranarray(np.asarray(s["coefs"], dtype=np.float64).reshape(len(x), -1))
            keep += [x, c]
            args += [x.ctypes.data_as(dp), cen.ctypes.data_as(dp), s["l"], ns[k], int(strides[k]), len(x),
                     c.shape[1], int(cont_strides[k]), c.ctypes.data_as(dp)]
            if self.spherical:
                S = np.asfortranarray(s["sph"], dtype=np.float64)
                keep.append(S)
                args.append(S.ctypes.data_as(dp))
        if self.kind == "nuclear":
            cc = np.ascontiguousarray(C, dtype=np.float64)
            keep.append(cc)
            args.append(cc.ctypes.data_as(dp))
        work = np.zeros(1)
        lib.check(fn(*args, work.ctypes.data_as(dp), output_array.ctypes.data_as(dp)))
        return output_array

    def plan(self, shells, items, centres=None, device="cuda"):
        from .runtime import OneEPlan
        self._fn()
        return OneEPlan(0 if self.kind == "overlap" else 1, shells, items, centres=centres,
                        spherical=self.spherical, device=device)

    def header(self):
        n = self.full_name
        pre = "intc_overlap_" if self.kind == "overlap" else "intc_nuclear_"
        fn = pre + ("prim_cart_main" if not self.contracted else
                    "cont_sph_main" if self.spherical else "cont_cart_main")
        return "\n".join([
            "/* Generated by intception_amd.generator (hip backend) for integral '%s'. */" % self.integral.name(),
            "#ifndef %s_H" % n.upper(), "#define %s_H" % n.upper(), '#include "intc.h"',
            "/* Reference name mapped onto the libintc C-ABI; the function returns an",
            " * int status (0 = success) where the reference's returns void. */",
            "#define %s_main %s" % (n, fn), "#endif", ""])


class generator:
    def __init__(self, *integrals, opt=None):
        self._opt = opt or generator_options()
        self._bindings = {}
        for itg in integrals:
            if not isinstance(itg, dsl.dsl_integral):
                raise TypeError("generator takes dsl_integral objects")
            if len(itg.index_list()) == 2:
                b = OneElectronBinding(itg, recognize_onee(itg), self._opt)
            else:
                if not self._opt.contracted():
                    raise NotImplementedError("hip backend generates contracted ERIs")
                recognize_eri(itg)
                b = EriBinding(itg, self._opt)
            self._bindings[b.full_name] = b

    def generator_options(self):
        return self._opt

    def names(self):
        return list(self._bindings)

    def binding(self, name=None):
        if name is None:
            assert len(self._bindings) == 1
            return next(iter(self._bindings.values()))
        return self._bindings[name]

    def out(self):
        d = self._opt.output_directory()
        os.makedirs(d, exist_ok=True)
        if self._opt.backend() == "c99":
            # CPU C99 of each ERI class in the reference's file set (global_data,
            # boys_function, global_functions, <full> and its class files, a
            # Makefile for libintception.a), reference interface
            from . import c99gen
            written = {}
            for n, b in self._bindings.items():
                if not isinstance(b, EriBinding):
                    raise NotImplementedError("the c99 backend emits the ERI class only")
                written[n] = c99gen.write(d, n, b.spherical, self._opt.global_user_lmax())
            return written
        manifest = {}
        for n, b in self._bindings.items():
            with open(os.path.join(d, n + ".h"), "w") as fh:
                fh.write(b.header())
            manifest[n] = {"integral": b.integral.name(), "kind": getattr(b, "kind", "eri"),
                           "spherical": b.spherical, "backend": "hip",
                           "boys": self._opt.boys(),
                           "library": "libintc.so", "header": "intc.h"}
        with open(os.path.join(d, "intception_hip_manifest.json"), "w") as fh:
            json.dump(manifest, fh, indent=1)
        return manifest
