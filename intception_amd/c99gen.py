"""C99 backend of the generator: CPU source for the contracted ERI class.

``generator_options(backend="c99")`` makes ``gen.out()`` write ``<full>.c`` and
``<full>.h`` next to each other: a self-contained C99 implementation of the
reference's generated interface for the ERI input script
(src/intception/generator/callables/main_source_algorithms/contracted_general.py:31-73):

    void <full>_main(per index: const double *xX, const double cX[3], const int lX,
                     const int nX, const int iskipX, const int nX_prim,
                     const int nX_cont, const int iskipX_cont,
                     const double *contract_array_X [, const double *sph_trans_array_X],
                     double *work_array, double *output_array);
    int <full>_work_array_size(la, lb, lc, ld, na_prim..nd_prim, na_cont..nd_cont);

The code is generated from the recurrence structure of the script, the same
one the HIP kernels are generated from, and is serial straight-line C per
canonical class (la >= lb, lc >= ld, smaller pair in the bra):

* per primitive quartet: F_m(T) (Chebyshev interpolation per unit interval
  and downward recursion, the device's mode-0 scheme, table embedded), the
  Obara-Saika VRR (e|0)^(m) for |e| <= la + lb, then (e|f)^(m) for
  |f| <= lc + ld with the direction rule "first non-zero component"
  (src/intception/generator/generator.py:547-558), and the contraction sums;
* on the contracted values: the HGP HRR on the ket (c, d) and on the bra (a, b);
* the dispatcher permutes shells into the canonical class, applies the
  caller's sph_trans_array matrices per index (spherical) and scatters with
  the caller's iskip / iskip_cont strides (contracted_general.py:1808-1875).

The work array is accepted and not used (the generated code keeps its
intermediates in locals).
"""
import os

from .basis import cart_components, ncc, ncc0
from .codegen import (_hrr_lines, canonical, canonical_classes, cidx, comps_upto, dec, first_nonzero)

TWO_PI52 = 34.98683665524972  # 2 pi^(5/2)


def _prune(lines, roots):
    """Keep the `const double NAME = ...;` lines that the names in `roots`
    depend on (the HRR helpers emit every intermediate of their levels)."""
    import re
    need = set(roots)
    keep = []
    for ln in reversed(lines):
        m = re.match(r"\s*const double (\w+) = (.*);", ln)
        if m is None or m.group(1) in need:
            keep.append(ln)
            if m is not None:
                need.update(re.findall(r"[A-Za-z_]\w*", m.group(2)))
    return keep[::-1]


def _class_fn(cls, prefix="eri"):
    return "%s_c_%d%d%d%d" % ((prefix,) + tuple(cls))


CLASS_ARGS = ("const double *xa, const double *ka, int na, const double *A, const double *xb, const double *kb, "
              "int nb, const double *B, const double *xc, const double *kc, int nc, const double *C, "
              "const double *xd, const double *kd, int nd, const double *D, double *out")


def emit_class(cls, prefix="eri"):
    """Serial C for one canonical class: contracted Cartesian block
    out[a][b][c][d] of shells (A, B, C, D) in canonical order."""
    LA, LB, LC, LD = cls
    LX, LR = LA + LB, LC + LD
    M = LX + LR
    L = []
    o = L.append
    o("void %s(const double *xa, const double *ka, int na, const double *A," % _class_fn(cls, prefix))
    o("    const double *xb, const double *kb, int nb, const double *B,")
    o("    const double *xc, const double *kc, int nc, const double *C,")
    o("    const double *xd, const double *kd, int nd, const double *D, double *out)")
    o("{")
    o("    const double ABx = A[0] - B[0], ABy = A[1] - B[1], ABz = A[2] - B[2];")
    o("    const double CDx = C[0] - D[0], CDy = C[1] - D[1], CDz = C[2] - D[2];")
    o("    const double RAB2 = ABx * ABx + ABy * ABy + ABz * ABz, RCD2 = CDx * CDx + CDy * CDy + CDz * CDz;")
    o("    (void)ABx; (void)ABy; (void)ABz; (void)CDx; (void)CDy; (void)CDz;")
    acc_e = [e for e in comps_upto(LX) if sum(e) >= LA]
    acc_f = [f for f in comps_upto(LR) if sum(f) >= LC]
    for e in acc_e:
        o("    double " + ", ".join("a_%d_%d = 0.0" % (cidx(e), cidx(f)) for f in acc_f) + ";")
    o("    for (int i = 0; i < na; ++i) for (int j = 0; j < nb; ++j) {")
    o("        const double zeta = xa[i] + xb[j], oz = 1.0 / zeta;")
    o("        const double Px = (xa[i] * A[0] + xb[j] * B[0]) * oz, Py = (xa[i] * A[1] + xb[j] * B[1]) * oz,"
      " Pz = (xa[i] * A[2] + xb[j] * B[2]) * oz;")
    o("        const double kab = ka[i] * kb[j] * exp(-xa[i] * xb[j] * oz * RAB2);")
    o("        for (int k = 0; k < nc; ++k) for (int l = 0; l < nd; ++l) {")
    o("            const double eta = xc[k] + xd[l], oe = 1.0 / eta;")
    o("            const double Qx = (xc[k] * C[0] + xd[l] * D[0]) * oe, Qy = (xc[k] * C[1] + xd[l] * D[1]) * oe,"
      " Qz = (xc[k] * C[2] + xd[l] * D[2]) * oe;")
    o("            const double kcd = kc[k] * kd[l] * exp(-xc[k] * xd[l] * oe * RCD2);")
    o("            const double os = 1.0 / (zeta + eta), rho = zeta * eta * os;")
    o("            const double PQx = Px - Qx, PQy = Py - Qy, PQz = Pz - Qz;")
    o("            double F[%d];" % (M + 1))
    o("            intception_boys(F, 1, 0, %d, rho * (PQx * PQx + PQy * PQy + PQz * PQz));" % M)
    o("            const double pref = %r * oz * oe * sqrt(os) * kab * kcd;" % TWO_PI52)
    o("            const double PAx = Px - A[0], PAy = Py - A[1], PAz = Pz - A[2];")
    o("            const double WPx = -eta * os * PQx, WPy = -eta * os * PQy, WPz = -eta * os * PQz;")
    o("            const double QCx = Qx - C[0], QCy = Qy - C[1], QCz = Qz - C[2];")
    o("            const double WQx = zeta * os * PQx, WQy = zeta * os * PQy, WQz = zeta * os * PQz;")
    o("            const double o2z = 0.5 * oz, o2e = 0.5 * oe, o2s = 0.5 * os, roz = rho * oz, roe = rho * oe;")
    o("            (void)PAx; (void)PAy; (void)PAz; (void)WPx; (void)WPy; (void)WPz; (void)QCx; (void)QCy;"
      " (void)QCz; (void)WQx; (void)WQy; (void)WQz; (void)o2z; (void)o2e; (void)o2s; (void)roz; (void)roe;")

    def bname(e, m):
        return "b_%d_%d" % (cidx(e), m)

    def kname(e, f, m):
        return bname(e, m) if sum(f) == 0 else "k_%d_%d_%d" % (cidx(e), cidx(f), m)

    # the recurrence terms of every (e|f)^(m) and (e|0)^(m), then only the
    # values the accumulated targets depend on, in dependency order
    bra_all = comps_upto(LX)

    def ket_terms(e, f, m):
        i = first_nonzero(f)
        f1 = dec(f, i)
        ax = "xyz"[i]
        terms = [("QC" + ax, (e, f1, m)), ("WQ" + ax, (e, f1, m + 1))]
        if f1[i] > 0:
            f2 = dec(f1, i)
            terms += [("%d.0 * o2e" % f1[i], (e, f2, m)), ("-%d.0 * o2e * roe" % f1[i], (e, f2, m + 1))]
        if e[i] > 0:
            terms.append(("%d.0 * o2s" % e[i], (dec(e, i), f1, m + 1)))
        return terms

    def bra_terms(e, m):
        i = first_nonzero(e)
        e1 = dec(e, i)
        ax = "xyz"[i]
        terms = [("PA" + ax, (e1, (0, 0, 0), m)), ("WP" + ax, (e1, (0, 0, 0), m + 1))]
        if e1[i] > 0:
            e2 = dec(e1, i)
            terms += [("%d.0 * o2z" % e1[i], (e2, (0, 0, 0), m)), ("-%d.0 * o2z * roz" % e1[i], (e2, (0, 0, 0), m + 1))]
        return terms

    need = set()
    stack = [(e, f, 0) for e in acc_e for f in acc_f]
    while stack:
        key = stack.pop()
        if key in need:
            continue
        need.add(key)
        e, f, m = key
        if sum(f) > 0:
            stack += [k for _, k in ket_terms(e, f, m)]
        elif sum(e) > 0:
            stack += [k for _, k in bra_terms(e, m)]

    def emit_value(e, f, m, terms):
        ex = " + ".join("%s * %s" % (c, kname(*k)) for c, k in terms)
        o("            const double %s = %s;" % (kname(e, f, m), ex.replace("+ -", "- ")))

    for m in range(M + 1):
        if ((0, 0, 0), (0, 0, 0), m) in need:
            o("            const double %s = pref * F[%d];" % (bname((0, 0, 0), m), m))
    for n in range(1, LX + 1):
        for e in cart_components(n):
            for m in range(M - n + 1):
                if (e, (0, 0, 0), m) in need:
                    emit_value(e, (0, 0, 0), m, bra_terms(e, m))
    for n in range(1, LR + 1):
        for f in cart_components(n):
            for e in bra_all:
                for m in range(LR - n + 1):
                    if (e, f, m) in need:
                        emit_value(e, f, m, ket_terms(e, f, m))
    for e in acc_e:
        for f in acc_f:
            o("            a_%d_%d += %s;" % (cidx(e), cidx(f), kname(e, f, 0)))
    o("        }")
    o("    }")
    # ket HRR per bra row e, then bra HRR per ket column (c, d)
    cc, dc = cart_components(LC), cart_components(LD)
    ac, bc = cart_components(LA), cart_components(LB)
    ket = {}
    hrr = []
    for e in acc_e:
        hl, H = _hrr_lines(lambda f, e=e: "a_%d_%d" % (cidx(e), cidx(f)), LC, LD, "CD", "hk%d_" % cidx(e), "ket")
        hrr += hl
        for c in cc:
            for d in dc:
                ket[(e, c, d)] = H(c, d)
    nB, nC, nD = ncc(LB), ncc(LC), ncc(LD)
    outs = []
    for ci, c in enumerate(cc):
        for di, d in enumerate(dc):
            hl, H = _hrr_lines(lambda e, c=c, d=d: ket[(e, c, d)], LA, LB, "AB", "hb%d_%d_" % (ci, di), "bra")
            hrr += hl
            for ai, a in enumerate(ac):
                for bi, b in enumerate(bc):
                    outs.append((((ai * nB + bi) * nC + ci) * nD + di, H(a, b)))
    for ln in _prune(hrr, [nm for _, nm in outs]):
        o("    " + ln)
    for k, nm in outs:
        o("    out[%d] = %s;" % (k, nm))
    o("}")
    return "\n".join(L) + "\n"


def _cheb_rows(mmax):
    """Chebyshev rows for m <= mmax: read from the build's generated table
    (csrc/boys_tables.inc, the numbers the device uses), computed with
    mpmath when that file is absent."""
    import re
    from . import boys_tables as bt
    inc = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "boys_tables.inc")
    n = (mmax + 1) * bt.NINT
    if os.path.exists(inc):
        txt = open(inc).read()
        body = txt[txt.index("intc_boys_cheb_host"):]
        body = body[body.index("{") + 1:body.index("}")]
        vals = [float(v) for v in re.split(r"[,\s]+", body.strip()) if v]
        return [vals[i * (bt.DEG + 1):(i + 1) * (bt.DEG + 1)] for i in range(n)]
    return [bt.cheb_interval_poly(m, k) for m in range(mmax + 1) for k in range(bt.NINT)]


def _boys_files(mmax):
    """boys_function.{h,c}: intception_boys with the reference's interface
    (src/intception/generator/callables/generator_boys_function.py:98-205:
    f[m * ifskip] = F_m(x) for nmin <= m <= nmax); the values come from this
    build's scheme, a degree-DEG polynomial per unit interval of [0, NINT) for
    F_nmax and downward recursion, the asymptotic F_0 and upward recursion
    beyond (coefficients from F_n(x) = gamma(n + 1/2, x) / (2 x^(n + 1/2)),
    40-digit arithmetic)."""
    from . import boys_tables as bt
    rows = _cheb_rows(mmax)
    h = "\n".join([
        "/* Generated by intception_amd (c99 backend): Boys function F_m(x), m <= %d. */" % mmax,
        "#ifndef INTCEPTION_BOYS_FUNCTION_H", "#define INTCEPTION_BOYS_FUNCTION_H",
        "/* f[m * ifskip] = F_m(x) for nmin <= m <= nmax (nmax <= %d) */" % mmax,
        "void intception_boys(double *f, const int ifskip, const int nmin, const int nmax, const double x);",
        "#endif", ""])
    L = ["/* Generated by intception_amd (c99 backend) -- do not edit.",
         " * F_m(x) for m <= %d: degree-%d polynomial per unit interval of [0, %d) for F_nmax," % (mmax, bt.DEG,
                                                                                               bt.NINT),
         " * downward recursion; beyond, the asymptotic F_0 and upward recursion.  Coefficients from",
         " * F_n(x) = gamma(n + 1/2, x) / (2 x^(n + 1/2)) in 40-digit arithmetic. */",
         "#include <math.h>", '#include "boys_function.h"', "",
         "static const double boys_cheb[%d] = {" % (len(rows) * (bt.DEG + 1))]
    for r in rows:
        L.append("    " + ", ".join(repr(float(v)) for v in r) + ",")
    L += ["};", "",
          "void intception_boys(double *f, const int ifskip, const int nmin, const int nmax, const double x)",
          "{",
          "    double F[%d];" % (mmax + 1),
          "    double T = x > 0.0 ? x : 0.0;",
          "    int m;",
          "    if (T < %d.0) {" % bt.NINT,
          "        const int k = (int)T;",
          "        const double t = 2.0 * (T - (double)k) - 1.0, ex = exp(-T);",
          "        const double *c = boys_cheb + (nmax * %d + k) * %d;" % (bt.NINT, bt.DEG + 1),
          "        double v = c[%d];" % bt.DEG,
          "        int j;",
          "        for (j = %d; j >= 0; --j) v = v * t + c[j];" % (bt.DEG - 1),
          "        F[nmax] = v;",
          "        for (m = nmax - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + ex) / (2.0 * m + 1.0);",
          "    } else {",
          "        const double ex = exp(-T);",
          "        F[0] = 0.5 * sqrt(3.141592653589793238 / T);",
          "        for (m = 0; m < nmax; ++m) F[m + 1] = ((2.0 * m + 1.0) * F[m] - ex) / (2.0 * T);",
          "    }",
          "    for (m = nmin; m <= nmax; ++m) f[m * ifskip] = F[m];",
          "}", ""]
    return h, "\n".join(L)


INDEX_CHANGES = (-2, -1, 1)   # the ERI script's index changes: VRR a-1, a-2 (c-1, c-2); HRR a+1, b-1


def _global_data_files(lmax):
    """global_data.{h,c}: the reference's shared tables
    (src/intception/generator/generator.py:678-753, 1019-1157), computed here:
    ccs[n][3] the Cartesian exponents of flat component n (x^l first ... z^l
    last, every l <= 2 lmax, the VRR's highest level); jump[n][c][3] the change
    of the flat index when component d of n changes by INDEX_CHANGES[c] (0 where
    that would be negative); hrr_layer[la][lb] the largest HGP layer
    max_lq (ncc0(la + lb - lq) - ncc0(lb - 1)) ncc(lq), lq <= la."""
    lint = 2 * lmax
    comps = comps_upto(lint)
    n = len(comps)
    nch = len(INDEX_CHANGES)
    h = "\n".join([
        "/* Generated by intception_amd (c99 backend): shared index tables. */",
        "#ifndef INTCEPTION_GLOBAL_DATA_H", "#define INTCEPTION_GLOBAL_DATA_H",
        "#define INTCEPTION_NCC %d" % n, "#define INTCEPTION_NCHANGES %d" % nch,
        "#define INTCEPTION_HRR_LA %d" % (lint - lmax + 1), "#define INTCEPTION_HRR_LB %d" % (lmax + 1),
        "/* Cartesian exponents of flat component n (x^l first, ..., z^l last), l <= %d */" % lint,
        "extern const int ccs[INTCEPTION_NCC][3];",
        "/* flat-index change of component n when its d exponent changes by %s (0 if negative) */"
        % ", ".join(str(c) for c in INDEX_CHANGES),
        "extern const int jump[INTCEPTION_NCC][INTCEPTION_NCHANGES][3];",
        "/* largest HGP HRR layer for final angular momenta (la, lb) */",
        "extern const int hrr_layer[INTCEPTION_HRR_LA][INTCEPTION_HRR_LB];", "#endif", ""])
    L = ["/* Generated by intception_amd (c99 backend) -- do not edit. */", '#include "global_data.h"', "",
         "const int ccs[INTCEPTION_NCC][3] = {"]
    L += ["    {%d, %d, %d}," % c for c in comps]
    L += ["};", "", "const int jump[INTCEPTION_NCC][INTCEPTION_NCHANGES][3] = {"]
    for c in comps:
        rows = []
        for ch in INDEX_CHANGES:
            vals = []
            for d in range(3):
                t = list(c)
                t[d] += ch
                vals.append(cidx(tuple(t)) - cidx(c) if t[d] >= 0 else 0)
            rows.append("{%d, %d, %d}" % tuple(vals))
        L.append("    {%s}," % ", ".join(rows))
    L += ["};", "", "const int hrr_layer[INTCEPTION_HRR_LA][INTCEPTION_HRR_LB] = {"]
    for la in range(lint - lmax + 1):
        vals = [max((ncc0(la + lb - lq) - ncc0(lb - 1)) * ncc(lq) for lq in range(la + 1)) for lb in range(lmax + 1)]
        L.append("    {%s}," % ", ".join(str(v) for v in vals))
    L += ["};", ""]
    return h, "\n".join(L)


GLOBAL_FUNCTIONS_H = """/* Generated by intception_amd (c99 backend): shared transforms. */
#ifndef INTCEPTION_GLOBAL_FUNCTIONS_H
#define INTCEPTION_GLOBAL_FUNCTIONS_H
/* HGP HRR step on contracted values, (a, b) = (a + 1_d, b - 1_d) + Rfromto_d (a, b - 1_d),
 * d the first non-zero component of b: for every b at level loop_l_move_to (x^l first ...),
 * every a at levels l_move_from..lmax_move_from and every pair of contractions,
 *   work[iwrk0 + ib iskip_move_to0 + ia iskip_move_from0 + ka iskip_cont_move_from0 + kb iskip_cont_move_to0]
 *   = work[iwrk1 + (b - 1_d) iskip_move_to1 + (a + 1_d) iskip_move_from1 + ...]
 *   + Rfromto[d] work[iwrk1 + (b - 1_d) iskip_move_to1 + a iskip_move_from1 + ...],
 * with a counted over the levels from l_move_from up and b - 1_d within level loop_l_move_to - 1. */
void hrr(const double Rfromto[3], const int lmax_move_from, const int l_move_from, const int loop_l_move_to,
         const int iskip_move_from0, const int iskip_move_to0, const int iskip_move_from1, const int iskip_move_to1,
         const int ncont_move_from, const int ncont_move_to, const int iskip_cont_move_from0,
         const int iskip_cont_move_to0, const int iskip_cont_move_from1, const int iskip_cont_move_to1,
         double *work_array, const int iwrk0, const int iwrk1);
/* work[iwrk0 + ipre + icont n_pre] = sum_iprim work[iwrk1 + ipre + iprim iskip_prim]
 *                                    * contract_array[iprim + icont n_prim] */
void contract(double *work_array, const int iwrk0, const int iwrk1, const double *contract_array, const int n_pre,
              const int n_prim, const int n_cont, const int iskip_prim);
/* work[iwrk0 + ipre + isph n_pre] = sum_icart work[iwrk1 + ipre + icart iskip_cart]
 *                                   * sph_trans_array[icart + isph n_cart] */
void sph_trans(double *work_array, const int iwrk0, const int iwrk1, const double *sph_trans_array,
               const int n_pre, const int n_cart, const int n_sph, const int iskip_cart);
#endif
"""

GLOBAL_FUNCTIONS_C = """/* Generated by intception_amd (c99 backend) -- do not edit.
 * The reference's shared functions (hrr, contract, sph_trans), same arguments and meaning
 * (src/intception/generator/callables/generator_hrr_function.py:227-435,
 * generator_contract_function.py:69-125, generator_sph_trans_function.py:69-125). */
#include "global_functions.h"

static int gf_ncc0(int l) { return l < 0 ? 0 : (l + 1) * (l + 2) * (l + 3) / 6; }
/* position of (x, y, z) among the components of its level (x^l first) */
static int gf_within(int y, int z) { return (y + z) * (y + z + 1) / 2 + z; }

void hrr(const double Rfromto[3], const int lmax_move_from, const int l_move_from, const int loop_l_move_to,
         const int iskip_move_from0, const int iskip_move_to0, const int iskip_move_from1, const int iskip_move_to1,
         const int ncont_move_from, const int ncont_move_to, const int iskip_cont_move_from0,
         const int iskip_cont_move_to0, const int iskip_cont_move_from1, const int iskip_cont_move_to1,
         double *work_array, const int iwrk0, const int iwrk1)
{
    const int lt = loop_l_move_to, base = gf_ncc0(l_move_from - 1);
    int kt, kf, bx, by, it;
    for (kt = 0; kt < ncont_move_to; ++kt) {
        it = 0;
        for (bx = lt; bx >= 0; --bx)
            for (by = lt - bx; by >= 0; --by, ++it) {
                const int bz = lt - bx - by;
                /* build direction: the first non-zero component of b */
                const int d = bx > 0 ? 0 : (by > 0 ? 1 : 2);
                const int im1 = gf_within(by - (d == 1), bz - (d == 2));
                const double R = Rfromto[d];
                for (kf = 0; kf < ncont_move_from; ++kf) {
                    const int o0 = iwrk0 + it * iskip_move_to0 + kf * iskip_cont_move_from0 + kt * iskip_cont_move_to0;
                    const int o1 = iwrk1 + im1 * iskip_move_to1 + kf * iskip_cont_move_from1 +
                                   kt * iskip_cont_move_to1;
                    int l, ax, ay, ia = 0;
                    for (l = l_move_from; l <= lmax_move_from; ++l)
                        for (ax = l; ax >= 0; --ax)
                            for (ay = l - ax; ay >= 0; --ay, ++ia) {
                                const int az = l - ax - ay;
                                /* a + 1_d sits at level l + 1 */
                                const int ip = gf_ncc0(l) + gf_within(ay + (d == 1), az + (d == 2)) - base;
                                work_array[o0 + ia * iskip_move_from0] =
                                    work_array[o1 + ip * iskip_move_from1] + R * work_array[o1 + ia * iskip_move_from1];
                            }
                }
            }
    }
}

void contract(double *work_array, const int iwrk0, const int iwrk1, const double *contract_array, const int n_pre,
              const int n_prim, const int n_cont, const int iskip_prim)
{
    int k, p, i;
    for (k = 0; k < n_cont; ++k) {
        double *dst = work_array + iwrk0 + k * n_pre;
        for (i = 0; i < n_pre; ++i) dst[i] = 0.0;
        for (p = 0; p < n_prim; ++p) {
            const double w = contract_array[p + k * n_prim];
            const double *src = work_array + iwrk1 + p * iskip_prim;
            for (i = 0; i < n_pre; ++i) dst[i] += src[i] * w;
        }
    }
}

void sph_trans(double *work_array, const int iwrk0, const int iwrk1, const double *sph_trans_array,
               const int n_pre, const int n_cart, const int n_sph, const int iskip_cart)
{
    int s, c, i;
    for (s = 0; s < n_sph; ++s) {
        double *dst = work_array + iwrk0 + s * n_pre;
        for (i = 0; i < n_pre; ++i) dst[i] = 0.0;
        for (c = 0; c < n_cart; ++c) {
            const double w = sph_trans_array[c + s * n_cart];
            const double *src = work_array + iwrk1 + c * iskip_cart;
            if (w == 0.0) continue;
            for (i = 0; i < n_pre; ++i) dst[i] += src[i] * w;
        }
    }
}
"""

MAKEFILE = """# Generated by intception_amd (c99 backend): builds lib{lib}.a from every .c file here,
# as examples/Makefile of the reference does for its output directory.
CC ?= gcc
CSTD := -std=c99
COPT ?= -O3
WARN := -Wall
SRC := $(wildcard *.c)
OBJ := $(SRC:.c=.o)
lib{lib}.a : $(OBJ)
\tar rcs $@ $(OBJ)
%.o : %.c
\t$(CC) $(CSTD) $(COPT) $(WARN) -fPIC -MMD -c $< -o $@
.PHONY : clean
clean :
\trm -f $(OBJ) $(OBJ:.o=.d) lib{lib}.a
-include $(OBJ:.o=.d)
"""


def emit_file_set(full_name, spherical, lmax, classes=None, lib="intception"):
    """The reference's output file set for the ERI class
    (src/intception/generator/generator.py:1504-1702), from this build's
    recurrence structure: global_data.{c,h}, boys_function.{c,h},
    global_functions.{c,h}, <full>.{c,h} (the <full>_main dispatcher and
    <full>_work_array_size), one <full>_<abcd>.c per canonical class (so a
    parallel make compiles them side by side, declared in <full>_classes.h),
    and a Makefile for lib<lib>.a.  Returns {file name: text}."""
    classes = canonical_classes(lmax) if classes is None else sorted(canonical(*c)[0] for c in classes)
    kind = "sph" if spherical else "cart"
    files = {}
    files["global_data.h"], files["global_data.c"] = _global_data_files(lmax)
    files["boys_function.h"], files["boys_function.c"] = _boys_files(4 * lmax)
    files["global_functions.h"], files["global_functions.c"] = GLOBAL_FUNCTIONS_H, GLOBAL_FUNCTIONS_C
    files["Makefile"] = MAKEFILE.replace("{lib}", lib)
    args = []
    for X in "abcd":
        args.append("const double *x{0}, const double c{0}[3], const int l{0}, const int n{0}, const int iskip{0}, "
                    "const int n{0}_prim, const int n{0}_cont, const int iskip{0}_cont, "
                    "const double *contract_array_{0}".format(X))
        if spherical:
            args.append("const double *sph_trans_array_%s" % X)
    proto_main = "void %s_main(%s, double *work_array, double *output_array)" % (full_name, ", ".join(args))
    proto_size = ("int %s_work_array_size(const int la, const int lb, const int lc, const int ld, const int na_prim, "
                  "const int nb_prim, const int nc_prim, const int nd_prim, const int na_cont, const int nb_cont, "
                  "const int nc_cont, const int nd_cont)" % full_name)
    guard = full_name.upper() + "_H"
    files[full_name + ".h"] = "\n".join([
        "/* Generated by intception_amd (c99 backend) for the contracted ERI class, %s output, l <= %d. */"
        % ("spherical" if spherical else "Cartesian", lmax),
        "#ifndef %s" % guard, "#define %s" % guard, proto_main + ";", proto_size + ";", "#endif", ""])
    cguard = full_name.upper() + "_CLASSES_H"
    files[full_name + "_classes.h"] = "\n".join(
        ["/* Generated by intception_amd (c99 backend): canonical-class functions of %s. */" % full_name,
         "#ifndef %s" % cguard, "#define %s" % cguard,
         "/* contracted Cartesian block out[a][b][c][d] of shells A, B, C, D in canonical order",
         " * (la >= lb, lc >= ld, (la + lb, la) <= (lc + ld, lc)); coefficients ka.. normalised */"]
        + ["void %s(%s);" % (_class_fn(c, full_name), CLASS_ARGS) for c in classes] + ["#endif", ""])
    for cls in classes:
        files["%s_%d%d%d%d.c" % ((full_name,) + tuple(cls))] = "\n".join([
            "/* Generated by intception_amd (c99 backend) -- do not edit.",
            " * Contracted ERI class (%d %d | %d %d): Obara-Saika VRR on a then c, HGP HRR on the" % tuple(cls),
            " * contracted ket (c, d) and bra (a, b). */",
            "#include <math.h>", '#include "boys_function.h"', '#include "%s_classes.h"' % full_name, "",
            emit_class(cls, full_name)])
    S = ["/* Generated by intception_amd (c99 backend) -- do not edit.",
         " * <full>_main and <full>_work_array_size of the contracted ERI, %s output, shells up to l = %d. */"
         % (kind, lmax),
         "#include <stdio.h>", '#include "global_data.h"', '#include "global_functions.h"',
         '#include "%s.h"' % full_name, '#include "%s_classes.h"' % full_name, "",
         "typedef void (*eri_c_fn)(const double *, const double *, int, const double *, const double *, "
         "const double *, int, const double *, const double *, const double *, int, const double *, "
         "const double *, const double *, int, const double *, double *);",
         "static eri_c_fn class_fn(int la, int lb, int lc, int ld)", "{"]
    for cls in classes:
        S.append("    if (la == %d && lb == %d && lc == %d && ld == %d) return %s;" % (cls + (_class_fn(cls, full_name),)))
    S += ["    return 0;", "}", "",
          "static int c99_ncc(int l) { return (l + 1) * (l + 2) / 2; }",
          "static int c99_ncc0(int l) { return l < 0 ? 0 : (l + 1) * (l + 2) * (l + 3) / 6; }", ""]
    nbuf = ncc(lmax) ** 4
    main = [proto_main, "{",
            "    const double *xs[4] = {xa, xb, xc, xd}, *cs[4] = {ca, cb, cc, cd};",
            "    const double *ks[4] = {contract_array_a, contract_array_b, contract_array_c, contract_array_d};",
            "    const int ls[4] = {la, lb, lc, ld}, np[4] = {na_prim, nb_prim, nc_prim, nd_prim};",
            "    const int nco[4] = {na_cont, nb_cont, nc_cont, nd_cont};",
            "    const int isk[4] = {iskipa, iskipb, iskipc, iskipd};",
            "    const int iskc[4] = {iskipa_cont, iskipb_cont, iskipc_cont, iskipd_cont};",
            "    const int nn[4] = {na, nb, nc, nd};"]
    if spherical:
        main.append("    const double *ss[4] = {sph_trans_array_a, sph_trans_array_b, sph_trans_array_c, "
                    "sph_trans_array_d};")
    main += [
        "    int p[4], q, k0, k1, k2, k3;",
        "    int bra0 = 0, bra1 = 1, ket0 = 2, ket1 = 3;",
        "    eri_c_fn f;",
        "    /* blk: the canonical block; tmp: caller order; two buffers for the transforms */",
        "    double blk[%d], tmp[%d];" % (nbuf, 2 * nbuf),
        "    (void)work_array;",
        "    /* canonical class: la >= lb, lc >= ld, (la+lb, la) <= (lc+ld, lc) in the bra */",
        "    if (ls[0] < ls[1]) { bra0 = 1; bra1 = 0; }",
        "    if (ls[2] < ls[3]) { ket0 = 3; ket1 = 2; }",
        "    if (ls[bra0] + ls[bra1] > ls[ket0] + ls[ket1] ||",
        "        (ls[bra0] + ls[bra1] == ls[ket0] + ls[ket1] && ls[bra0] > ls[ket0])) {",
        "        p[0] = ket0; p[1] = ket1; p[2] = bra0; p[3] = bra1;",
        "    } else {",
        "        p[0] = bra0; p[1] = bra1; p[2] = ket0; p[3] = ket1;",
        "    }",
        "    for (q = 0; q < 4; ++q)",
        "        if (ls[q] < 0 || ls[q] > %d || nn[q] != %s) {" % (
            lmax, "2 * ls[q] + 1" if spherical else "c99_ncc(ls[q])"),
        '            fprintf(stderr, "%s_main: unsupported l or component count\\n");' % full_name,
        "            return;",
        "        }",
        "    f = class_fn(ls[p[0]], ls[p[1]], ls[p[2]], ls[p[3]]);",
        "    if (!f) {",
        '            fprintf(stderr, "%s_main: class not generated\\n");' % full_name,
        "        return;",
        "    }",
        "    for (k0 = 0; k0 < nco[0]; ++k0) for (k1 = 0; k1 < nco[1]; ++k1)",
        "    for (k2 = 0; k2 < nco[2]; ++k2) for (k3 = 0; k3 < nco[3]; ++k3) {",
        "        const int kk[4] = {k0, k1, k2, k3};",
        "        int nct[4], st[4], i0, i1, i2, i3, x;",
        "        double *res = tmp;",
        "        f(xs[p[0]], ks[p[0]] + kk[p[0]] * np[p[0]], np[p[0]], cs[p[0]],",
        "          xs[p[1]], ks[p[1]] + kk[p[1]] * np[p[1]], np[p[1]], cs[p[1]],",
        "          xs[p[2]], ks[p[2]] + kk[p[2]] * np[p[2]], np[p[2]], cs[p[2]],",
        "          xs[p[3]], ks[p[3]] + kk[p[3]] * np[p[3]], np[p[3]], cs[p[3]], blk);",
        "        /* canonical block [p0][p1][p2][p3] -> caller order [a][b][c][d] (Cartesian) */",
        "        for (x = 0; x < 4; ++x) nct[x] = c99_ncc(ls[x]);",
        "        st[3] = 1; st[2] = nct[3]; st[1] = nct[2] * nct[3]; st[0] = nct[1] * nct[2] * nct[3];",
        "        {",
        "            int n0 = nct[p[0]], n1 = nct[p[1]], n2 = nct[p[2]], n3 = nct[p[3]], t = 0;",
        "            for (i0 = 0; i0 < n0; ++i0) for (i1 = 0; i1 < n1; ++i1)",
        "            for (i2 = 0; i2 < n2; ++i2) for (i3 = 0; i3 < n3; ++i3)",
        "                tmp[i0 * st[p[0]] + i1 * st[p[1]] + i2 * st[p[2]] + i3 * st[p[3]]] = blk[t++];",
        "        }"]
    if spherical:
        main += [
            "        /* caller's S on each index with sph_trans, one call per slab of the slower",
            "         * indices: a block [outer][n_cart][inner] becomes [outer][n_sph][inner] */",
            "        {",
            "            int dims[4], ax, o, y, src = 0, dst = %d;" % nbuf,
            "            for (x = 0; x < 4; ++x) dims[x] = nct[x];",
            "            for (ax = 3; ax >= 0; --ax) {",
            "                const int nsp = 2 * ls[ax] + 1, nca = dims[ax];",
            "                int inner = 1, outer = 1;",
            "                for (y = ax + 1; y < 4; ++y) inner *= dims[y];",
            "                for (y = 0; y < ax; ++y) outer *= dims[y];",
            "                for (o = 0; o < outer; ++o)",
            "                    sph_trans(tmp, dst + o * nsp * inner, src + o * nca * inner, ss[ax], inner, nca, nsp,",
            "                              inner);",
            "                dims[ax] = nsp;",
            "                y = src; src = dst; dst = y;",
            "            }",
            "            res = tmp + src;",
            "        }"]
    main += [
        "        {",
        "            const int n0 = nn[0], n1 = nn[1], n2 = nn[2], n3 = nn[3];",
        "            int t = 0;",
        "            for (i0 = 0; i0 < n0; ++i0) for (i1 = 0; i1 < n1; ++i1)",
        "            for (i2 = 0; i2 < n2; ++i2) for (i3 = 0; i3 < n3; ++i3, ++t)",
        "                output_array[i0 * isk[0] + kk[0] * iskc[0] + i1 * isk[1] + kk[1] * iskc[1] +",
        "                             i2 * isk[2] + kk[2] * iskc[2] + i3 * isk[3] + kk[3] * iskc[3]] = res[t];",
        "        }",
        "    }",
        "}"]
    S += main
    S += ["", proto_size, "{",
          "    /* the reference's formula: 2-layer VRR buffer, early contractions in either order,",
          "     * HRR layer (hrr_layer of global_data), late transform",
          "     * (contracted_vrr_hrr_with_auxiliary.py:248-351) */",
          "    const int L = la + lb, M = L + lc + ld;",
          "    const long nvrr = (long)c99_ncc0(L) * c99_ncc0(lc) * c99_ncc0(ld);",
          "    const long nprim = (long)na_prim * nb_prim * nc_prim * nd_prim;",
          "    const int np[4] = {na_prim, nb_prim, nc_prim, nd_prim}, nk[4] = {na_cont, nb_cont, nc_cont, nd_cont};",
          "    const long ncont = (long)na_cont * nb_cont * nc_cont * nd_cont;",
          "    const int nc_c = %s, nc_d = %s;" % (("2 * lc + 1", "2 * ld + 1") if spherical else
                                                ("c99_ncc(lc)", "c99_ncc(ld)")),
          "    long mem = 2 * (nvrr * nprim + M + 1), nmax = nprim, late, layer;",
          "    int dir, s, cur[4];",
          "    if (lb < 0 || lb >= INTCEPTION_HRR_LA || la < 0 || la >= INTCEPTION_HRR_LB) return -1;",
          "    layer = hrr_layer[lb][la];",
          "    for (dir = 0; dir < 2; ++dir) {",
          "        for (s = 0; s < 4; ++s) cur[s] = np[s];",
          "        for (s = 0; s < 4; ++s) {",
          "            const int k = dir == 0 ? s : 3 - s;",
          "            long pr;",
          "            cur[k] = nk[k];",
          "            pr = (long)cur[0] * cur[1] * cur[2] * cur[3];",
          "            if (pr > nmax) nmax = pr;",
          "        }",
          "    }",
          "    if (2 * nvrr * nmax > mem) mem = 2 * nvrr * nmax;",
          "    if (2 * layer * nc_c * nc_d * ncont > mem) mem = 2 * layer * nc_c * nc_d * ncont;",
          "    late = %s;" % ("2L * c99_ncc(la) * c99_ncc(lb) * c99_ncc(lc) * c99_ncc(ld) * ncont" if spherical else
                               "2L * c99_ncc(la) * c99_ncc(lb) * nc_c * nc_d * ncont"),
          "    if (late > mem) mem = late;",
          "    return (int)mem;", "}", ""]
    files[full_name + ".c"] = "\n".join(S)
    return files


def write(outdir, full_name, spherical, lmax, classes=None, lib="intception"):
    """Write the file set of emit_file_set into outdir; returns the paths of
    its .c files."""
    os.makedirs(outdir, exist_ok=True)
    out = []
    for name, text in emit_file_set(full_name, spherical, lmax, classes, lib).items():
        path = os.path.join(outdir, name)
        old = open(path).read() if os.path.exists(path) else None
        if old != text:
            with open(path, "w") as fh:
                fh.write(text)
        if name.endswith(".c"):
            out.append(path)
    return sorted(out)
