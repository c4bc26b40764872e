"""HIP backend: emits one gfx950 kernel per canonical ERI class.

The recurrences are those of the reference's ERI input script -- Obara-Saika
VRR (bra index a, ket index c, one auxiliary index m) and the Head-Gordon-Pople
HRR (src/intception/dsl_extensions.py:822-950) -- with the reference's direction
rule, "first non-zero component of the target"
(src/intception/generator/generator.py:547-558).  The schedule is MI355X's
own rather than the reference's 2-layer work-array scheme
(src/intception/generator/callables/main_source_algorithms/contracted_general.py:720-1080):

* lanes over bra components.  A *group* of G = ncc0(la+lb) lanes works on one
  primitive quartet; lane e owns bra component e (all |e| <= la+lb).  Lane e
  builds (e 0|0 0)^(m) along the chain 0 -> ... -> e of the direction rule,
  with per-lane direction selects (one straight-line program for all lanes).
* ket in registers.  Every lane then runs the same unrolled OS recursion on the
  ket index c, (e 0|f 0)^(m) for |f| <= lc+ld; the only cross-lane term,
  e_i/(2(zeta+eta)) (e-1_i 0|f 0)^(m+1), is a ds_bpermute from the lane that
  owns e-1_i.
* contraction in registers.  Coefficients are folded into the pair prefactor,
  so contraction is a plain sum over primitive quartets; a *team* of T groups
  splits the primitive quartets of one contracted quartet and is reduced with
  cross-lane reads (tree over groups).
* HRR and transform.  Group 0 applies the ket HRR (c,d) and the ket
  Cartesian->spherical step in registers, writes (e|cd) to LDS, and the team's
  lanes take one cd column each for the bra HRR (a,b) and the bra transform,
  writing the output rows coalesced.

Canonical classes have la >= lb, lc >= ld and the smaller pair in the bra
(lanes); the batch planner permutes shells and output strides accordingly.
"""
import itertools
import os

import numpy as np

from .basis import SPH, ncc, ncc0, nsph, cart_components, sph_matrix

# built-in S_l (real solid harmonics), l = 0..7: compiled into the "_bs"
# kernels as literals and into the runtime as the device table used when a
# caller passes no matrices
SPH_BUILTIN = {l: (SPH[l] if l in SPH else sph_matrix(l)) for l in range(8)}


def comps_upto(l):
    out = []
    for k in range(l + 1):
        out += cart_components(k)
    return out


def cidx(c):
    """0-based index of component c among all components with |c'| <= |c|."""
    x, y, z = c
    l = x + y + z
    yz = y + z
    return ncc0(l - 1) + yz * (yz + 1) // 2 + z


def first_nonzero(c):
    for i in range(3):
        if c[i] > 0:
            return i
    return 0


def dec(c, i, k=1):
    c = list(c)
    c[i] -= k
    return tuple(c)


def canonical(la, lb, lc, ld):
    """Canonical class and the permutation applied.

    Returns (cls, perm) where cls = (LA, LB, LC, LD) and perm[k] is the original
    position (0..3 for a,b,c,d) of canonical position k."""
    bra, ket = [(la, 0), (lb, 1)], [(lc, 2), (ld, 3)]
    if bra[0][0] < bra[1][0]:
        bra = [bra[1], bra[0]]
    if ket[0][0] < ket[1][0]:
        ket = [ket[1], ket[0]]
    kb = (bra[0][0] + bra[1][0], bra[0][0])
    kk = (ket[0][0] + ket[1][0], ket[0][0])
    if kb > kk:
        bra, ket = ket, bra
    cls = (bra[0][0], bra[1][0], ket[0][0], ket[1][0])
    perm = (bra[0][1], bra[1][1], ket[0][1], ket[1][1])
    return cls, perm


def canonical_np(ls):
    """Vectorised canonical class for an (n, 4) array of angular momenta."""
    ls = np.asarray(ls)
    a, b = np.maximum(ls[:, 0], ls[:, 1]), np.minimum(ls[:, 0], ls[:, 1])
    c, d = np.maximum(ls[:, 2], ls[:, 3]), np.minimum(ls[:, 2], ls[:, 3])
    sw = (a + b > c + d) | ((a + b == c + d) & (a > c))
    return np.stack([np.where(sw, c, a), np.where(sw, d, b), np.where(sw, a, c), np.where(sw, b, d)], axis=1)


def canonical_classes(lmax):
    out = set()
    for c in itertools.product(range(lmax + 1), repeat=4):
        out.add(canonical(*c)[0])
    return sorted(out)


class Emitter:
    def __init__(self):
        self.lines = []
        self.ind = 1

    def __call__(self, s=""):
        self.lines.append("    " * self.ind + s)

    def text(self):
        return "\n".join(self.lines)


# groups of at least this many lanes whose ket reaches lc + ld in BATCH_LK use
# the batched setup loop (measured on MI355X, config-3 bins: 1.06-1.37x for
# ket 2-3, 2.0x for (ps|ps), 1.05x for (pp|dd) and (ps|dd)); a per-kernel
# choice (emit_class_source's batch) overrides the rule
BATCH_MIN_G = 4
BATCH_LK = (1, 2, 3, 4)
# batched kernels: each lane reads its bra-chain step's PA / WP entry from the
# setup slot (instead of selecting among the x, y, z entries)
SLOT_PW = True
# the top ket level's cross-lane term is added once per contracted quartet
# (_ket_lines defer_top) instead of once per primitive quartet
DEFER_TOP = True
# FP64 kernels: the top ket level goes into its contraction sums by fused
# multiply-adds instead of being formed and then added
FUSE_TOP = True


def _lds_conflict(addrs):
    """Largest number of distinct dwords that one LDS access maps to a single
    bank (64 banks of 4 bytes)."""
    banks = {}
    for a in addrs:
        banks.setdefault(a % 64, set()).add(a)
    return max(len(v) for v in banks.values())


def slot_stride(G, ns, ipa=None):
    """Slot length in doubles for the batched setup: the smallest length
    >= ns for which the wave's reads of a slot entry hit the fewest LDS bank
    conflicts.  At step st the 64 // G groups of a wave read slot
    (group * G + st): one entry each (broadcast within the group), or, for the
    bra chain's PA / WP entries, one of the four at ipa.. per lane."""
    ng = 64 // G
    best = None
    for n in range(ns, ns + 9):
        worst = 0
        for st in range(G):
            for i in range(n if ipa is None else ipa + 1):
                ad = [2 * ((g * G + st) * n + i) + h for g in range(ng) for h in (0, 1)]
                worst = max(worst, _lds_conflict(ad))
            if ipa is not None:
                ad = [2 * ((g * G + st) * n + ipa + k) + h for g in range(ng) for k in range(4) for h in (0, 1)]
                worst = max(worst, _lds_conflict(ad))
        if best is None or worst < best[0]:
            best = (worst, n)
    return best[1]


def slot_stride_f(G, ns, ipa=None):
    """Float-slot length (even, so the prefactor double at floats 0-1 stays
    8-byte aligned) for the FP32 kernels' batched setup: the smallest length
    >= ns with the fewest LDS bank conflicts for the wave's reads at a step
    (slot group * G + st, one dword each; the bra chain's PA / WP reads, four
    dwords from ipa per group)."""
    ng = 64 // G
    best = None
    for n in range(ns + (ns & 1), ns + 18, 2):
        worst = 0
        for st in range(G):
            for i in range(ns):
                worst = max(worst, _lds_conflict([(g * G + st) * n + i for g in range(ng)]))
            if ipa is not None:
                ad = [(g * G + st) * n + ipa + k for g in range(ng) for k in range(4)]
                worst = max(worst, _lds_conflict(ad))
        if best is None or worst < best[0]:
            best = (worst, n)
    return best[1]


def kernel_name(cls, spherical, prec="d"):
    return "intc_eri_%s%d%d%d%d_%s" % (prec, *cls, "s" if spherical else "c")


def _chain_lines(LX, M, R="double", slot_pw=None):
    """Bra chain: c{s}_{m} for s = 0..LX, m = 0..M-s; c0 given (type R).

    slot_pw = (iPA, iWP): the setup slot u holds PA_x, PA_y, PA_z, 1 at iPA
    and WP_x, WP_y, WP_z, 0 at iWP, and each lane reads its step direction's
    entry directly instead of selecting among the three components."""
    one = "1.0f" if R == "float" else "1.0"
    zero = "0.0f" if R == "float" else "0.0"
    L = []
    for s in range(1, LX + 1):
        L.append("const int d%d = (cdir >> %d) & 3;" % (s, 2 * (s - 1)))
        if slot_pw:
            L.append("const %s pa%d = (%s)u[%d + d%d];" % (R, s, R, slot_pw[0], s))
            L.append("const %s wp%d = (%s)u[%d + d%d];" % (R, s, R, slot_pw[1], s))
        else:
            L.append("const %s pa%d = d%d == 0 ? PAx : (d%d == 1 ? PAy : (d%d == 2 ? PAz : %s));"
                     % (R, s, s, s, s, one))
            L.append("const %s wp%d = d%d == 0 ? WPx : (d%d == 1 ? WPy : (d%d == 2 ? WPz : %s));"
                     % (R, s, s, s, s, zero))
        if s >= 2:
            L.append("const %s kk%d = (%s)((ckk >> %d) & 15) * o2z;" % (R, s, R, 4 * (s - 1)))
        for m in range(M - s + 1):
            expr = "pa%d * c%d_%d + wp%d * c%d_%d" % (s, s - 1, m, s, s - 1, m + 1)
            if s >= 2:
                expr += " + kk%d * (c%d_%d - roz * c%d_%d)" % (s, s - 2, m, s - 2, m + 1)
            L.append("const %s c%d_%d = %s;" % (R, s, m, expr))
    return L


def _ket_lines(LX, LR, R="double", defer_top=False, alt=False, fuse_top=None):
    """Ket OS VRR in registers: V{f}_{m} for |f| <= LR, m <= LR-|f| (type R).

    defer_top: the top level |f| = LR leaves out its cross-lane term
    e_i/(2(zeta+eta)) (e-1_i 0|f-1_i 0)^(1).  That level only feeds the
    contraction sums, so the term is linear in them: the caller accumulates
    o2s (e 0|f' 0)^(1) per lane (f' of level LR-1) over the primitive quartets
    and adds e_i times the sum of lane e-1_i once per contracted quartet
    (top_coupling).

    alt (la+lb = 2): the first level's neighbour value in direction y or z
    comes from the lane's own second chain, alt_m = (1_ja 0|0 0)^(m), where
    amask says the predecessor e - 1_i is off the lane's chain.

    fuse_top (FP64 only): name(f) of the contraction sum of each top-level
    value; the top level is then not formed, its terms go into the sum by one
    fused multiply-add each (one addition fewer per value and primitive
    quartet).

    Returns (lines, name(f, m))."""
    L = []
    shfl = "intc_shflf" if R == "float" else "intc_shfl"
    lit = "f" if R == "float" else ""

    def V(f, m):
        return "v%d_%d" % (cidx(f), m)
    for m in range(LR + 1):
        L.append("const %s %s = c%d_%d;" % (R, V((0, 0, 0), m), LX, m))
    for n in range(1, LR + 1):
        for t in cart_components(n):
            i = first_nonzero(t)
            f1 = dec(t, i)
            k1 = f1[i]
            ax = "xyz"[i]
            for m in range(LR - n + 1):
                if fuse_top is not None and n == LR:
                    a = fuse_top(t)
                    L.append("%s = fma(QC%s, %s, %s);" % (a, ax, V(f1, m), a))
                    L.append("%s = fma(WQ%s, %s, %s);" % (a, ax, V(f1, m + 1), a))
                    if k1 > 0:
                        f2 = dec(f1, i)
                        L.append("%s = fma(%d.0 * o2e, %s - roe * %s, %s);" % (a, k1, V(f2, m), V(f2, m + 1), a))
                    assert LX == 0 or defer_top, "fused top level without its deferred coupling"
                    continue
                expr = "QC%s * %s + WQ%s * %s" % (ax, V(f1, m), ax, V(f1, m + 1))
                if k1 > 0:
                    f2 = dec(f1, i)
                    expr += " + %d.0%s * o2e * (%s - roe * %s)" % (k1, lit, V(f2, m), V(f2, m + 1))
                if LX > 0 and not (defer_top and n == LR):
                    if n == 1 and (LX == 1 or i == 0):
                        # the first level's neighbour value (e - 1_i 0|0 0)^(m+1) is
                        # the lane's own chain value before the last (the chain
                        # builds e from e - 1_i0, i0 = first non-zero of e; i = x
                        # is i0 wherever e_x > 0, and with |e| <= 1 every i is)
                        expr += " + cpl%s * c%d_%d" % (ax, LX - 1, m + 1)
                    elif n == 1 and alt:
                        expr += " + cpl%s * ((amask & %d) ? alt_%d : c%d_%d)" % (ax, 1 << i, m + 1, LX - 1, m + 1)
                    else:
                        expr += " + cpl%s * %s(%s, s4%s)" % (ax, shfl, V(f1, m + 1), ax)
                L.append("const %s %s = %s;" % (R, V(t, m), expr))
    return L, V


def fp32_err_c(cls):
    """Constant C of the FP32 kernels' error estimate C 2^-24 max|(e 0|f 0)|:
    8 * 2^(la+lb+lc+ld).  On quartets of the config-3 molecule's shells the
    largest measured ratio of the FP32 error to 2^-24 max|(e 0|f 0)| was 8
    to 130 by class (oracle reference), at least 2x below this constant."""
    return 8.0 * 2.0 ** sum(cls)


def top_coupling(LR):
    """The deferred top-level terms: (target t, direction i, source f1 = t - 1_i)
    for every |t| = LR, i = first non-zero of t."""
    out = []
    for t in cart_components(LR):
        i = first_nonzero(t)
        out.append((t, i, dec(t, i)))
    return out


def _hrr_lines(src, lo, nmove, shift_vec, prefix, kind):
    """Generic HGP HRR on one index pair.

    src(e) gives the expression of (e, 0) for |e| in [lo, lo+nmove]; returns
    lines and a function H(a, b) giving (a, b) for |a| = lo, |b| = nmove.
    (a, b+1_i) = (a+1_i, b) + R_i (a, b), i = first non-zero of the target b.
    """
    L = []
    names = {}
    for e in comps_upto(lo + nmove):
        if sum(e) >= lo:
            names[(e, (0, 0, 0))] = src(e)
    for bl in range(1, nmove + 1):
        for e in comps_upto(lo + nmove - bl):
            if sum(e) < lo:
                continue
            for b in cart_components(bl):
                i = first_nonzero(b)
                b1 = dec(b, i)
                e1 = list(e); e1[i] += 1; e1 = tuple(e1)
                nm = "%s%d_%d" % (prefix, cidx(e), cidx(b))
                L.append("const double %s = %s + %s%s * %s;" % (nm, names[(e1, b1)], shift_vec, "xyz"[i],
                                                                names[(e, b1)]))
                names[(e, b)] = nm
    return L, (lambda a_, b_: names[(a_, b_)])


def emit_class_source(cls, spherical, prec="d", waves_per_eu=None, batch=None):
    """prec "d": FP64 throughout.  prec "f": the per-primitive recurrences (bra
    chain, ket VRR, cross-lane terms) in FP32 on the Boys values F_m; FP64 for
    the Boys function, the prefactor exp(-mu R^2) terms, the contraction sums
    (acc += prefactor * (double) V) and the contracted HRR/transform."""
    R = "double" if prec == "d" else "float"
    LA, LB, LC, LD = cls
    LX, LR = LA + LB, LC + LD
    M = LX + LR
    G = ncc0(LX)
    eoff = ncc0(LA - 1)
    NE = G - eoff
    nC = nsph(LC) if spherical else ncc(LC)
    nD = nsph(LD) if spherical else ncc(LD)
    NCD = nC * nD
    nA = nsph(LA) if spherical else ncc(LA)
    nB = nsph(LB) if spherical else ncc(LB)
    name = kernel_name(cls, spherical, prec)
    bra_comps = comps_upto(LX)
    ket_comps = comps_upto(LR)

    lane_tab = []
    for e in bra_comps:
        src = [cidx(dec(e, i)) if e[i] > 0 else cidx(e) for i in range(3)]
        steps = []
        t = e
        while sum(t) > 0:
            i = first_nonzero(t)
            steps.append((i, t[i] - 1))
            t = dec(t, i)
        steps = steps[::-1]
        # padding steps (direction 3: PA 1, WP 0) come first, so that the
        # chain's value before the last is e's predecessor e - 1_i0 in every
        # lane (i0 = first non-zero of e); the ket's first level reads it
        # there instead of from another lane
        pad = LX - len(steps)
        dirs, ks = 0, 0
        for s in range(LX):
            d, k = (3, 0) if s < pad else steps[s - pad]
            dirs |= d << (2 * s)
            ks |= k << (4 * s)
        # first-level neighbour values off the lane's chain: for i with e_i > 0
        # and i != i0, the predecessor e - 1_i; with la+lb = 2 that is at most
        # one |p| = 1 component per lane, (1_ja 0|0 0) = one chain step
        alt_dir, alt_mask, alt_p = 3, 0, None
        i0 = first_nonzero(e)
        for i in range(3 if LX == 2 else 0):
            if e[i] > 0 and i != i0:
                p = dec(e, i)
                assert alt_p in (None, p) and sum(p) == 1
                alt_p, alt_mask = p, alt_mask | (1 << i)
                alt_dir = first_nonzero(p)
        lane_tab.append((e[0], e[1], e[2], src[0], src[1], src[2], dirs, ks, alt_dir, alt_mask))

    o = Emitter()
    o.ind = 0
    o("// Generated by intception_amd/codegen.py -- do not edit.")
    o("// Canonical class (%d %d | %d %d), %s output, %s recurrences." % (
        LA, LB, LC, LD, "spherical" if spherical else "Cartesian", "FP64" if prec == "d" else "FP32"))
    o('#include "../intc_common.h"')
    o("")
    # Batched setup (groups of G >= BATCH_MIN_G lanes): lane e of a group
    # computes the Boys values and recurrence coefficients of the group's
    # primitive quartet kb + e into its LDS slot, then the group walks the G
    # slots.  The setup (pair loads, Boys, exp, divisions) is then issued once
    # per G primitive quartets instead of once per quartet in every lane.
    batched = (G >= BATCH_MIN_G and LR in BATCH_LK) if batch is None else (bool(batch) and G >= 2)
    slot = []          # (name, C type in the recurrences) in slot order
    for m in range(M + 1):
        slot.append(("c0_%d" % m, R))
    if LX > 0:
        # PA and WP each padded with the value of the 'no step' direction (3)
        slot += [(n, R) for n in ("PAx", "PAy", "PAz", "ONE", "WPx", "WPy", "WPz", "ZERO", "roz")]
        slot += [("o2z", "double")]
    if LR > 0:
        slot += [(n, R) for n in ("QCx", "QCy", "QCz", "WQx", "WQy", "WQz", "roe")] + [("o2e", "double")]
        if LX > 0:
            slot.append(("o2s", R))
    if prec == "f":
        slot.append(("pref", "double"))
    names = [nm for nm, _ in slot]
    # FP32 recurrences: a float slot, the prefactor as one aligned double at
    # floats 0-1 and every other entry one float (o2z and o2e as well: the
    # recurrences read them as floats), so a step reads half the bytes and
    # converts nothing.  NS stays the slot length in doubles (LDS sizing).
    fslot = batched and prec == "f"
    foff, NSF = {}, 0
    if fslot:
        pos = 2
        foff["pref"] = 0
        for nm in names:
            if nm != "pref":
                foff[nm] = pos
                pos += 1
        NSF = slot_stride_f(G, pos, foff["PAx"] if LX > 0 else None)
        NS = NSF // 2
    else:
        NS = slot_stride(G, len(slot), names.index("PAx") if LX > 0 else None) if batched else 0

    # The body is shared by the per-class kernels below and by the persistent
    # multi-class kernel, which includes this file with INTC_BODIES_ONLY
    # defined (file-scope names have internal linkage for that reason).
    o("static __constant__ int %s_lane[%d][10] = {" % (name, G))
    for r in lane_tab:
        o("    {%s}," % ", ".join(str(v) for v in r))
    o("};")
    o("// G lanes per group, NE kept bra rows, NCD ket columns, NS setup-slot length")
    o("static constexpr int %s_G = %d, %s_NE = %d, %s_NCD = %d, %s_NS = %d;" % (name, G, name, NE, name, NCD,
                                                                                name, NS))
    o("// LDS doubles per wave for team size T: max(epilogue transpose of its teams, 64 setup slots)")
    o("__host__ __device__ constexpr int %s_reg(int T)" % name)
    o("{")
    o("    return (64 / (%d * T)) * %d > 64 * %d ? (64 / (%d * T)) * %d : 64 * %d;" % (G, NE * NCD, NS, G, NE * NCD,
                                                                                     NS))
    o("}")
    o("")
    # BM: Boys scheme (0 Chebyshev, 1 reference); BS: 1 = the built-in real
    # solid-harmonic matrices are compiled in as literals (zero terms dropped),
    # 0 = the caller's matrices are read from a.sph.  The calling wave works on
    # queue positions qq0 + team, qq0 + team + qstep, ... below qend (and below
    # the bin's length); wlds is its LDS region of NAME_reg(T) doubles.
    o("template <int BM, int BS>")
    o("__device__ __forceinline__ void %s_body(const IntcEriArgs &a, double *wlds, const int qq0, const int qstep,"
      " int qend)" % name)
    o("{")
    o.ind = 1
    o("constexpr int G = %d, NE = %d, NCD = %d, M = %d, NS = %d;" % (G, NE, NCD, M, NS))
    if fslot:
        o("constexpr int NSF = %d;  // float-slot length" % NSF)
    o("const int T = a.team_groups, TL = G * T, TPW = 64 / TL;")
    o("const int lane = threadIdx.x & 63;")
    o("const int tw = lane / TL, tl = lane - tw * TL;")
    o("const int grp = tl / G, e = tl - grp * G;")
    o("if (tw >= TPW) return;")
    o("// per-wave LDS region: the setup slots (batched loop), then the epilogue transpose")
    o("double *tlds = wlds + (size_t)tw * (NE * NCD);")
    o("const int *lt = %s_lane[e];" % name)
    o("const %s ex_d = (%s)lt[0], ey_d = (%s)lt[1], ez_d = (%s)lt[2];" % (R, R, R, R))
    o("const int lbase = lane - e;")
    o("const int s4x = (lbase + lt[3]) * 4, s4y = (lbase + lt[4]) * 4, s4z = (lbase + lt[5]) * 4;")
    o("const int cdir = lt[6], ckk = lt[7], adir = lt[8], amask = lt[9];")
    o("(void)adir; (void)amask;")
    o("(void)ex_d; (void)ey_d; (void)ez_d; (void)s4x; (void)s4y; (void)s4z; (void)cdir; (void)ckk;")
    o("const int nq = a.nq_dev ? *a.nq_dev : a.nq;")
    o("if (qend > nq) qend = nq;")
    # software pipeline over the team's quartets: each pass loads the pair ids
    # of the next quartet and the queue entry of the one after it, so the
    # pass starts with its pair ids at hand (the same entries, under the same
    # guards, read one pass earlier).  A deeper pipeline (pair offsets a pass
    # ahead too) measured no faster.  The multi-class kernel, which includes
    # these bodies and holds every class at its register limit, keeps the
    # plain loop head.
    o("#ifdef INTC_BODIES_ONLY")
    o("for (int qq = qq0 + tw; qq < qend; qq += qstep) {")
    o("    const int q = a.q_idx ? a.q_idx[qq] : qq;")
    o("    if (q < 0 || q >= a.nq) continue;  // queue entries index the bin")
    o("    const int pb = a.q_bra[q], pk = a.q_ket[q];")
    o("#else")
    o("int q1 = qq0 + tw < qend ? (a.q_idx ? a.q_idx[qq0 + tw] : qq0 + tw) : -1;")
    o("int pb1 = 0, pk1 = 0;")
    o("if (q1 >= 0 && q1 < a.nq) { pb1 = a.q_bra[q1]; pk1 = a.q_ket[q1]; }")
    o("int q2 = qq0 + tw + qstep < qend ? (a.q_idx ? a.q_idx[qq0 + tw + qstep] : qq0 + tw + qstep) : -1;")
    o("for (int qq = qq0 + tw; qq < qend; qq += qstep) {")
    o("    const int q = q1, pb = pb1, pk = pk1;")
    o("    q1 = q2;")
    o("    if (q1 >= 0 && q1 < a.nq) { const unsigned o4 = (unsigned)q1 * 4u; pb1 = intc_ld(a.q_bra, o4); pk1 = intc_ld(a.q_ket, o4); }")
    o("    q2 = qq + 2 * qstep < qend ? (a.q_idx ? a.q_idx[qq + 2 * qstep] : qq + 2 * qstep) : -1;")
    o("    if (q < 0 || q >= a.nq) continue;  // queue entries index the bin")
    o("    const long long qoff = intc_ld(a.q_off, (unsigned)q * 8u);")
    o("    const unsigned long long qs = intc_ld(a.q_str, (unsigned)q * 8u);")
    o("#endif")
    o.ind = 2
    o("const int ob = intc_ld(a.pair_off, (unsigned)pb * 4u), Pb = intc_ld(a.pair_np, (unsigned)pb * 4u);")
    o("const int okk = intc_ld(a.pair_off, (unsigned)pk * 4u), Pk = intc_ld(a.pair_np, (unsigned)pk * 4u);")
    o("const double Ax = a.pair_A[3 * pb], Ay = a.pair_A[3 * pb + 1], Az = a.pair_A[3 * pb + 2];")
    o("const double Cx = a.pair_A[3 * pk], Cy = a.pair_A[3 * pk + 1], Cz = a.pair_A[3 * pk + 2];")
    o("(void)Ax; (void)Ay; (void)Az; (void)Cx; (void)Cy; (void)Cz;")
    o("const int P = Pb * Pk;")
    acc_f = [f for f in ket_comps if sum(f) >= LC]
    for f in acc_f:
        o("double acc%d = 0.0;" % cidx(f))
    # deferred top-level coupling: per lane, sum over primitive quartets of
    # o2s (e 0|f1 0)^(1) for every f1 of level LR - 1 (FP64, prefactor included)
    defer = DEFER_TOP and LX > 0 and LR > 0
    defer_src = cart_components(LR - 1) if defer else []
    for f1 in defer_src:
        o("double bq%d = 0.0;" % cidx(f1))
    def emit_setup(load_idx):
        """pair loads + Boys + coefficients for primitive quartet (ib, ik)"""
        # one 32-bit byte offset per primitive pair, shared by its six arrays
        o("const unsigned bo = (unsigned)(ob + ib) * 8u, ko = (unsigned)(okk + ik) * 8u;")
        o("const double zeta = intc_ld(a.pp_zeta, bo), o2z = intc_ld(a.pp_o2z, bo);")
        o("const double Px = intc_ld(a.pp_px, bo), Py = intc_ld(a.pp_py, bo), Pz = intc_ld(a.pp_pz, bo);")
        o("const double kab = intc_ld(a.pp_kappa, bo);")
        o("const double eta = intc_ld(a.pp_zeta, ko), o2e = intc_ld(a.pp_o2z, ko);")
        o("const double Qx = intc_ld(a.pp_px, ko), Qy = intc_ld(a.pp_py, ko), Qz = intc_ld(a.pp_pz, ko);")
        o("const double kcd = intc_ld(a.pp_kappa, ko);")
        o("(void)o2z; (void)o2e;")
        if load_idx:
            o("ik += T; while (ik >= Pk) { ik -= Pk; ++ib; }")
        # 1/sqrt(zeta + eta) gives both 1/(zeta + eta) and the prefactor's root
        o("const double ros = intc_rsqrt(zeta + eta), os = ros * ros;")
        o("const double rho = zeta * eta * os;")
        o("const double PQx = Px - Qx, PQy = Py - Qy, PQz = Pz - Qz;")
        o("const double Tb = rho * (PQx * PQx + PQy * PQy + PQz * PQz);")
        o("const double pref = kab * kcd * ros;")
        o("double F[M + 1];")
        o("intc_boys<M, BM>(Tb, F, a);")
        if prec == "f":
            # FP32 recurrences on s^m F_m with s = max(1, 2T): F_m(T) ~ (2m-1)!! /
            # (2T)^m F_0 underflows FP32 for large T, s^m F_m stays O(F_0).  Every
            # term that raises m by one carries 1/s; results at m = 0 are unscaled.
            # The prefactor multiplies in FP64 at accumulation.
            o("const float o2zf = (float)o2z, o2ef = (float)o2e;")
            o("(void)o2zf; (void)o2ef;")
            o("const double sc = Tb > 0.5 ? 2.0 * Tb : 1.0, isc = 1.0 / sc;")
        else:
            o("const double isc = 1.0;")
        if LX > 0:
            o("const double eo = eta * os;")
            o("const %s PAx = (%s)(Px - Ax), PAy = (%s)(Py - Ay), PAz = (%s)(Pz - Az);" % (R, R, R, R))
            o("const %s WPx = (%s)(-eo * PQx * isc), WPy = (%s)(-eo * PQy * isc), WPz = (%s)(-eo * PQz * isc);"
              % (R, R, R, R))
            o("const %s roz = (%s)(eo * isc);" % (R, R))
        if LR > 0:
            o("const double zo = zeta * os;")
            o("const %s QCx = (%s)(Qx - Cx), QCy = (%s)(Qy - Cy), QCz = (%s)(Qz - Cz);" % (R, R, R, R))
            o("const %s WQx = (%s)(zo * PQx * isc), WQy = (%s)(zo * PQy * isc), WQz = (%s)(zo * PQz * isc);"
              % (R, R, R, R))
            o("const %s roe = (%s)(zo * isc);" % (R, R))
            if LX > 0:
                o("const %s o2s = (%s)(0.5 * os * isc);" % (R, R))
        if prec == "f":
            o("double scm = 1.0;")
            for m in range(M + 1):
                o("const float c0_%d = (float)(F[%d] * scm);%s" % (m, m, " scm *= sc;" if m < M else ""))
        else:
            for m in range(M + 1):
                o("const double c0_%d = pref * F[%d];" % (m, m))

    # second chain of the lanes whose first-level neighbour is off their own
    # chain (la+lb = 2, a ket with a first level below its top)
    use_alt = LX == 2 and LR >= 1 and not (DEFER_TOP and LR == 1)

    def _alt_lines(R, slot_pw):
        """alt_m = (1_ja 0|0 0)^(m) = PA_ja c0_m + WP_ja c0_{m+1}, m = 1..LR"""
        L = []
        if slot_pw:
            L.append("const %s apa = (%s)u[%d + adir], awp = (%s)u[%d + adir];" % (R, R, slot_pw[0], R, slot_pw[1]))
        else:
            one, zero = ("1.0f", "0.0f") if R == "float" else ("1.0", "0.0")
            L.append("const %s apa = adir == 0 ? PAx : (adir == 1 ? PAy : (adir == 2 ? PAz : %s));" % (R, one))
            L.append("const %s awp = adir == 0 ? WPx : (adir == 1 ? WPy : (adir == 2 ? WPz : %s));" % (R, zero))
        for m in range(1, LR + 1):
            L.append("const %s alt_%d = apa * c0_%d + awp * c0_%d;" % (R, m, m, m + 1))
        return L

    def emit_recurrences(slot_pw=None):
        """bra chain + ket VRR + accumulation on the names of the setup"""
        if LX > 0 and LR > 0:
            o("const %s cplx = ex_d * o2s, cply = ey_d * o2s, cplz = ez_d * o2s;" % R)
        if prec == "f":
            body = _chain_lines(LX, M, R, slot_pw)
            body += _alt_lines(R, slot_pw) if use_alt else []
            lines, V = _ket_lines(LX, LR, R, defer, use_alt)
            body += lines
            body = [ln.replace("o2z", "o2zf").replace("o2e", "o2ef") for ln in body]
            for ln in body:
                o(ln)
            for f in acc_f:
                o("acc%d = fma(pref, (double)%s, acc%d);" % (cidx(f), V(f, 0), cidx(f)))
            if defer:
                # o2s and the m = 1 values carry the FP32 scale 1/s and s
                o("const double po2s = pref * (double)o2s;")
                for f1 in defer_src:
                    o("bq%d = fma(po2s, (double)%s, bq%d);" % (cidx(f1), V(f1, 1), cidx(f1)))
        else:
            for ln in _chain_lines(LX, M, "double", slot_pw):
                o(ln)
            for ln in (_alt_lines("double", slot_pw) if use_alt else []):
                o(ln)
            fuse = FUSE_TOP and LR > 0 and (LX == 0 or defer)
            lines, V = _ket_lines(LX, LR, "double", defer, use_alt,
                                  (lambda f: "acc%d" % cidx(f)) if fuse else None)
            for ln in lines:
                o(ln)
            for f1 in defer_src:
                o("bq%d = fma(o2s, %s, bq%d);" % (cidx(f1), V(f1, 1), cidx(f1)))
            for f in acc_f:
                if not (fuse and sum(f) == LR):
                    o("acc%d += %s;" % (cidx(f), V(f, 0)))

    if not batched:
        o("int ib = grp / Pk, ik = grp - ib * Pk;")
        o("for (int p = grp; p < P; p += T) {")
        o.ind = 3
        emit_setup(True)
        emit_recurrences()
    else:
        o("// this group's primitive quartets are p = grp + k T, k < K")
        o("const int K = P > grp ? (P - grp + T - 1) / T : 0;")
        if fslot:
            o("float *myslot = (float *)wlds + (size_t)lane * NSF;")
            o("const float *gslot = (const float *)wlds + (size_t)(lane - e) * NSF;")
        else:
            o("double *myslot = wlds + (size_t)lane * NS;")
            o("const double *gslot = wlds + (size_t)(lane - e) * NS;")
        o("for (int kb = 0; kb < K; kb += G) {")
        o.ind = 3
        o("if (kb + e < K) {")
        o.ind = 4
        o("const int pq = grp + (kb + e) * T;")
        o("const int ib = pq / Pk, ik = pq - ib * Pk;")
        emit_setup(False)
        for i, (nm, ty) in enumerate(slot):
            if fslot and nm == "pref":
                o("*(double *)(myslot + %d) = pref;" % foff[nm])
            elif fslot:
                val = {"ONE": "1.0f", "ZERO": "0.0f"}.get(nm, "(float)" + nm)
                o("myslot[%d] = %s;" % (foff[nm], val))
            else:
                val = {"ONE": "1.0", "ZERO": "0.0"}.get(nm, "(double)" + nm)
                o("myslot[%d] = %s;" % (i, val))
        o.ind = 3
        o("}")
        o("__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, \"wavefront\");")
        o("__builtin_amdgcn_wave_barrier();")
        o("const int ns = (K - kb < G) ? K - kb : G;")
        o("for (int st = 0; st < ns; ++st) {")
        o.ind = 4
        o("const %s *u = gslot + st * %s;" % (("float", "NSF") if fslot else ("double", "NS")))
        # PA / WP are read per lane by the bra chain (its step direction)
        lane_read = {"PAx", "PAy", "PAz", "ONE", "WPx", "WPy", "WPz", "ZERO"} if SLOT_PW else {"ONE", "ZERO"}
        for i, (nm, ty) in enumerate(slot):
            if nm in lane_read:
                continue
            if fslot and nm == "pref":
                o("const double pref = *(const double *)(u + %d);" % foff[nm])
            elif fslot and nm in ("o2z", "o2e"):
                o("const float %sf = u[%d];" % (nm, foff[nm]))
            elif fslot:
                o("const float %s = u[%d];" % (nm, foff[nm]))
            elif ty == "double":
                o("const double %s = u[%d];" % (nm, i))
            else:
                o("const %s %s = (%s)u[%d];" % (ty, nm, ty, i))
        if prec == "f" and not fslot:
            if LX > 0:
                o("const float o2zf = (float)o2z;")
            if LR > 0:
                o("const float o2ef = (float)o2e;")
        names = [nm for nm, _ in slot]
        if fslot:
            emit_recurrences((foff["PAx"], foff["WPx"]) if LX > 0 and SLOT_PW else None)
        else:
            emit_recurrences((names.index("PAx"), names.index("WPx")) if LX > 0 and SLOT_PW else None)
        o.ind = 3
        o("}")
        o("__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, \"wavefront\");")
        o("__builtin_amdgcn_wave_barrier();")
    o.ind = 2
    o("}")
    # team reduction over groups (tree towards group 0)
    o("for (int s = 1; s < T; s <<= 1) {")
    o.ind = 3
    o("const int src4 = (lane + s * G < 64 ? lane + s * G : lane) * 4;")
    o("const bool take = ((grp & (2 * s - 1)) == 0) && (grp + s < T);")
    for f in acc_f:
        o("{ const double r = intc_shfl(acc%d, src4); if (take) acc%d += r; }" % (cidx(f), cidx(f)))
    for f1 in defer_src:
        o("{ const double r = intc_shfl(bq%d, src4); if (take) bq%d += r; }" % (cidx(f1), cidx(f1)))
    o.ind = 2
    o("}")
    if defer:
        # the deferred top-level term, every lane (cross-lane reads of group 0's sums)
        o("// top ket level: + e_i (sum of o2s (e-1_i 0|t-1_i 0)^(1)) from lane e - 1_i")
        for t, i, f1 in top_coupling(LR):
            ax = "xyz"[i]
            o("acc%d = fma((double)e%s_d, intc_shfl(bq%d, s4%s), acc%d);" % (cidx(t), ax, cidx(f1), ax, cidx(t)))
    if prec == "f":
        # a-posteriori error estimate of the FP32 recurrences: C 2^-24 times
        # the largest contracted (e 0|f 0) of the quartet (FP32_ERR_C: the
        # rounding of the recurrence values, amplified by the contraction, the
        # HRR and the transforms, stays below that); over the tolerance the
        # quartet goes to the FP64 kernel's redo queue
        o("if (a.fp32_tol > 0.0) {")
        o.ind += 1
        o("float mx = 0.0f;")
        o("if (grp == 0 && e >= %d) {" % eoff)
        # a NaN or infinite intermediate reads as infinite (always redone)
        for f in acc_f:
            o("    { const float v = (float)fabs(acc%d); mx = fmaxf(mx, v < 3.0e38f ? v : INFINITY); }" % cidx(f))
        o("}")
        o("for (int s = 1; s < G; s <<= 1) mx = fmaxf(mx, intc_shflf(mx, (e + s < G ? lane + s : lane) * 4));")
        o("if (grp == 0 && e == 0 && (double)mx * %.17g > a.fp32_tol) a.redo_q[atomicAdd(a.redo_n, 1)] = q;"
          % (fp32_err_c(cls) * 2.0 ** -24))
        o.ind -= 1
        o("}")
    o("#ifdef INTC_BODIES_ONLY")
    o("const long long qoff = a.q_off[q];")
    o("const unsigned long long qs = a.q_str[q];")
    o("#endif")
    o("const int sA = (int)(qs & 0xffff), sB = (int)((qs >> 16) & 0xffff);")
    o("const int sC = (int)((qs >> 32) & 0xffff), sD = (int)((qs >> 48) & 0xffff);")
    o("(void)sA; (void)sB;")
    # ket HRR + ket sph by group 0 lanes with |e| >= LA
    o("if (grp == 0 && e >= %d) {" % eoff)
    o.ind = 3
    # the lane's LDS row: an offset opaque to the compiler, so that the row's
    # addresses are formed here from one base (non-negative immediate offsets)
    # rather than hoisted out of the quartet loop as NCD live registers
    o("int trow_off = (e - %d) * NCD;" % eoff)
    o('asm volatile("" : "+v"(trow_off));')
    o("double *trow = tlds + trow_off;")
    if LD > 0:
        o("const double CDx = a.pair_AB[3 * pk], CDy = a.pair_AB[3 * pk + 1], CDz = a.pair_AB[3 * pk + 2];")
    hl, H = _hrr_lines(lambda f: "acc%d" % cidx(f), LC, LD, "CD", "hk", "ket")
    for ln in hl:
        o(ln)
    cc, dc = cart_components(LC), cart_components(LD)
    if spherical:
        o("if constexpr (BS) {")
        o.ind += 1
        # d first: K1(c, sd) = sum_d H(c,d) S_LD[d, sd], nonzero terms only
        for ci, c in enumerate(cc):
            for sd in range(nsph(LD)):
                terms = _sparse_terms([H(c, d) for d in dc], SPH_BUILTIN[LD][:, sd])
                o("const double k1_%d_%d = %s;" % (ci, sd, terms))
        for sc in range(nsph(LC)):
            for sd in range(nsph(LD)):
                terms = _sparse_terms(["k1_%d_%d" % (ci, sd) for ci in range(len(cc))], SPH_BUILTIN[LC][:, sc])
                o("trow[%d] = %s;" % (sc * nD + sd, terms))
        o.ind -= 1
        o("} else {")
        o.ind += 1
        sC, sD = "a.sph + %d" % intc_sph_off(LC), "a.sph + %d" % intc_sph_off(LD)
        o("const double *SC = %s, *SD = %s;" % (sC, sD))
        # d first: K1(c, sd) = sum_d H(c,d) SD[d + sd*ncc(LD)]
        for ci, c in enumerate(cc):
            for sd in range(nsph(LD)):
                terms = " + ".join("%s * SD[%d]" % (H(c, d), di + sd * ncc(LD)) for di, d in enumerate(dc))
                o("const double k1_%d_%d = %s;" % (ci, sd, terms))
        for sc in range(nsph(LC)):
            for sd in range(nsph(LD)):
                terms = " + ".join("SC[%d] * k1_%d_%d" % (ci + sc * ncc(LC), ci, sd) for ci in range(len(cc)))
                o("trow[%d] = %s;" % (sc * nD + sd, terms))
        o.ind -= 1
        o("}")
    else:
        for ci, c in enumerate(cc):
            for di, d in enumerate(dc):
                o("trow[%d] = %s;" % (ci * nD + di, H(c, d)))
    o.ind = 2
    o("}")
    o('__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");')
    o("__builtin_amdgcn_wave_barrier();")
    o('__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");')
    # column phase: bra HRR + bra sph
    o("for (int j = tl; j < NCD; j += TL) {")
    o.ind = 3
    o("const int jc = j / %d, jd = j - jc * %d;" % (nD, nD))
    o("double *op = a.out + qoff + (long long)jc * sC + (long long)jd * sD;")
    bra_keep = [ec for ec in bra_comps if sum(ec) >= LA]
    for ec in bra_keep:
        o("const double x%d = tlds[%d * NCD + j];" % (cidx(ec), cidx(ec) - eoff))
    if LB > 0:
        o("const double ABx = a.pair_AB[3 * pb], ABy = a.pair_AB[3 * pb + 1], ABz = a.pair_AB[3 * pb + 2];")
    hl, HB = _hrr_lines(lambda ec: "x%d" % cidx(ec), LA, LB, "AB", "hb", "bra")
    for ln in hl:
        o(ln)
    ac, bc = cart_components(LA), cart_components(LB)
    if spherical:
        o("if constexpr (BS) {")
        o.ind += 1
        for ai, aa in enumerate(ac):
            for sb in range(nsph(LB)):
                terms = _sparse_terms([HB(aa, bb) for bb in bc], SPH_BUILTIN[LB][:, sb])
                o("const double m1_%d_%d = %s;" % (ai, sb, terms))
        for sa in range(nsph(LA)):
            for sb in range(nsph(LB)):
                terms = _sparse_terms(["m1_%d_%d" % (ai, sb) for ai in range(len(ac))], SPH_BUILTIN[LA][:, sa])
                o("op[%d * sA + %d * sB] = %s;" % (sa, sb, terms))
        o.ind -= 1
        o("} else {")
        o.ind += 1
        o("const double *SA = a.sph + %d, *SB = a.sph + %d;" % (intc_sph_off(LA), intc_sph_off(LB)))
        for ai, aa in enumerate(ac):
            for sb in range(nsph(LB)):
                terms = " + ".join("%s * SB[%d]" % (HB(aa, bb), bi + sb * ncc(LB)) for bi, bb in enumerate(bc))
                o("const double m1_%d_%d = %s;" % (ai, sb, terms))
        for sa in range(nsph(LA)):
            for sb in range(nsph(LB)):
                terms = " + ".join("SA[%d] * m1_%d_%d" % (ai + sa * ncc(LA), ai, sb) for ai in range(len(ac)))
                o("op[%d * sA + %d * sB] = %s;" % (sa, sb, terms))
        o.ind -= 1
        o("}")
    else:
        for ai, aa in enumerate(ac):
            for bi, bb in enumerate(bc):
                o("op[%d * sA + %d * sB] = %s;" % (ai, bi, HB(aa, bb)))
    o.ind = 2
    o("}")
    o('__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");')
    o("__builtin_amdgcn_wave_barrier();")
    o.ind = 1
    o("}")
    o.ind = 0
    o("}")
    o("")
    o("#ifndef INTC_BODIES_ONLY")
    # one-class grid: wave w of block b starts at team index (b * waves + w) * TPW
    # and strides over every team of the grid
    o("template <int BM, int BS>")
    o("__device__ __forceinline__ void %s_grid(const IntcEriArgs &a)" % name)
    o("{")
    o("    extern __shared__ double lds[];")
    o("    const int tpw = 64 / (%d * a.team_groups), wave = threadIdx.x >> 6;" % G)
    o("    %s_body<BM, BS>(a, lds + (size_t)wave * %s_reg(a.team_groups), (blockIdx.x * (blockDim.x >> 6) + wave)"
      " * tpw, gridDim.x * (blockDim.x >> 6) * tpw, a.nq);" % (name, name))
    o("}")
    # waves_per_eu: occupancy target for the Chebyshev-mode kernel (the compiler
    # then caps its registers at 512 / waves VGPRs)
    occ = " __attribute__((amdgpu_waves_per_eu(%d)))" % waves_per_eu if waves_per_eu else ""
    o('extern "C" __global__ void __launch_bounds__(256)%s %s(const IntcEriArgs a) { %s_grid<0, 0>(a); }'
      % (occ, name, name))
    o('extern "C" __global__ void __launch_bounds__(256) %s_refboys(const IntcEriArgs a) { %s_grid<1, 0>(a); }'
      % (name, name))
    if spherical:
        o('extern "C" __global__ void __launch_bounds__(256)%s %s_bs(const IntcEriArgs a) { %s_grid<0, 1>(a); }'
          % (occ, name, name))
    o('extern "C" hipError_t %s_launch(const IntcEriArgs *a, int grid, int block, size_t lds, hipStream_t st)'
      % name)
    o("{")
    o("    lds = (size_t)(block / 64) * %s_reg(a->team_groups) * sizeof(double);" % name)
    if spherical:
        o("    if (a->boys_mode == 0 && a->sph_builtin)")
        o("        hipLaunchKernelGGL(%s_bs, dim3(grid), dim3(block), lds, st, *a);" % name)
        o("    else if (a->boys_mode == 0)")
    else:
        o("    if (a->boys_mode == 0)")
    o("        hipLaunchKernelGGL(%s, dim3(grid), dim3(block), lds, st, *a);" % name)
    o("    else")
    o("        hipLaunchKernelGGL(%s_refboys, dim3(grid), dim3(block), lds, st, *a);" % name)
    o("    return hipGetLastError();")
    o("}")
    o("#endif  // INTC_BODIES_ONLY")
    meta = dict(cls=cls, spherical=spherical, G=G, NE=NE, NCD=NCD, M=M, NS=NS, NSF=NSF, name=name, prec=prec)
    return o.text() + "\n", meta


def _sparse_terms(xs, col):
    """sum_i xs[i] * col[i] over the nonzero col[i], coefficients as exact
    double literals (1.0 and -1.0 become plain terms)."""
    out = []
    for x, v in zip(xs, col):
        v = float(v)
        if v == 0.0:
            continue
        if v == 1.0:
            out.append(x)
        elif v == -1.0:
            out.append("-" + x)
        else:
            out.append("%s * %s" % (x, repr(v)))
    return " + ".join(out) if out else "0.0"


def intc_sph_off(l):
    return sum(ncc(k) * (2 * k + 1) for k in range(l))


def class_set(lmax_pair_side=2, extra=()):
    """Canonical classes for shells up to lmax, plus extra (la,lb,lc,ld)."""
    out = set(canonical_classes(lmax_pair_side))
    for c in extra:
        out.add(canonical(*c)[0])
    return sorted(out)


def wg_classes(lmax=3, covered=(), extra=((4, 0, 4, 0),)):
    """Canonical 4-centre classes with shells up to lmax that have no register
    kernel (an f shell makes the bra group or the per-lane ket table too
    large): they run on the workgroup-per-quartet kernels.  extra: further
    classes for those kernels, by default (g s|g s), the Schwarz diagonal of
    a g auxiliary shell with the unit s-shell."""
    cov = set(covered)
    out = [c for c in canonical_classes(lmax) if c not in cov]
    for c in extra:
        c = canonical(*c)[0]
        if c not in cov and c not in out:
            out.append(c)
    return out


def _nout(l, sph):
    return 2 * l + 1 if sph else (l + 1) * (l + 2) // 2


# ---------------------------------------------------------------------------
# one-electron integrals: overlap (a|b) and nuclear attraction (a|C|b)
# ---------------------------------------------------------------------------
ONEE_LMAX = 3
ONEE_KINDS = ("o", "v")   # overlap, nuclear attraction (point charge at C)


def onee_kernel_name(kind, la, lb, spherical):
    return "intc_onee_%s%d%d_%s" % (kind, la, lb, "s" if spherical else "c")


def emit_onee_kernel(o, kind, la, lb, spherical):
    """One thread per contracted item.  Obara-Saika VRR on a, (e|0)^(m) for
    |e| <= la + lb in registers, with the reference's direction rule (first
    non-zero component); contraction into la <= |e| accumulators; HGP HRR a -> b
    on the contracted values; optional Cartesian -> spherical step with the
    S_l of a.sph.  Overlap: base (pi/zeta)^(3/2) exp(-mu R^2); nuclear
    attraction: base 2 pi/zeta exp(-mu R^2) F_m(zeta |P-C|^2) and the
    -(P-C)_i (e-1_i)^(m+1) terms (the recurrences of the reference's example
    script, examples/example_input.py:79-110, with the VRR on a and the
    transfer to b by the HRR)."""
    L = la + lb
    nuc = kind == "v"
    M = L if nuc else 0
    name = onee_kernel_name(kind, la, lb, spherical)
    o("template <int BM>")
    o("__device__ __forceinline__ void %s_body(const IntcOneArgs &a)" % name)
    o("{")
    o.ind = 1
    o("constexpr int M = %d;" % M)
    o("(void)M;")
    o("for (int it = blockIdx.x * blockDim.x + threadIdx.x; it < a.n; it += gridDim.x * blockDim.x) {")
    o.ind = 2
    o("const int sa = a.it_a[it], sb = a.it_b[it];")
    o("const double Ax = a.shell_cen[3 * sa], Ay = a.shell_cen[3 * sa + 1], Az = a.shell_cen[3 * sa + 2];")
    o("const double Bx = a.shell_cen[3 * sb], By = a.shell_cen[3 * sb + 1], Bz = a.shell_cen[3 * sb + 2];")
    if nuc:
        o("const int ic = a.it_c[it];")
        o("const double Cx = a.centres[3 * ic], Cy = a.centres[3 * ic + 1], Cz = a.centres[3 * ic + 2];")
    o("const int oa = a.shell_off[sa], na = a.shell_np[sa], ob = a.shell_off[sb], nb = a.shell_np[sb];")
    o("const double ABx = Ax - Bx, ABy = Ay - By, ABz = Az - Bz;")
    o("const double R2 = ABx * ABx + ABy * ABy + ABz * ABz;")
    o("(void)ABx; (void)ABy; (void)ABz;")
    acc = [e for e in comps_upto(L) if sum(e) >= la]
    for e in acc:
        o("double acc%d = 0.0;" % cidx(e))
    o("for (int i = 0; i < na; ++i) {")
    o.ind = 3
    o("const double al = a.exps[oa + i], ca = a.coefs[oa + i];")
    o("for (int j = 0; j < nb; ++j) {")
    o.ind = 4
    o("const double be = a.exps[ob + j], cb = a.coefs[ob + j];")
    o("const double z = al + be, oz = 1.0 / z, o2z = 0.5 * oz;")
    o("(void)o2z;")
    o("const double Px = (al * Ax + be * Bx) * oz, Py = (al * Ay + be * By) * oz, Pz = (al * Az + be * Bz) * oz;")
    o("const double PAx = Px - Ax, PAy = Py - Ay, PAz = Pz - Az;")
    o("(void)PAx; (void)PAy; (void)PAz;")
    o("const double kab = ca * cb * exp(-al * be * oz * R2);")
    if nuc:
        o("const double PCx = Px - Cx, PCy = Py - Cy, PCz = Pz - Cz;")
        o("(void)PCx; (void)PCy; (void)PCz;")
        o("const double Tb = z * (PCx * PCx + PCy * PCy + PCz * PCz);")
        o("double F[M + 1];")
        o("intc_boys<M, BM>(Tb, F, a);")
        o("const double pref = kab * (2.0 * 3.141592653589793238) * oz;")
        for m in range(M + 1):
            o("const double v0_%d = pref * F[%d];" % (m, m))
    else:
        # (pi / zeta)^(3/2)
        o("const double v0_0 = kab * 5.568327996831707845 * oz * sqrt(oz);")
    for n in range(1, L + 1):
        for e in cart_components(n):
            i = first_nonzero(e)
            e1 = dec(e, i)
            k = e1[i]
            ax = "xyz"[i]
            for m in (range(M - n + 1) if nuc else range(1)):
                if nuc:
                    expr = "PA%s * v%d_%d - PC%s * v%d_%d" % (ax, cidx(e1), m, ax, cidx(e1), m + 1)
                else:
                    expr = "PA%s * v%d_%d" % (ax, cidx(e1), m)
                if k > 0:
                    e2 = dec(e1, i)
                    if nuc:
                        expr += " + %d.0 * o2z * (v%d_%d - v%d_%d)" % (k, cidx(e2), m, cidx(e2), m + 1)
                    else:
                        expr += " + %d.0 * o2z * v%d_%d" % (k, cidx(e2), m)
                o("const double v%d_%d = %s;" % (cidx(e), m, expr))
    for e in acc:
        o("acc%d += v%d_0;" % (cidx(e), cidx(e)))
    o.ind = 3
    o("}")
    o.ind = 2
    o("}")
    hl, H = _hrr_lines(lambda e: "acc%d" % cidx(e), la, lb, "AB", "h", "bra")
    for ln in hl:
        o(ln)
    o("double *op = a.out + a.it_off[it];")
    o("const unsigned long long st = a.it_str[it];")
    o("const long long s_a = (long long)(st & 0xffffffffull), s_b = (long long)(st >> 32);")
    ac, bc = cart_components(la), cart_components(lb)
    if spherical:
        o("const double *SA = a.sph + %d, *SB = a.sph + %d;" % (intc_sph_off(la), intc_sph_off(lb)))
        for ai, aa in enumerate(ac):
            for sbb in range(nsph(lb)):
                terms = " + ".join("%s * SB[%d]" % (H(aa, bb), bi + sbb * ncc(lb)) for bi, bb in enumerate(bc))
                o("const double m1_%d_%d = %s;" % (ai, sbb, terms))
        for saa in range(nsph(la)):
            for sbb in range(nsph(lb)):
                terms = " + ".join("SA[%d] * m1_%d_%d" % (ai + saa * ncc(la), ai, sbb) for ai in range(len(ac)))
                o("op[%d * s_a + %d * s_b] = %s;" % (saa, sbb, terms))
    else:
        for ai, aa in enumerate(ac):
            for bi, bb in enumerate(bc):
                o("op[%d * s_a + %d * s_b] = %s;" % (ai, bi, H(aa, bb)))
    o.ind = 1
    o("}")
    o.ind = 0
    o("}")
    o('extern "C" __global__ void __launch_bounds__(256) %s(const IntcOneArgs a) { %s_body<0>(a); }' % (name, name))
    if nuc:
        o('extern "C" __global__ void __launch_bounds__(256) %s_refboys(const IntcOneArgs a) { %s_body<1>(a); }'
          % (name, name))
    o("static hipError_t %s_launch(const IntcOneArgs *a, int grid, hipStream_t st)" % name)
    o("{")
    if nuc:
        o("    if (a->boys_mode == 1)")
        o("        hipLaunchKernelGGL(%s_refboys, dim3(grid), dim3(256), 0, st, *a);" % name)
        o("    else")
        o("        hipLaunchKernelGGL(%s, dim3(grid), dim3(256), 0, st, *a);" % name)
    else:
        o("    hipLaunchKernelGGL(%s, dim3(grid), dim3(256), 0, st, *a);" % name)
    o("    return hipGetLastError();")
    o("}")
    o("")


def emit_onee_source(lmax=ONEE_LMAX):
    """All one-electron class kernels (both kinds, l <= lmax on each shell,
    Cartesian and spherical) and their dispatch table, in one file."""
    o = Emitter()
    o.ind = 0
    o("// Generated by intception_amd/codegen.py -- do not edit.")
    o("// One-electron class kernels: overlap (o) and nuclear attraction (v).")
    o('#include "../intc_common.h"')
    o("")
    entries = []
    for kind in ONEE_KINDS:
        for la in range(lmax + 1):
            for lb in range(lmax + 1):
                for sph in (False, True):
                    emit_onee_kernel(o, kind, la, lb, sph)
                    entries.append((kind, la, lb, sph))
    o("typedef hipError_t (*intc_one_launch_fn)(const IntcOneArgs *, int, hipStream_t);")
    o("struct IntcOneEntry { int kind, la, lb, sph; intc_one_launch_fn fn; };")
    o("static const IntcOneEntry intc_one_table[] = {")
    for kind, la, lb, sph in entries:
        o("    {%d, %d, %d, %d, %s_launch}," % (ONEE_KINDS.index(kind), la, lb, int(sph),
                                               onee_kernel_name(kind, la, lb, sph)))
    o("};")
    o("static const int intc_n_one = %d;" % len(entries))
    o("static const int intc_one_lmax = %d;" % lmax)
    return o.text() + "\n"


def multi_members(metas):
    """Class kernels compiled into the persistent multi-class kernel: every
    spherical register kernel, FP64 and FP32 recurrences."""
    return [m for m in metas if m["spherical"]]


def fp32_register_class(cls):
    """Whether a register class gets an FP32-recurrence kernel: shells up to
    l = 2 only.  The three-centre classes with an f or g auxiliary shell have
    none (one routed run of those kernels stopped responding and no cause was
    found); the runtime gives their FP32-routed quartets the FP64 kernel."""
    return max(cls) <= 2


def emit_multi_source(members):
    """The persistent multi-class kernel: waves take tiles (one bin, a range
    of its quartet positions) from a device counter and run that class's body
    on them with its team size, until the tiles run out."""
    L = ["// Generated by intception_amd/codegen.py -- do not edit.",
         "// Persistent multi-class kernel: every spherical register-class body (built-in",
         "// transforms, Chebyshev Boys) behind one tile queue.",
         "#define INTC_BODIES_ONLY"]
    for m in members:
        L.append('#include "%s.hip"' % m["name"])
    L += ["#undef INTC_BODIES_ONLY",
          "",
          "// tiles[t] = {bin, kind, q_lo, q_hi}: queue positions [q_lo, q_hi) of bin `bin`,",
          "// whose body is case `kind`; args[bin] holds that bin's kernel arguments.",
          "// `next` must be 0 at launch (intc_multi_reset); every wave stops once the",
          "// counter has passed n_tiles.  reg: LDS doubles per wave (max over the bins).",
          "// Two waves per SIMD: the largest bodies need up to 256 VGPRs (a few spill).",
          'extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) '
          "intc_eri_multi_s_bs(const IntcEriArgs *args, const int4 *tiles, int n_tiles, int *next, int reg)",
          "{",
          "    extern __shared__ double lds[];",
          "    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;",
          "    double *wlds = lds + (size_t)wave * reg;",
          "    for (;;) {",
          "        int t = 0;",
          "        if (lane == 0) t = atomicAdd(next, 1);",
          "        t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));",
          "        if (t >= n_tiles) break;",
          "        const int4 w = tiles[t];",
          "        const int bin = __builtin_amdgcn_readfirstlane(w.x), kind = __builtin_amdgcn_readfirstlane(w.y);",
          "        const int q_lo = __builtin_amdgcn_readfirstlane(w.z), q_hi = __builtin_amdgcn_readfirstlane(w.w);",
          "        const IntcEriArgs &a = args[bin];",
          "        switch (kind) {"]
    for k, m in enumerate(members):
        L.append("        case %d: %s_body<0, 1>(a, wlds, q_lo, 64 / (%d * a.team_groups), q_hi); break;"
                 % (k, m["name"], m["G"]))
    L += ["        default: break;",
          "        }",
          '        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");',
          "        __builtin_amdgcn_wave_barrier();",
          "    }",
          "}",
          "",
          'extern "C" __global__ void intc_multi_reset(int *next) { if (threadIdx.x == 0) *next = 0; }',
          "",
          "// LDS beyond 64 KB per block needs the attribute on this device",
          'extern "C" int intc_multi_prepare()',
          "{",
          "    return hipFuncSetAttribute((const void *)intc_eri_multi_s_bs, "
          "hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) == hipSuccess ? 0 : -1;",
          "}",
          "",
          "// counter reset (a kernel, so that graphs hold no memset node), then the persistent grid",
          'extern "C" hipError_t intc_multi_launch(const IntcEriArgs *args, const int4 *tiles, int n_tiles, '
          "int *next, int reg, int grid, hipStream_t st)",
          "{",
          "    hipLaunchKernelGGL(intc_multi_reset, dim3(1), dim3(64), 0, st, next);",
          "    hipError_t e = hipGetLastError();",
          "    if (e != hipSuccess) return e;",
          "    hipLaunchKernelGGL(intc_eri_multi_s_bs, dim3(grid), dim3(256), (size_t)4 * reg * sizeof(double), st, "
          "args, tiles, n_tiles, next, reg);",
          "    return hipGetLastError();",
          "}",
          ""]
    return "\n".join(L)


def generate(outdir, classes, sph_modes=(False, True), precs=("d", "f"), wg=(), occupancy=None, batch=None):
    os.makedirs(outdir, exist_ok=True)
    metas = []
    written = []
    for cls in classes:
        for sph, prec in itertools.product(sph_modes, precs):
            if prec == "f" and not fp32_register_class(cls):
                continue
            nm = kernel_name(cls, sph, prec)
            src, meta = emit_class_source(cls, sph, prec, (occupancy or {}).get(nm), (batch or {}).get(nm))
            fn = os.path.join(outdir, meta["name"] + ".hip")
            old = open(fn).read() if os.path.exists(fn) else None
            if old != src:
                with open(fn, "w") as fh:
                    fh.write(src)
            metas.append(meta)
            written.append(fn)
    # dispatch table
    lines = ["// Generated by intception_amd/codegen.py -- do not edit.", '#include "../intc_common.h"',
             "typedef hipError_t (*intc_launch_fn)(const IntcEriArgs *, int, int, size_t, hipStream_t);"]
    for m in metas:
        lines.append('extern "C" hipError_t %s_launch(const IntcEriArgs *, int, int, size_t, hipStream_t);'
                     % m["name"])
    for pr in precs:
        if wg:
            lines.append('extern "C" hipError_t intc_eri_wg_%s_launch(const IntcEriArgs *, int, int, size_t, '
                         'hipStream_t);' % pr)
    # G = 64 and one group per team for the workgroup kernels (wg = 1): one
    # quartet per 256-thread block, LDS sized by the launcher
    # NS: setup-slot length (LDS per wave is max(teams per wave * NE * NCD,
    # 64 * NS) doubles); multi: case of the class in intc_eri_multi_s_bs, -1
    # when the persistent kernel does not hold it
    lines.append("struct IntcClassEntry { int la, lb, lc, ld, sph, prec, G, NE, NCD, M, wg, NS, multi; "
                 "intc_launch_fn fn; };")
    lines.append("static const IntcClassEntry intc_class_table[] = {")
    members = multi_members(metas)
    kind = {m["name"]: k for k, m in enumerate(members)}
    n = 0
    for m in metas:
        la, lb, lc, ld = m["cls"]
        lines.append("    {%d, %d, %d, %d, %d, %d, %d, %d, %d, %d, 0, %d, %d, %s_launch}," % (
            la, lb, lc, ld, int(m["spherical"]), 0 if m["prec"] == "d" else 1, m["G"], m["NE"], m["NCD"], m["M"],
            m["NS"], kind.get(m["name"], -1), m["name"]))
        n += 1
    for cls in wg:
        la, lb, lc, ld = cls
        for sph, pr in itertools.product(sph_modes, precs):
            lines.append("    {%d, %d, %d, %d, %d, %d, 64, %d, %d, %d, 1, 0, -1, intc_eri_wg_%s_launch}," % (
                la, lb, lc, ld, int(sph), 0 if pr == "d" else 1, _nout(la, sph) * _nout(lb, sph),
                _nout(lc, sph) * _nout(ld, sph), la + lb + lc + ld, pr))
            n += 1
    lines.append("};")
    lines.append("static const int intc_n_class = %d;" % n)
    # built-in S_l, l = 0..7, column-major, concatenated (the a.sph layout)
    blob = np.concatenate([SPH_BUILTIN[l].reshape(-1, order="F") for l in range(8)])
    lines.append("static const double intc_builtin_sph[%d] = {" % len(blob))
    for i in range(0, len(blob), 4):
        lines.append("    " + ", ".join(repr(float(v)) for v in blob[i:i + 4]) + ",")
    lines.append("};")
    # the workgroup kernel's class list (its template instantiations)
    wsrc = "// Generated by intception_amd/codegen.py -- do not edit.\n" + "".join(
        "INTC_WG_CLASS(%d, %d, %d, %d)\n" % tuple(c) for c in wg)
    fn = os.path.join(outdir, "wg_classes.inc")
    old = open(fn).read() if os.path.exists(fn) else None
    if old != wsrc:
        with open(fn, "w") as fh:
            fh.write(wsrc)
    tab = os.path.join(outdir, "class_table.inc")
    src = "\n".join(lines) + "\n"
    old = open(tab).read() if os.path.exists(tab) else None
    if old != src:
        with open(tab, "w") as fh:
            fh.write(src)
    src = emit_onee_source()
    fn = os.path.join(outdir, "onee_kernels.inc")
    old = open(fn).read() if os.path.exists(fn) else None
    if old != src:
        with open(fn, "w") as fh:
            fh.write(src)
    if members:
        src = emit_multi_source(members)
        fn = os.path.join(outdir, "intc_eri_multi.hip")
        old = open(fn).read() if os.path.exists(fn) else None
        if old != src:
            with open(fn, "w") as fh:
                fh.write(src)
        written.append(fn)
    return written, metas
