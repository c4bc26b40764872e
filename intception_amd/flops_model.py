"""Algorithmic work per contracted quartet (the roofline numerator).

The count is the minimal work of the reference's own scheme for the ERI input
script (VRR a -> c -> d per primitive quartet with the m-group bounds of
src/intception/generator/integral_wrapper.py:1153-1254, contraction, one HRR
a -> b on contracted integrals, four Cartesian->spherical transforms), with
FMA = 2 flops, +,-,x,/ = 1, exp and sqrt = 1 each:

    P * [40 (pair and quartet scalars) + (14 + 8 + 4M) (Boys: Taylor, exp,
         recursion) + (M + 1) (base) + V]
    + 2 P [ncc0(L) - ncc0(la-1)] ncc(lc) ncc(ld)          (contraction)
    + sum_{l=1..lb} 2 ncc(l) [ncc0(L-l) - ncc0(la-1)] n_c n_d   (HRR)
    + 2 sum(rows * n_cart * n_sph) over the four transforms

V = 2 x (number of surviving OS terms), each term one FMA, with the
direction rule "first non-zero component of the target"
(src/intception/generator/generator.py:547-558).  Bytes per quartet =
8 x (integrals written) + 16 (quartet descriptor).

``hgp_flops`` counts the work of the device kernels' own schedule (bra chain
per lane, ket OS recursion, ket and bra HRR, transforms) for comparison.
"""
from functools import lru_cache

import numpy as np

from .basis import cart_components, ncc, ncc0, nsph


def _comps_upto(l):
    out = []
    for k in range(l + 1):
        out += cart_components(k)
    return out


def _fnz(c):
    for i in range(3):
        if c[i] > 0:
            return i
    return 0


@lru_cache(maxsize=None)
def vrr_terms(la, lb, lc, ld):
    """Surviving OS terms per primitive quartet in the reference scheme."""
    L, M = la + lb, la + lb + lc + ld
    ca, cc, cd = _comps_upto(L), _comps_upto(lc), _comps_upto(ld)
    n = 0
    for m in range(M, -1, -1):
        amax = min(L, M - m)
        for a in ca:
            if 0 < sum(a) <= amax:
                i = _fnz(a)
                n += 2 + (2 if a[i] >= 2 else 0)
        cmax = min(lc, lc + ld - m)
        for c in cc:
            if 0 < sum(c) <= cmax:
                i = _fnz(c)
                for a in ca:
                    n += 2 + (2 if c[i] >= 2 else 0) + (1 if a[i] > 0 else 0)
        dmax = ld - m
        for d in cd:
            if 0 < sum(d) <= dmax:
                i = _fnz(d)
                for a in ca:
                    for c in cc:
                        n += 2 + (2 if d[i] >= 2 else 0) + (2 if c[i] > 0 else 0) + (1 if a[i] > 0 else 0)
    return n


@lru_cache(maxsize=None)
def _per_prim(la, lb, lc, ld):
    L, M = la + lb, la + lb + lc + ld
    V = 2 * vrr_terms(la, lb, lc, ld)
    return 40 + (14 + 8 + 4 * M) + (M + 1) + V


@lru_cache(maxsize=None)
def _per_contracted(la, lb, lc, ld, spherical):
    L = la + lb
    nc = nsph(lc) if spherical else ncc(lc)
    nd = nsph(ld) if spherical else ncc(ld)
    contract_row = 2 * (ncc0(L) - ncc0(la - 1)) * ncc(lc) * ncc(ld)
    hrr = 0
    for l in range(1, lb + 1):
        hrr += 2 * ncc(l) * (ncc0(L - l) - ncc0(la - 1)) * nc * nd
    tr = 0
    if spherical:
        # early c, d on [a in la..L] x c x d rows, late a, b
        ne = ncc0(L) - ncc0(la - 1)
        tr += 2 * ne * ncc(lc) * ncc(ld) * nsph(ld)          # d
        tr += 2 * ne * ncc(lc) * nsph(ld) * nsph(lc)          # c
        tr += 2 * ncc(la) * ncc(lb) * nc * nd * nsph(lb)       # b
        tr += 2 * ncc(la) * nsph(lb) * nc * nd * nsph(la)      # a
    return contract_row, hrr + tr


def quartet_flops(ls, nprims, spherical):
    """Reference-scheme flops of one contracted quartet."""
    la, lb, lc, ld = (int(v) for v in ls)
    P = int(np.prod(nprims))
    crow, fixed = _per_contracted(la, lb, lc, ld, bool(spherical))
    return P * (_per_prim(la, lb, lc, ld) + crow) + fixed


def batch_flops(batch, idx=None):
    """Vector of reference-scheme flops per quartet of a QuartetBatch."""
    q = batch.quartets if idx is None else batch.quartets[idx]
    ls = batch.shell_l[q]
    P = batch.shell_nprim[q].astype(np.int64).prod(axis=1)
    out = np.empty(len(q), dtype=np.float64)
    keys = ls[:, 0] * 1000 + ls[:, 1] * 100 + ls[:, 2] * 10 + ls[:, 3]
    for k in np.unique(keys):
        sel = keys == k
        la, lb, lc, ld = k // 1000, (k // 100) % 10, (k // 10) % 10, k % 10
        crow, fixed = _per_contracted(la, lb, lc, ld, bool(batch.spherical))
        out[sel] = P[sel] * (_per_prim(la, lb, lc, ld) + crow) + fixed
    return out


def batch_bytes(batch, idx=None):
    sz = batch.sizes() if idx is None else batch.sizes()[idx]
    return 8.0 * sz + 16.0


# ---------------------------------------------------------------------------
# one-electron kinds (overlap, nuclear attraction)
# ---------------------------------------------------------------------------
@lru_cache(maxsize=None)
def _onee_parts(kind, la, lb, spherical):
    """(per primitive pair, per contracted item) flops of the one-electron
    scheme: VRR (e|0)^(m) on a for |e| <= la + lb with the direction rule
    (nuclear attraction: PA V^m - PC V^(m+1) [+ k/2zeta (V^m - V^(m+1))];
    overlap: PA V [+ k/2zeta V]), accumulation of la <= |e|, then the HRR to b
    and the transforms on the contracted block.  Setup per primitive pair:
    zeta, P, PA, the Gaussian-product prefactor (30); Boys as for the ERI
    (14 + 8 + 4M) and M + 1 prefactor products for the nuclear attraction."""
    L = la + lb
    nuc = kind == 1
    M = L if nuc else 0
    vrr = 0
    for n in range(1, L + 1):
        for e in cart_components(n):
            i = _fnz(e)
            k = e[i] - 1
            for _ in (range(M - n + 1) if nuc else range(1)):
                vrr += (3 if nuc else 1) + ((3 if nuc else 2) if k > 0 else 0)
    setup = 30 + ((14 + 8 + 4 * M) + (M + 1) if nuc else 4)
    acc = ncc0(L) - ncc0(la - 1)
    hrr = sum(2 * ncc(bl) * (ncc0(L - bl) - ncc0(la - 1)) for bl in range(1, lb + 1))
    if spherical:
        tr = 2 * ncc(la) * ncc(lb) * nsph(lb) + 2 * ncc(la) * nsph(la) * nsph(lb)
    else:
        tr = 0
    return setup + vrr + acc, hrr + tr


def onee_flops(kind, la, lb, nprim_pairs, spherical):
    """Algorithmic flops of one contracted one-electron item."""
    p, c = _onee_parts(int(kind), int(la), int(lb), bool(spherical))
    return nprim_pairs * p + c
