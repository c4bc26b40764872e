"""Independent arbitrary-precision ERI, overlap and nuclear attraction
integrals (McMurchie-Davidson) with mpmath.

TEST INFRASTRUCTURE ONLY.  This is a different algorithm from the reference's
Obara-Saika/Head-Gordon-Pople scheme: Hermite expansion coefficients E_t^{ij}
and Hermite Coulomb integrals R_{tuv} (Helgaker, Jorgensen, Olsen, "Molecular
Electronic-Structure Theory", sections 9.5 and 9.9), evaluated at 40 digits
with the Boys function taken from the incomplete gamma function,
F_n(x) = gamma(n+1/2, x) / (2 x^(n+1/2)).  It pins the C restatement
(oracle/eri_oracle.c) on small classes: a dropped VRR term, a wrong index or a
transposed operand there changes the value by far more than the 1e-12 check.
"""
import itertools

import mpmath as mp

mp.mp.dps = 40


def boys(n, x):
    x = mp.mpf(x)
    if x == 0:
        return mp.mpf(1) / (2 * n + 1)
    return mp.gammainc(n + mp.mpf("0.5"), 0, x) / (2 * x ** (n + mp.mpf("0.5")))


def cart_components(l):
    """Cartesian components of shell l, x^l first ... z^l last."""
    out = []
    for i in range(l, -1, -1):
        for j in range(l - i, -1, -1):
            out.append((i, j, l - i - j))
    return out


def _hermite_E(i, j, a, b, XAB):
    """E_t^{ij} for t = 0..i+j (1D), including exp(-mu X_AB^2)."""
    p = a + b
    mu = a * b / p
    XPA = -b / p * XAB
    XPB = a / p * XAB
    E = {}
    E[(0, 0, 0)] = mp.e ** (-mu * XAB * XAB)

    def get(ii, jj, t):
        if t < 0 or t > ii + jj or ii < 0 or jj < 0:
            return mp.mpf(0)
        return E[(ii, jj, t)]

    for ii in range(i + 1):
        for jj in range(j + 1):
            if ii == 0 and jj == 0:
                continue
            for t in range(ii + jj + 1):
                if ii > 0:
                    v = get(ii - 1, jj, t - 1) / (2 * p) + XPA * get(ii - 1, jj, t) + (t + 1) * get(ii - 1, jj, t + 1)
                else:
                    v = get(ii, jj - 1, t - 1) / (2 * p) + XPB * get(ii, jj - 1, t) + (t + 1) * get(ii, jj - 1, t + 1)
                E[(ii, jj, t)] = v
    return [get(i, j, t) for t in range(i + j + 1)]


def _hermite_R(tmax, alpha, PQ):
    """R_{tuv}^{(0)} for t+u+v <= tmax."""
    X, Y, Z = PQ
    T = alpha * (X * X + Y * Y + Z * Z)
    Fn = [boys(n, T) for n in range(tmax + 1)]
    cache = {}

    def R(n, t, u, v):
        if t < 0 or u < 0 or v < 0:
            return mp.mpf(0)
        key = (n, t, u, v)
        if key in cache:
            return cache[key]
        if t == 0 and u == 0 and v == 0:
            r = (-2 * alpha) ** n * Fn[n]
        elif t > 0:
            r = (t - 1) * R(n + 1, t - 2, u, v) + X * R(n + 1, t - 1, u, v)
        elif u > 0:
            r = (u - 1) * R(n + 1, t, u - 2, v) + Y * R(n + 1, t, u - 1, v)
        else:
            r = (v - 1) * R(n + 1, t, u, v - 2) + Z * R(n + 1, t, u, v - 1)
        cache[key] = r
        return r

    return lambda t, u, v: R(0, t, u, v)


def primitive_eri(ca, cb, cc, cd, a, b, c, d, A, B, C, D):
    """(ab|cd) for unnormalised primitive Cartesian Gaussians.

    ca..cd: component tuples; a..d: exponents; A..D: centres."""
    A = [mp.mpf(v) for v in A]; B = [mp.mpf(v) for v in B]
    C = [mp.mpf(v) for v in C]; D = [mp.mpf(v) for v in D]
    a, b, c, d = (mp.mpf(v) for v in (a, b, c, d))
    p, q = a + b, c + d
    P = [(a * A[k] + b * B[k]) / p for k in range(3)]
    Q = [(c * C[k] + d * D[k]) / q for k in range(3)]
    alpha = p * q / (p + q)
    Eab = [_hermite_E(ca[k], cb[k], a, b, A[k] - B[k]) for k in range(3)]
    Ecd = [_hermite_E(cc[k], cd[k], c, d, C[k] - D[k]) for k in range(3)]
    R = _hermite_R(sum(ca) + sum(cb) + sum(cc) + sum(cd), alpha, [P[k] - Q[k] for k in range(3)])
    s = mp.mpf(0)
    for t, u, v in itertools.product(range(len(Eab[0])), range(len(Eab[1])), range(len(Eab[2]))):
        e1 = Eab[0][t] * Eab[1][u] * Eab[2][v]
        if e1 == 0:
            continue
        for tau, nu, phi in itertools.product(range(len(Ecd[0])), range(len(Ecd[1])), range(len(Ecd[2]))):
            e2 = Ecd[0][tau] * Ecd[1][nu] * Ecd[2][phi]
            s += e1 * e2 * (-1) ** (tau + nu + phi) * R(t + tau, u + nu, v + phi)
    return 2 * mp.pi ** mp.mpf(2.5) / (p * q * mp.sqrt(p + q)) * s


def primitive_overlap(ca, cb, a, b, A, B):
    """(a|b) for unnormalised primitive Cartesian Gaussians: per direction
    S_ij = E_0^{ij} sqrt(pi/p) (Helgaker et al. eq. 9.5.41)."""
    A = [mp.mpf(v) for v in A]; B = [mp.mpf(v) for v in B]
    a, b = mp.mpf(a), mp.mpf(b)
    p = a + b
    s = mp.mpf(1)
    for k in range(3):
        s *= _hermite_E(ca[k], cb[k], a, b, A[k] - B[k])[0] * mp.sqrt(mp.pi / p)
    return s


def primitive_nuclear(ca, cb, a, b, A, B, C):
    """(a|1/r_C|b) for unnormalised primitive Cartesian Gaussians:
    2 pi/p sum_tuv E_t E_u E_v R_tuv(p, P - C) (Helgaker et al. eq. 9.9.32)."""
    A = [mp.mpf(v) for v in A]; B = [mp.mpf(v) for v in B]; C = [mp.mpf(v) for v in C]
    a, b = mp.mpf(a), mp.mpf(b)
    p = a + b
    P = [(a * A[k] + b * B[k]) / p for k in range(3)]
    E = [_hermite_E(ca[k], cb[k], a, b, A[k] - B[k]) for k in range(3)]
    R = _hermite_R(sum(ca) + sum(cb), p, [P[k] - C[k] for k in range(3)])
    s = mp.mpf(0)
    for t, u, v in itertools.product(range(len(E[0])), range(len(E[1])), range(len(E[2]))):
        s += E[0][t] * E[1][u] * E[2][v] * R(t, u, v)
    return 2 * mp.pi / p * s


def contracted_eri(shells, spherical=False, sph=None):
    """Contracted quartet; shells as in oracle.eri_quartet (n_cont = 1).

    Returns a flat list (row-major [a][b][c][d]) of mpf values."""
    comps = [cart_components(s["l"]) for s in shells]
    out = {}
    prims = [list(zip(s["exps"], [float(v) for v in list(s["coefs"]) ])) for s in shells]
    for ka in comps[0]:
        for kb in comps[1]:
            for kc in comps[2]:
                for kd in comps[3]:
                    tot = mp.mpf(0)
                    for (ea, wa) in prims[0]:
                        for (eb, wb) in prims[1]:
                            for (ec, wc) in prims[2]:
                                for (ed, wd) in prims[3]:
                                    tot += (mp.mpf(wa) * wb * wc * wd *
                                            primitive_eri(ka, kb, kc, kd, ea, eb, ec, ed,
                                                          shells[0]["cen"], shells[1]["cen"],
                                                          shells[2]["cen"], shells[3]["cen"]))
                    out[(ka, kb, kc, kd)] = tot
    if not spherical:
        return [out[k] for k in itertools.product(*comps)]
    # apply S (ncc x nsph, column-major S[icart + isph*ncart]) on each index
    ns = [2 * s["l"] + 1 for s in shells]
    res = []
    for sa in range(ns[0]):
        for sb in range(ns[1]):
            for sc in range(ns[2]):
                for sd in range(ns[3]):
                    v = mp.mpf(0)
                    for ia, ka in enumerate(comps[0]):
                        wa = sph[shells[0]["l"]][ia][sa]
                        if wa == 0:
                            continue
                        for ib, kb in enumerate(comps[1]):
                            wb = sph[shells[1]["l"]][ib][sb]
                            if wb == 0:
                                continue
                            for ic, kc in enumerate(comps[2]):
                                wc = sph[shells[2]["l"]][ic][sc]
                                if wc == 0:
                                    continue
                                for idd, kd in enumerate(comps[3]):
                                    wd = sph[shells[3]["l"]][idd][sd]
                                    if wd == 0:
                                        continue
                                    v += mp.mpf(wa) * wb * wc * wd * out[(ka, kb, kc, kd)]
                    res.append(v)
    return res


def contracted_eri_elements(shells, elements, spherical=False, sph=None):
    """Selected elements (flat row-major [a][b][c][d] indices into the block)
    of a contracted quartet, same definition as contracted_eri but evaluated
    per primitive quartet with one Hermite R table and one set of E tables
    shared by every Cartesian component the elements need (for blocks too
    large to evaluate whole, e.g. (ff|ff) with several primitives)."""
    comps = [cart_components(s["l"]) for s in shells]
    ncomp = [(2 * s["l"] + 1) if spherical else len(c) for s, c in zip(shells, comps)]
    # Cartesian terms each element needs: (weight, (ia, ib, ic, id))
    need = {}
    plans = []
    for el in elements:
        idx = []
        r = int(el)
        for n in reversed(ncomp):
            idx.append(r % n)
            r //= n
        idx = idx[::-1]
        terms = []
        if spherical:
            lists = []
            for k in range(4):
                col = [(ic, sph[shells[k]["l"]][ic][idx[k]]) for ic in range(len(comps[k]))]
                lists.append([(ic, w) for ic, w in col if w != 0])
            for (ia, wa), (ib, wb), (ic, wc), (idd, wd) in itertools.product(*lists):
                terms.append((mp.mpf(wa) * wb * wc * wd, (ia, ib, ic, idd)))
        else:
            terms.append((mp.mpf(1), tuple(idx)))
        for _, key in terms:
            need[key] = mp.mpf(0)
        plans.append(terms)
    cen = [[mp.mpf(v) for v in s["cen"]] for s in shells]
    A, B, C, D = cen
    lsum = sum(s["l"] for s in shells)
    for (ea, wa), (eb, wb), (ec, wc), (ed, wd) in itertools.product(
            *[list(zip(s["exps"], s["coefs"])) for s in shells]):
        a, b, c, d = (mp.mpf(float(v)) for v in (ea, eb, ec, ed))
        p, q = a + b, c + d
        P = [(a * A[k] + b * B[k]) / p for k in range(3)]
        Q = [(c * C[k] + d * D[k]) / q for k in range(3)]
        alpha = p * q / (p + q)
        R = _hermite_R(lsum, alpha, [P[k] - Q[k] for k in range(3)])
        Eab = {}
        Ecd = {}
        w = mp.mpf(float(wa)) * mp.mpf(float(wb)) * mp.mpf(float(wc)) * mp.mpf(float(wd))
        pref = 2 * mp.pi ** mp.mpf(2.5) / (p * q * mp.sqrt(p + q)) * w
        for key in need:
            ka, kb, kc, kd = (comps[k][key[k]] for k in range(4))
            e1s = []
            for k in range(3):
                if (k, ka[k], kb[k]) not in Eab:
                    Eab[(k, ka[k], kb[k])] = _hermite_E(ka[k], kb[k], a, b, A[k] - B[k])
                if (k, kc[k], kd[k]) not in Ecd:
                    Ecd[(k, kc[k], kd[k])] = _hermite_E(kc[k], kd[k], c, d, C[k] - D[k])
                e1s.append((Eab[(k, ka[k], kb[k])], Ecd[(k, kc[k], kd[k])]))
            s = mp.mpf(0)
            (Ex, Fx), (Ey, Fy), (Ez, Fz) = e1s
            for t, u, v in itertools.product(range(len(Ex)), range(len(Ey)), range(len(Ez))):
                e1 = Ex[t] * Ey[u] * Ez[v]
                if e1 == 0:
                    continue
                for tau, nu, phi in itertools.product(range(len(Fx)), range(len(Fy)), range(len(Fz))):
                    e2 = Fx[tau] * Fy[nu] * Fz[phi]
                    s += e1 * e2 * (-1) ** (tau + nu + phi) * R(t + tau, u + nu, v + phi)
            need[key] += pref * s
    return [sum(wt * need[key] for wt, key in terms) for terms in plans]
