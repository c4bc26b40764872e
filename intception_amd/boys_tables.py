"""Generate the Boys-function tables used by the HIP kernels.

Two tables, both computed here with mpmath from the definition
F_n(x) = gamma(n + 1/2, x) / (2 x^(n + 1/2)) (nothing is taken from the
reference's data files):

* Chebyshev table (the default device path).  For every m_max = 0..MMAX and
  every unit interval [k, k+1), k = 0..NINT-1, the degree-DEG Chebyshev
  interpolant of F_{m_max} on that interval, rewritten as a polynomial in
  t = 2 (x - k) - 1 in [-1, 1] (Horner on device).  Lower orders come from
  downward recursion F_m = (2 x F_{m+1} + e^-x) / (2m + 1); for x >= NINT the
  device uses F_0 = sqrt(pi / x) / 2 and upward recursion with the e^-x term.

* Reference-grid tables, for the device's reference-algorithm mode (the
  three-branch scheme of
  src/intception/generator/callables/generator_boys_function.py:98-205):
  F_n(k/10) for k = 0..150, n = 0..32 and F_n(15 + k/10) for k = 0..200,
  n = 0..6, the same grids as the reference's own tables.

``write_inc(path)`` writes both as a C include file.
"""
import os

import mpmath as mp

MMAX = 16      # highest F_m the Chebyshev table serves (config 2: M = 12)
NINT = 36      # unit intervals on [0, 36)
DEG = 12       # polynomial degree per interval


def boys_mp(n, x):
    x = mp.mpf(x)
    if x == 0:
        return mp.mpf(1) / (2 * n + 1)
    return mp.gammainc(n + mp.mpf("0.5"), 0, x) / (2 * x ** (n + mp.mpf("0.5")))


def cheb_interval_poly(m, k, deg=DEG):
    """Coefficients a_0..a_deg with F_m(k + (t+1)/2) ~ sum a_j t^j."""
    mp.mp.dps = 40
    n = deg + 1
    nodes = [mp.cos(mp.pi * (j + mp.mpf("0.5")) / n) for j in range(n)]
    vals = [boys_mp(m, k + (t + 1) / 2) for t in nodes]
    # Chebyshev coefficients
    c = []
    for i in range(n):
        s = mp.fsum(vals[j] * mp.cos(mp.pi * i * (j + mp.mpf("0.5")) / n) for j in range(n))
        c.append(s * (1 if i == 0 else 2) / n)
    # convert sum c_i T_i(t) to monomials in t
    Tprev, Tcur = [mp.mpf(1)], [mp.mpf(0), mp.mpf(1)]
    poly = [mp.mpf(0)] * n
    poly[0] += c[0]
    if n > 1:
        poly[1] += c[1]
    for i in range(2, n):
        Tnext = [mp.mpf(0)] * (i + 1)
        for j, v in enumerate(Tcur):
            Tnext[j + 1] += 2 * v
        for j, v in enumerate(Tprev):
            Tnext[j] -= v
        for j, v in enumerate(Tnext):
            poly[j] += c[i] * v
        Tprev, Tcur = Tcur, Tnext
    return poly


def cheb_table():
    out = []
    for m in range(MMAX + 1):
        for k in range(NINT):
            out.append([float(v) for v in cheb_interval_poly(m, k)])
    return out


def grid_tables():
    mp.mp.dps = 40
    small = [[float(boys_mp(n, mp.mpf(k) / 10)) for n in range(33)] for k in range(151)]
    medium = [[float(boys_mp(n, 15 + mp.mpf(k) / 10)) for n in range(7)] for k in range(201)]
    return small, medium


def _fmt(rows):
    return ",\n".join("  " + ", ".join(repr(v) for v in r) for r in rows)


def write_inc(path):
    cheb = cheb_table()
    small, medium = grid_tables()
    with open(path, "w") as fh:
        fh.write("/* Generated by intception_amd/boys_tables.py -- do not edit. */\n")
        fh.write(f"#define INTC_BOYS_MMAX {MMAX}\n#define INTC_BOYS_NINT {NINT}\n"
                 f"#define INTC_BOYS_DEG {DEG}\n")
        fh.write(f"static const double intc_boys_cheb_host[{len(cheb) * (DEG + 1)}] = {{\n{_fmt(cheb)}\n}};\n")
        fh.write(f"static const double intc_boys_smallx_host[{151 * 33}] = {{\n{_fmt(small)}\n}};\n")
        fh.write(f"static const double intc_boys_mediumx_host[{201 * 7}] = {{\n{_fmt(medium)}\n}};\n")


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    write_inc(os.path.join(here, "csrc", "boys_tables.inc"))
