/*
 * onee_oracle.c -- CPU restatement of the one-electron integrals of the
 * reference's example input script (examples/example_input.py:79-110):
 * the 2-index overlap (a|b) and nuclear attraction (a|C|b), primitive
 * (the script's contracted=False) and contracted, Cartesian or spherical.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (intception_amd) never links, loads or calls it.
 *
 * The recurrences are the script's, applied in its order through the
 * reference's VRR-only algorithms (src/intception/generator/callables/
 * main_source_algorithms/primitive_vrr_only_no_auxiliary.py and
 * primitive_vrr_only_with_auxiliary.py: the VRRs build their indices in the
 * order they are added, so the last-added VRR is the outermost step), with the
 * direction rule "first non-zero component of the target"
 * (src/intception/generator/generator.py:547-558):
 *   overlap (example_input.py:79-88): base sqrt(pi^3 / zeta^3) exp(-xa xb/zeta |A-B|^2);
 *     vrr1 on a: (a|b) = PA_i (a-1|b) + a_i'/2zeta (a-2|b) + b_i/2zeta (a-1|b-1)
 *     vrr2 on b: (a|b) = PB_i (a|b-1) + b_i'/2zeta (a|b-2) + a_i/2zeta (a-1|b-1)
 *     (b built last);
 *   nuclear attraction (example_input.py:90-110): base 2 pi/zeta exp(-xa xb/zeta |A-B|^2)
 *     F_m(zeta |P-C|^2);
 *     vrr1 on b: (a|b)^m = PB_i (a|b-1)^m - PC_i (a|b-1)^(m+1)
 *                + b_i'/2zeta [(a|b-2)^m - (a|b-2)^(m+1)] + a_i/2zeta [(a-1|b-1)^m - (a-1|b-1)^(m+1)]
 *     vrr2 on a: the same with a and b exchanged (a built last);
 * where x_i' is the target's component minus one.  Contraction, the
 * spherical transform (out = S_a^T cart S_b, S[icart + isph * ncc]) and the
 * output layout follow contracted_general.py as for the ERI.
 *
 * The device path builds (e|0) by the VRR on a only and moves angular
 * momentum to b with the HRR on contracted values, so the two agree to
 * rounding, not bit for bit.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "eri_oracle.h"

#define OE_LMAX 3
#define OE_N 20                  /* components with |c| <= OE_LMAX */
#define OE_M (2 * OE_LMAX + 1)

static int oe_ncc(int l) { return (l + 1) * (l + 2) / 2; }
static int oe_ncc0(int l) { return l < 0 ? 0 : (l + 1) * (l + 2) * (l + 3) / 6; }

static int oe_cidx(const int c[3])
{
    int l = c[0] + c[1] + c[2], yz = c[1] + c[2];
    return oe_ncc0(l - 1) + yz * (yz + 1) / 2 + c[2];
}

/* k-th Cartesian component of shell l in the reference order (x^l first) */
static void oe_comp(int l, int k, int c[3])
{
    int n = 0;
    for (int yz = 0; yz <= l; ++yz)
        for (int z = 0; z <= yz; ++z, ++n)
            if (n == k) {
                c[0] = l - yz;
                c[1] = yz - z;
                c[2] = z;
                return;
            }
}

static int oe_first(const int c[3]) { return c[0] > 0 ? 0 : (c[1] > 0 ? 1 : 2); }

typedef struct {
    int kind;
    double PA[3], PB[3], PC[3], o2z, base;
    double F[OE_M];
    double memo[OE_N][OE_N][OE_M];
    unsigned char have[OE_N][OE_N][OE_M];
} oe_ctx;

static double oe_int(oe_ctx *c, const int a[3], const int b[3], int m)
{
    for (int i = 0; i < 3; ++i)
        if (a[i] < 0 || b[i] < 0) return 0.0;
    const int ia = oe_cidx(a), ib = oe_cidx(b);
    if (c->have[ia][ib][m]) return c->memo[ia][ib][m];
    const int la = a[0] + a[1] + a[2], lb = b[0] + b[1] + b[2];
    double v;
    int a1[3], b1[3], a2[3], b2[3];
    memcpy(a1, a, sizeof a1);
    memcpy(b1, b, sizeof b1);
    if (c->kind == 0) {
        if (lb > 0) {           /* vrr2 on b (added last, outermost) */
            int i = oe_first(b);
            b1[i] -= 1;
            memcpy(b2, b1, sizeof b2);
            b2[i] -= 1;
            memcpy(a2, a, sizeof a2);
            a2[i] -= 1;
            v = c->PB[i] * oe_int(c, a, b1, 0) + c->o2z * b1[i] * oe_int(c, a, b2, 0) +
                c->o2z * a[i] * oe_int(c, a2, b1, 0);
        } else if (la > 0) {    /* vrr1 on a (b = 0) */
            int i = oe_first(a);
            a1[i] -= 1;
            memcpy(a2, a1, sizeof a2);
            a2[i] -= 1;
            v = c->PA[i] * oe_int(c, a1, b, 0) + c->o2z * a1[i] * oe_int(c, a2, b, 0);
        } else {
            v = c->base;
        }
    } else {
        if (la > 0) {           /* vrr2 on a (added last, outermost) */
            int i = oe_first(a);
            a1[i] -= 1;
            memcpy(a2, a1, sizeof a2);
            a2[i] -= 1;
            memcpy(b2, b, sizeof b2);
            b2[i] -= 1;
            v = c->PA[i] * oe_int(c, a1, b, m) - c->PC[i] * oe_int(c, a1, b, m + 1) +
                c->o2z * b[i] * (oe_int(c, a1, b2, m) - oe_int(c, a1, b2, m + 1)) +
                c->o2z * a1[i] * (oe_int(c, a2, b, m) - oe_int(c, a2, b, m + 1));
        } else if (lb > 0) {    /* vrr1 on b (a = 0) */
            int i = oe_first(b);
            b1[i] -= 1;
            memcpy(b2, b1, sizeof b2);
            b2[i] -= 1;
            v = c->PB[i] * oe_int(c, a, b1, m) - c->PC[i] * oe_int(c, a, b1, m + 1) +
                c->o2z * b1[i] * (oe_int(c, a, b2, m) - oe_int(c, a, b2, m + 1));
        } else {
            v = c->base * c->F[m];
        }
    }
    c->memo[ia][ib][m] = v;
    c->have[ia][ib][m] = 1;
    return v;
}

/* primitive Cartesian block out[ia * ncc(lb) + ib] (unnormalised primitives) */
void oracle_onee_prim(int kind, int la, int lb, double xa, const double ca[3], double xb, const double cb[3],
                      const double cc[3], double *out)
{
    oe_ctx *c = (oe_ctx *)calloc(1, sizeof(oe_ctx));
    const double z = xa + xb, oz = 1.0 / z;
    double r2 = 0.0, pc2 = 0.0;
    c->kind = kind;
    c->o2z = 0.5 * oz;
    for (int i = 0; i < 3; ++i) {
        const double P = (xa * ca[i] + xb * cb[i]) * oz;
        c->PA[i] = P - ca[i];
        c->PB[i] = P - cb[i];
        r2 += (ca[i] - cb[i]) * (ca[i] - cb[i]);
        if (kind == 1) {
            c->PC[i] = P - cc[i];
            pc2 += c->PC[i] * c->PC[i];
        }
    }
    const double kab = exp(-xa * xb * oz * r2);
    if (kind == 0) {
        c->base = sqrt(M_PI * M_PI * M_PI * oz * oz * oz) * kab;
    } else {
        c->base = 2.0 * M_PI * oz * kab;
        oracle_boys_eval(c->F, la + lb, z * pc2);
    }
    int a[3], b[3];
    for (int i = 0; i < oe_ncc(la); ++i) {
        oe_comp(la, i, a);
        for (int j = 0; j < oe_ncc(lb); ++j) {
            oe_comp(lb, j, b);
            out[i * oe_ncc(lb) + j] = oe_int(c, a, b, 0);
        }
    }
    free(c);
}

/* contracted block (one contraction per shell), Cartesian or spherical,
 * row-major [a][b] */
static void oe_cont(int kind, int spherical, int la, const double *xa, const double *coa, int nap,
                    const double ca[3], const double *Sa, int lb, const double *xb, const double *cob, int nbp,
                    const double cb[3], const double *Sb, const double *cc, double *out)
{
    const int na = oe_ncc(la), nb = oe_ncc(lb);
    double cart[OE_N * OE_N], prim[OE_N * OE_N];
    memset(cart, 0, sizeof cart);
    for (int i = 0; i < nap; ++i)
        for (int j = 0; j < nbp; ++j) {
            oracle_onee_prim(kind, la, lb, xa[i], ca, xb[j], cb, cc, prim);
            for (int k = 0; k < na * nb; ++k) cart[k] += coa[i] * cob[j] * prim[k];
        }
    if (!spherical) {
        memcpy(out, cart, sizeof(double) * na * nb);
        return;
    }
    const int sa = 2 * la + 1, sb = 2 * lb + 1;
    for (int p = 0; p < sa; ++p)
        for (int q = 0; q < sb; ++q) {
            double s = 0.0;
            for (int i = 0; i < na; ++i)
                for (int j = 0; j < nb; ++j) s += Sa[i + p * na] * cart[i * nb + j] * Sb[j + q * nb];
            out[p * sb + q] = s;
        }
}

int oracle_onee_batch(int kind, int spherical, int n_shell, const int *shell_l, const int *shell_nprim,
                      const int *shell_off, const double *shell_cen, const double *exps, const double *coefs,
                      const double *const *sph_by_l, const double *centres, long n_item, const int *items,
                      const long *out_off, double *out, int nthreads)
{
    (void)n_shell;
    const int w = kind == 0 ? 2 : 3;
    for (int s = 0; s < n_shell; ++s)
        if (shell_l[s] < 0 || shell_l[s] > OE_LMAX) return -1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (long q = 0; q < n_item; ++q) {
        const int a = items[w * q], b = items[w * q + 1];
        const double *cc = kind == 1 ? centres + 3 * items[w * q + 2] : NULL;
        oe_cont(kind, spherical, shell_l[a], exps + shell_off[a], coefs + shell_off[a], shell_nprim[a],
                shell_cen + 3 * a, spherical ? sph_by_l[shell_l[a]] : NULL, shell_l[b], exps + shell_off[b],
                coefs + shell_off[b], shell_nprim[b], shell_cen + 3 * b, spherical ? sph_by_l[shell_l[b]] : NULL,
                cc, out + out_off[q]);
    }
    return 0;
}
