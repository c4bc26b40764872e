/*
 * eri_oracle.c -- CPU restatement of the contracted ERI code that Intception
 * generates for the (ab|cd) input script (VRR on a, c, d with one auxiliary
 * index m, then one HGP HRR a->b).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (intception_amd) never links, loads or calls it.
 *
 * What each part restates (reference paths are relative to the reference
 * repository root):
 *   oracle_boys_ref ........ src/intception/generator/callables/generator_boys_function.py:98-205
 *                            (3 branches: x<=15 Taylor + downward recursion,
 *                            15<x<35 Taylor of F0 + upward recursion, x>=35
 *                            asymptotic F0 + upward recursion without exp(-x));
 *                            tables: src/intception/boys.py:27-97 + the two boys_data JSON files
 *   base ................... src/intception/generator/callables/integral_wrapper_base_function.py:116-155
 *   2-layer m-loop VRR ..... src/intception/generator/callables/main_source_algorithms/contracted_general.py:720-1080
 *                            with the m-group bounds of src/intception/generator/integral_wrapper.py:1153-1254
 *                            (per m, high to low: a to min(L, M-m), then c to
 *                            min(lc, lc+ld-m), then d to ld-m) and the direction
 *                            rule "first non-zero component of the target"
 *                            (src/intception/generator/generator.py:547-558)
 *   contract ............... generator_contract_function.py:69-125 (out = in . C,
 *                            C[iprim + icont*n_prim]); truncation of a to
 *                            la <= |a| <= L (contracted_general.py:1083-1260)
 *   hrr .................... generator_hrr_function.py:265-435:
 *                            (a,b) = (a+1i,b-1i) + AB_i (a,b-1i), AB = A-B,
 *                            direction = first non-zero component of b
 *   sph_trans .............. generator_sph_trans_function.py:69-125
 *                            (out = in . S, S[icart + isph*n_cart])
 *   output copy ............ contracted_general.py:1808-1875
 *                            out[sum_X(iX*iskipX + icontX*iskipX_cont)]
 *   work_array_size ........ contracted_vrr_hrr_with_auxiliary.py:248-351 and
 *                            contracted_general.py:246-501, hrr_layer from
 *                            generator.py:1112-1157
 *
 * Summation order differs from the generated code (this file accumulates the
 * contraction inside the primitive loop instead of storing every primitive
 * block first), so results agree with it to rounding, not bit for bit.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "eri_oracle.h"

/* ------------------------------------------------------------------ */
/* Cartesian component bookkeeping (src/intception/dsl.py:988-1041)    */
/* ------------------------------------------------------------------ */
static int ncc(int l) { return l < 0 ? 0 : (l + 1) * (l + 2) / 2; }
static int ncc0(int l) { return l < 0 ? 0 : (l + 1) * (l + 2) * (l + 3) / 6; }

/* 0-based index of component (x,y,z) among all components with |c| <= l,
 * ordered by l then x^l first ... z^l last. */
static int cidx(int x, int y, int z)
{
    int l = x + y + z;
    int yz = y + z;
    return ncc0(l - 1) + yz * (yz + 1) / 2 + z;
}

typedef struct { int x[3]; } comp_t;

static void comp_of(int idx, int out[3])
{
    int l = 0;
    while (ncc0(l) <= idx) l++;
    int r = idx - ncc0(l - 1);
    /* within-shell index r = yz*(yz+1)/2 + z */
    int yz = 0;
    while ((yz + 1) * (yz + 2) / 2 <= r) yz++;
    int z = r - yz * (yz + 1) / 2;
    out[0] = l - yz;
    out[1] = yz - z;
    out[2] = z;
}

static int first_nonzero(const int c[3])
{
    for (int i = 0; i < 3; i++)
        if (c[i] > 0) return i;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Boys function                                                       */
/* ------------------------------------------------------------------ */
static const double *g_smallx = NULL;   /* [151][33], x = k/10, k = 0..150 */
static const double *g_mediumx = NULL;  /* [201][7],  x = 15 + k/10        */
static int g_boys_mode = ORACLE_BOYS_REFERENCE;

void oracle_set_boys_tables(const double *smallx, const double *mediumx)
{
    g_smallx = smallx;
    g_mediumx = mediumx;
}

void oracle_set_boys_mode(int mode) { g_boys_mode = mode; }
int oracle_get_boys_mode(void) { return g_boys_mode; }

/* generator_boys_function.py:98-205.  oof[k] = 1/k! (boys.py:62-69). */
void oracle_boys_ref(double *f, int ifskip, int nmin, int nmax, double x)
{
    static const double oof[7] = {1.0, 1.0, 1.0 / 2.0, 1.0 / 6.0, 1.0 / 24.0,
                                  1.0 / 120.0, 1.0 / 720.0};
    const int ntaylor = 7;
    if (x <= 15.0) {
        const double ooxstep = 1.0 / 0.1;
        int i = (int)round(x * ooxstep);             /* nint(): (int) round() */
        double dx = (double)i * 0.1 - x;
        double xn = 1.0;
        const double *row = g_smallx + (size_t)i * 33;
        f[nmax * ifskip] = row[nmax];
        for (int ii = 1; ii < ntaylor; ii++) {
            xn = xn * dx;
            f[nmax * ifskip] = f[nmax * ifskip] + row[nmax + ii] * xn * oof[ii];
        }
        double expx = exp(-x);
        for (int ii = nmax - 1; ii >= nmin; ii--)
            f[ii * ifskip] = (2.0 * x * f[(ii + 1) * ifskip] + expx) / (2.0 * (double)ii + 1.0);
        return;
    } else if (x > 15.0 && x < 35.0) {
        const double ooxstep = 1.0 / 0.1;
        int i = (int)round(x * ooxstep);
        double dx = (double)i * 0.1 - x;
        i = i - 150;                                 /* xmin / xstep */
        double oox = 1.0 / x;
        double xn = 1.0;
        const double *row = g_mediumx + (size_t)i * 7;
        f[0] = row[0];
        for (int ii = 1; ii < ntaylor; ii++) {
            xn = xn * dx;
            f[0] = f[0] + row[ii] * xn * oof[ii];
        }
        double expx = exp(-x);
        for (int ii = 0; ii < nmax; ii++)
            f[ii * ifskip + ifskip] = ((double)ii + 0.5) * oox * f[ii * ifskip] - expx * 0.5 * oox;
        return;
    } else {
        double oox = 1.0 / x;
        f[0] = 0.5 * sqrt(M_PI * oox);
        for (int ii = 0; ii < nmax; ii++)
            f[ii * ifskip + ifskip] = ((double)ii + 0.5) * oox * f[ii * ifskip];
        return;
    }
}

/* Accurate Boys function (not the reference's algorithm): power series
 * F_n(x) = exp(-x) sum_k (2x)^k / ((2n+1)(2n+3)...(2n+2k+1)) for x < 30 and
 * downward recursion; F_0 = sqrt(pi/x) erf(sqrt x)/2 with upward recursion
 * (including the exp(-x) term) for x >= 30.  Used to separate the reference's
 * large-x approximation from the rest of the algorithm in parity checks. */
void oracle_boys_exact(double *f, int nmin, int nmax, double x)
{
    double expx = exp(-x);
    if (x < 30.0) {
        double term = 1.0 / (2.0 * nmax + 1.0), sum = term;
        for (int k = 1; k < 400; k++) {
            term *= 2.0 * x / (2.0 * nmax + 2.0 * k + 1.0);
            sum += term;
            if (term < 1e-18 * sum) break;
        }
        f[nmax] = expx * sum;
        for (int m = nmax - 1; m >= nmin; m--)
            f[m] = (2.0 * x * f[m + 1] + expx) / (2.0 * m + 1.0);
    } else {
        double sx = sqrt(x);
        f[0] = 0.5 * sqrt(M_PI) * erf(sx) / sx;
        for (int m = 0; m < nmax; m++)
            f[m + 1] = ((2.0 * m + 1.0) * f[m] - expx) / (2.0 * x);
    }
}

static void boys_eval(double *f, int nmax, double x)
{
    if (g_boys_mode == ORACLE_BOYS_REFERENCE)
        oracle_boys_ref(f, 1, 0, nmax, x);
    else
        oracle_boys_exact(f, 0, nmax, x);
}

/* F_0..F_nmax in the current mode (shared with onee_oracle.c) */
void oracle_boys_eval(double *f, int nmax, double x) { boys_eval(f, nmax, x); }

/* ------------------------------------------------------------------ */
/* work_array_size                                                     */
/* ------------------------------------------------------------------ */
static long hrr_layer(int la, int lb)
{
    /* generator.py:1114-1141: max over lq of (ncc0(la+lb-lq)-ncc0(lb-1))*ncc(lq) */
    long best = 0;
    for (int lq = 0; lq <= la; lq++) {
        long s = (long)(ncc0(la + lb - lq) - ncc0(lb - 1)) * ncc(lq);
        if (s > best) best = s;
    }
    return best;
}

int oracle_eri_work_array_size(int spherical, int la, int lb, int lc, int ld,
                               int na_prim, int nb_prim, int nc_prim, int nd_prim,
                               int na_cont, int nb_cont, int nc_cont, int nd_cont)
{
    int L = la + lb, M = la + lb + lc + ld;
    long nvrr = (long)ncc0(L) * ncc0(lc) * ncc0(ld);
    long nprim = (long)na_prim * nb_prim * nc_prim * nd_prim;
    /* VRR phase */
    long mem = 2 * (nvrr * nprim + M + 1);
    /* early transform: largest prim/cont product in the contraction
     * sequence; the maximum over both sequence orders bounds it. */
    int np[4] = {na_prim, nb_prim, nc_prim, nd_prim};
    int nc[4] = {na_cont, nb_cont, nc_cont, nd_cont};
    long nmax = nprim;
    for (int dir = 0; dir < 2; dir++) {
        int cur[4] = {np[0], np[1], np[2], np[3]};
        for (int s = 0; s < 4; s++) {
            int k = dir == 0 ? s : 3 - s;
            cur[k] = nc[k];
            long v = (long)cur[0] * cur[1] * cur[2] * cur[3];
            if (v > nmax) nmax = v;
        }
    }
    long m2 = 2 * nvrr * nmax;
    if (m2 > mem) mem = m2;
    long ncont = (long)na_cont * nb_cont * nc_cont * nd_cont;
    int ncomp_c = spherical ? 2 * lc + 1 : ncc(lc);
    int ncomp_d = spherical ? 2 * ld + 1 : ncc(ld);
    /* HRR phase: hrr_layer[l_move_to][l_move_from], move_to = b, move_from = a */
    m2 = 2 * hrr_layer(lb, la) * ncomp_c * ncomp_d * ncont;
    if (m2 > mem) mem = m2;
    /* late transform phase */
    int ncomp_a = ncc(la), ncomp_b = ncc(lb);
    m2 = 2L * ncomp_a * ncomp_b * ncc(lc) * ncc(ld) * ncont;
    if (!spherical) m2 = 2L * ncomp_a * ncomp_b * ncomp_c * ncomp_d * ncont;
    if (m2 > mem) mem = m2;
    return (int)mem;
}

/* ------------------------------------------------------------------ */
/* one contracted quartet                                              */
/* ------------------------------------------------------------------ */
typedef struct {
    const double *x;      /* exponents, n_prim */
    const double *cen;    /* centre[3] */
    int l, n, iskip, n_prim, n_cont, iskip_cont;
    const double *contract;  /* n_prim x n_cont, column-major */
    const double *sph;       /* ncc(l) x (2l+1), column-major, or NULL */
} shell_arg_t;

static void eri_quartet(const shell_arg_t *sa, const shell_arg_t *sb,
                        const shell_arg_t *sc, const shell_arg_t *sd,
                        int spherical, double *out)
{
    const int la = sa->l, lb = sb->l, lc = sc->l, ld = sd->l;
    const int L = la + lb, M = L + lc + ld;
    const int nA = ncc0(L), nC = ncc0(lc), nD = ncc0(ld);
    const size_t layer = (size_t)nA * nC * nD;
    double *cur = (double *)calloc(layer, sizeof(double));
    double *prv = (double *)calloc(layer, sizeof(double));
    double F[64];
    int (*ca)[3] = malloc(sizeof(int[3]) * nA);
    int (*cc)[3] = malloc(sizeof(int[3]) * nC);
    int (*cd)[3] = malloc(sizeof(int[3]) * nD);
    for (int i = 0; i < nA; i++) comp_of(i, ca[i]);
    for (int i = 0; i < nC; i++) comp_of(i, cc[i]);
    for (int i = 0; i < nD; i++) comp_of(i, cd[i]);

    /* contracted block (|a| in [la, L], |c| = lc, |d| = ld) for every
     * contraction combination */
    const int aoff = ncc0(la - 1), nAe = nA - aoff;
    const int coff = ncc0(lc - 1), nCc = ncc(lc);
    const int doff = ncc0(ld - 1), nDc = ncc(ld);
    const int ncon = sa->n_cont * sb->n_cont * sc->n_cont * sd->n_cont;
    const size_t cblk = (size_t)nAe * nCc * nDc;
    double *acc = (double *)calloc(cblk * ncon, sizeof(double));

    const double *A = sa->cen, *B = sb->cen, *C = sc->cen, *D = sd->cen;
    double AB[3], CD[3];
    for (int i = 0; i < 3; i++) { AB[i] = A[i] - B[i]; CD[i] = C[i] - D[i]; }
    const double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
    const double rcd2 = CD[0] * CD[0] + CD[1] * CD[1] + CD[2] * CD[2];
    const double two_pi_52 = 2.0 * pow(M_PI, 2.5);

#define IDX(a, c, d) (((size_t)(a) * nC + (c)) * nD + (d))

    for (int ia = 0; ia < sa->n_prim; ia++)
    for (int ib = 0; ib < sb->n_prim; ib++) {
        const double xa = sa->x[ia], xb = sb->x[ib];
        const double zeta = xa + xb, ooz = 1.0 / zeta;
        double P[3], PA[3];
        for (int i = 0; i < 3; i++) { P[i] = (xa * A[i] + xb * B[i]) * ooz; PA[i] = P[i] - A[i]; }
        const double kab = exp(-xa * xb * ooz * rab2);
        for (int ic = 0; ic < sc->n_prim; ic++)
        for (int id = 0; id < sd->n_prim; id++) {
            const double xc = sc->x[ic], xd = sd->x[id];
            const double eta = xc + xd, ooe = 1.0 / eta;
            double Q[3], QC[3], QD[3], W[3], WP[3], WQ[3];
            for (int i = 0; i < 3; i++) {
                Q[i] = (xc * C[i] + xd * D[i]) * ooe;
                QC[i] = Q[i] - C[i];
                QD[i] = Q[i] - D[i];
            }
            const double kcd = exp(-xc * xd * ooe * rcd2);
            const double ozpe = 1.0 / (zeta + eta);
            const double rho = zeta * eta * ozpe;
            double rpq2 = 0.0;
            for (int i = 0; i < 3; i++) {
                W[i] = (zeta * P[i] + eta * Q[i]) * ozpe;
                WP[i] = W[i] - P[i];
                WQ[i] = W[i] - Q[i];
                rpq2 += (P[i] - Q[i]) * (P[i] - Q[i]);
            }
            const double T = rho * rpq2;
            const double o2z = 0.5 * ooz, o2e = 0.5 * ooe, o2ze = 0.5 * ozpe;
            const double roz = rho * ooz, roe = rho * ooe;
            /* base expression of the ERI input script:
             * 2 pi^(5/2) / (zeta eta sqrt(zeta+eta)) K_AB K_CD F_m(T) */
            const double pref = two_pi_52 * ooz * ooe / sqrt(zeta + eta) * kab * kcd;
            boys_eval(F, M, T);

            /* 2-layer algorithm: m from M down to 0 */
            for (int m = M; m >= 0; m--) {
                double *t = prv; prv = cur; cur = t;   /* prv = layer m+1 */
                cur[IDX(0, 0, 0)] = pref * F[m];
                /* a-build to min(L, M-m) */
                int amax = L < M - m ? L : M - m;
                for (int a = 1; a < ncc0(amax); a++) {
                    const int *e = ca[a];
                    int i = first_nonzero(e);
                    int e1[3] = {e[0], e[1], e[2]}; e1[i]--;
                    int a1 = cidx(e1[0], e1[1], e1[2]);
                    double v = PA[i] * cur[IDX(a1, 0, 0)];
                    if (m < M) v += WP[i] * prv[IDX(a1, 0, 0)];
                    if (e1[i] > 0) {
                        int e2[3] = {e1[0], e1[1], e1[2]}; e2[i]--;
                        int a2 = cidx(e2[0], e2[1], e2[2]);
                        double t2 = cur[IDX(a2, 0, 0)];
                        if (m < M) t2 -= roz * prv[IDX(a2, 0, 0)];
                        v += e1[i] * o2z * t2;
                    }
                    cur[IDX(a, 0, 0)] = v;
                }
                /* c-build to min(lc, lc+ld-m), all |a| <= L */
                int cmax = lc + ld - m;
                if (cmax > lc) cmax = lc;
                for (int c = 1; c < ncc0(cmax); c++) {
                    const int *g = cc[c];
                    int i = first_nonzero(g);
                    int g1[3] = {g[0], g[1], g[2]}; g1[i]--;
                    int c1 = cidx(g1[0], g1[1], g1[2]);
                    int c2 = -1;
                    if (g1[i] > 0) { int g2[3] = {g1[0], g1[1], g1[2]}; g2[i]--; c2 = cidx(g2[0], g2[1], g2[2]); }
                    for (int a = 0; a < nA; a++) {
                        const int *e = ca[a];
                        double v = QC[i] * cur[IDX(a, c1, 0)] + WQ[i] * prv[IDX(a, c1, 0)];
                        if (c2 >= 0)
                            v += g1[i] * o2e * (cur[IDX(a, c2, 0)] - roe * prv[IDX(a, c2, 0)]);
                        if (e[i] > 0) {
                            int e1[3] = {e[0], e[1], e[2]}; e1[i]--;
                            v += e[i] * o2ze * prv[IDX(cidx(e1[0], e1[1], e1[2]), c1, 0)];
                        }
                        cur[IDX(a, c, 0)] = v;
                    }
                }
                /* d-build to ld-m, all |a| <= L, |c| <= lc */
                int dmax = ld - m;
                for (int d = 1; d < ncc0(dmax); d++) {
                    const int *h = cd[d];
                    int i = first_nonzero(h);
                    int h1[3] = {h[0], h[1], h[2]}; h1[i]--;
                    int d1 = cidx(h1[0], h1[1], h1[2]);
                    int d2 = -1;
                    if (h1[i] > 0) { int h2[3] = {h1[0], h1[1], h1[2]}; h2[i]--; d2 = cidx(h2[0], h2[1], h2[2]); }
                    for (int a = 0; a < nA; a++) {
                        const int *e = ca[a];
                        int a1 = -1;
                        if (e[i] > 0) { int e1[3] = {e[0], e[1], e[2]}; e1[i]--; a1 = cidx(e1[0], e1[1], e1[2]); }
                        for (int c = 0; c < nC; c++) {
                            const int *g = cc[c];
                            double v = QD[i] * cur[IDX(a, c, d1)] + WQ[i] * prv[IDX(a, c, d1)];
                            if (d2 >= 0)
                                v += h1[i] * o2e * (cur[IDX(a, c, d2)] - roe * prv[IDX(a, c, d2)]);
                            if (g[i] > 0) {
                                int g1[3] = {g[0], g[1], g[2]}; g1[i]--;
                                int c1 = cidx(g1[0], g1[1], g1[2]);
                                v += g[i] * o2e * (cur[IDX(a, c1, d1)] - roe * prv[IDX(a, c1, d1)]);
                            }
                            if (a1 >= 0)
                                v += e[i] * o2ze * prv[IDX(a1, c, d1)];
                            cur[IDX(a, c, d)] = v;
                        }
                    }
                }
            }
            /* contraction (generator_contract_function.py:69-125): every
             * contraction combination gets coefficient product times the
             * truncated m=0 block */
            int k = 0;
            for (int ka = 0; ka < sa->n_cont; ka++)
            for (int kb = 0; kb < sb->n_cont; kb++)
            for (int kc = 0; kc < sc->n_cont; kc++)
            for (int kd = 0; kd < sd->n_cont; kd++, k++) {
                double w = sa->contract[ia + ka * sa->n_prim] * sb->contract[ib + kb * sb->n_prim] *
                           sc->contract[ic + kc * sc->n_prim] * sd->contract[id + kd * sd->n_prim];
                double *blk = acc + cblk * k;
                for (int a = 0; a < nAe; a++)
                    for (int c = 0; c < nCc; c++)
                        for (int d = 0; d < nDc; d++)
                            blk[((size_t)a * nCc + c) * nDc + d] += w * cur[IDX(a + aoff, c + coff, d + doff)];
            }
        }
    }
#undef IDX

    /* HRR a->b on contracted integrals, then sph_trans, then output copy */
    const int nAc = ncc(la), nBc = ncc(lb);
    const int nsa = spherical ? 2 * la + 1 : nAc, nsb = spherical ? 2 * lb + 1 : nBc;
    const int nsc = spherical ? 2 * lc + 1 : nCc, nsd = spherical ? 2 * ld + 1 : nDc;
    const int ncd = nCc * nDc;
    /* hrr work: h[bl][ (a in [la, L-bl]) x (b with |b| = bl) x cd ] */
    double *hb = (double *)malloc(sizeof(double) * (size_t)nAe * ncc(L) * ncd + 1);
    double *hn = (double *)malloc(sizeof(double) * (size_t)nAe * ncc(L) * ncd + 1);
    double *cart = (double *)malloc(sizeof(double) * (size_t)nAc * nBc * ncd);
    double *t1 = (double *)malloc(sizeof(double) * (size_t)nAc * nBc * ncd);
    double *t2 = (double *)malloc(sizeof(double) * (size_t)nAc * nBc * ncd);
    int k = 0;
    for (int ka = 0; ka < sa->n_cont; ka++)
    for (int kb = 0; kb < sb->n_cont; kb++)
    for (int kc = 0; kc < sc->n_cont; kc++)
    for (int kd = 0; kd < sd->n_cont; kd++, k++) {
        const double *blk = acc + cblk * k;
        /* level bl = 0: hb[a][0][cd] for |a| in [la, L] */
        int nb_cur = 1;
        memcpy(hb, blk, sizeof(double) * cblk);
        for (int bl = 1; bl <= lb; bl++) {
            /* targets: |a| in [la, L-bl], |b| = bl */
            int na_t = ncc0(L - bl) - aoff;
            int nb_t = ncc(bl);
            for (int a = 0; a < na_t; a++) {
                int e[3]; comp_of(a + aoff, e);
                for (int b = 0; b < nb_t; b++) {
                    int f[3]; comp_of(ncc0(bl - 1) + b, f);
                    int i = first_nonzero(f);
                    int f1[3] = {f[0], f[1], f[2]}; f1[i]--;
                    int b1 = cidx(f1[0], f1[1], f1[2]) - ncc0(bl - 2);
                    int e1[3] = {e[0], e[1], e[2]}; e1[i]++;
                    int a1 = cidx(e1[0], e1[1], e1[2]) - aoff;
                    for (int q = 0; q < ncd; q++)
                        hn[((size_t)a * nb_t + b) * ncd + q] =
                            hb[((size_t)a1 * nb_cur + b1) * ncd + q] + AB[i] * hb[((size_t)a * nb_cur + b1) * ncd + q];
                }
            }
            double *t = hb; hb = hn; hn = t;
            nb_cur = nb_t;
        }
        /* hb now [a (|a| = la)][b (|b| = lb)][c][d] */
        memcpy(cart, hb, sizeof(double) * (size_t)nAc * nBc * ncd);
        const double *res = cart;
        if (spherical) {
            /* d index, then c, then b, then a: out = in . S */
            int dims[4] = {nAc, nBc, nCc, nDc};
            const double *S[4] = {sa->sph, sb->sph, sc->sph, sd->sph};
            int nsph[4] = {nsa, nsb, nsc, nsd};
            const double *in = cart;
            double *bufs[2] = {t1, t2};
            for (int x = 3; x >= 0; x--) {
                double *o = bufs[x & 1];
                int pre = 1, post = 1;
                for (int y = 0; y < x; y++) pre *= dims[y];
                for (int y = x + 1; y < 4; y++) post *= dims[y];
                int n_cart = dims[x];
                for (int p = 0; p < pre; p++)
                    for (int s = 0; s < nsph[x]; s++)
                        for (int r = 0; r < post; r++) {
                            double v = 0.0;
                            for (int ct = 0; ct < n_cart; ct++)
                                v += in[((size_t)p * n_cart + ct) * post + r] * S[x][ct + s * n_cart];
                            o[((size_t)p * nsph[x] + s) * post + r] = v;
                        }
                dims[x] = nsph[x];
                in = o;
            }
            res = in;
        }
        for (int a = 0; a < nsa; a++)
            for (int b = 0; b < nsb; b++)
                for (int c = 0; c < nsc; c++)
                    for (int d = 0; d < nsd; d++)
                        out[(size_t)a * sa->iskip + ka * sa->iskip_cont + (size_t)b * sb->iskip + kb * sb->iskip_cont +
                            (size_t)c * sc->iskip + kc * sc->iskip_cont + (size_t)d * sd->iskip + kd * sd->iskip_cont] =
                            res[(((size_t)a * nsb + b) * nsc + c) * nsd + d];
    }
    free(hb); free(hn); free(cart); free(t1); free(t2);
    free(acc); free(cur); free(prv); free(ca); free(cc); free(cd);
}

#define SHELL_ARGS(X) const double *x##X, const double c##X[3], int l##X, int n##X, int iskip##X, \
    int n##X##_prim, int n##X##_cont, int iskip##X##_cont, const double *contract_array_##X

#define MK(X, S) shell_arg_t s##X = {x##X, c##X, l##X, n##X, iskip##X, n##X##_prim, n##X##_cont, \
    iskip##X##_cont, contract_array_##X, S}

void oracle_eri_cont_sph_main(
    SHELL_ARGS(a), const double *sph_trans_array_a,
    SHELL_ARGS(b), const double *sph_trans_array_b,
    SHELL_ARGS(c), const double *sph_trans_array_c,
    SHELL_ARGS(d), const double *sph_trans_array_d,
    double *work_array, double *output_array)
{
    (void)work_array;
    MK(a, sph_trans_array_a); MK(b, sph_trans_array_b);
    MK(c, sph_trans_array_c); MK(d, sph_trans_array_d);
    eri_quartet(&sa, &sb, &sc, &sd, 1, output_array);
}

void oracle_eri_cont_cart_main(
    SHELL_ARGS(a), SHELL_ARGS(b), SHELL_ARGS(c), SHELL_ARGS(d),
    double *work_array, double *output_array)
{
    (void)work_array;
    MK(a, NULL); MK(b, NULL); MK(c, NULL); MK(d, NULL);
    eri_quartet(&sa, &sb, &sc, &sd, 0, output_array);
}

/* ------------------------------------------------------------------ */
/* Batch harness: the same per-quartet routine over an SoA shell table, */
/* OpenMP over quartets (dynamic, 64).  This is how the CPU baseline    */
/* drives the per-quartet code; it is not part of the reference.       */
/* ------------------------------------------------------------------ */
int oracle_eri_batch(int spherical, int n_shell, const int *shell_l, const int *shell_nprim,
                     const int *shell_off, const double *shell_cen, const double *exps,
                     const double *coefs, const double *const *sph_by_l,
                     long n_quartet, const int *quartets, const long *out_off, double *out,
                     int nthreads)
{
    (void)n_shell;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (long q = 0; q < n_quartet; q++) {
        const int *s = quartets + 4 * q;
        shell_arg_t sh[4];
        int n[4];
        for (int k = 0; k < 4; k++) {
            int id = s[k];
            int l = shell_l[id];
            n[k] = spherical ? 2 * l + 1 : ncc(l);
            sh[k].x = exps + shell_off[id];
            sh[k].cen = shell_cen + 3 * (size_t)id;
            sh[k].l = l;
            sh[k].n = n[k];
            sh[k].n_prim = shell_nprim[id];
            sh[k].n_cont = 1;
            sh[k].iskip_cont = 0;
            sh[k].contract = coefs + shell_off[id];
            sh[k].sph = spherical ? sph_by_l[l] : NULL;
        }
        sh[3].iskip = 1;
        sh[2].iskip = n[3];
        sh[1].iskip = n[2] * n[3];
        sh[0].iskip = n[1] * n[2] * n[3];
        eri_quartet(&sh[0], &sh[1], &sh[2], &sh[3], spherical, out + out_off[q]);
    }
    return 0;
}
