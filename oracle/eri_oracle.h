/*
 * eri_oracle.h -- CPU restatement of Intception's generated contracted ERI
 * code.  TEST INFRASTRUCTURE ONLY (the checker); never linked by the product.
 */
#ifndef ERI_ORACLE_H
#define ERI_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_BOYS_REFERENCE = 0, ORACLE_BOYS_EXACT = 1 };

void oracle_set_boys_tables(const double *smallx, const double *mediumx);
void oracle_set_boys_mode(int mode);
int oracle_get_boys_mode(void);
void oracle_boys_ref(double *f, int ifskip, int nmin, int nmax, double x);
void oracle_boys_exact(double *f, int nmin, int nmax, double x);
void oracle_boys_eval(double *f, int nmax, double x);

int oracle_eri_work_array_size(int spherical, int la, int lb, int lc, int ld,
                               int na_prim, int nb_prim, int nc_prim, int nd_prim,
                               int na_cont, int nb_cont, int nc_cont, int nd_cont);

#define ORACLE_SHELL_ARGS(X) const double *x##X, const double c##X[3], int l##X, int n##X, \
    int iskip##X, int n##X##_prim, int n##X##_cont, int iskip##X##_cont, const double *contract_array_##X

void oracle_eri_cont_sph_main(
    ORACLE_SHELL_ARGS(a), const double *sph_trans_array_a,
    ORACLE_SHELL_ARGS(b), const double *sph_trans_array_b,
    ORACLE_SHELL_ARGS(c), const double *sph_trans_array_c,
    ORACLE_SHELL_ARGS(d), const double *sph_trans_array_d,
    double *work_array, double *output_array);

void oracle_eri_cont_cart_main(
    ORACLE_SHELL_ARGS(a), ORACLE_SHELL_ARGS(b), ORACLE_SHELL_ARGS(c), ORACLE_SHELL_ARGS(d),
    double *work_array, double *output_array);

int oracle_eri_batch(int spherical, int n_shell, const int *shell_l, const int *shell_nprim,
                     const int *shell_off, const double *shell_cen, const double *exps,
                     const double *coefs, const double *const *sph_by_l,
                     long n_quartet, const int *quartets, const long *out_off, double *out,
                     int nthreads);

/* one-electron kinds (onee_oracle.c): kind 0 overlap, 1 nuclear attraction */
void oracle_onee_prim(int kind, int la, int lb, double xa, const double ca[3], double xb, const double cb[3],
                      const double cc[3], double *out);
int oracle_onee_batch(int kind, int spherical, int n_shell, const int *shell_l, const int *shell_nprim,
                      const int *shell_off, const double *shell_cen, const double *exps, const double *coefs,
                      const double *const *sph_by_l, const double *centres, long n_item, const int *items,
                      const long *out_off, double *out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
