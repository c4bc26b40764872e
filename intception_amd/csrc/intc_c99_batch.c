/* intc_c99_batch.c -- OpenMP driver of the generated C99 backend over a batch
 * of quartets (the CPU baseline beside the GPU path).  The generated
 * <full>_main (INTC_C99_MAIN, spherical, n_cont = 1, the caller's S_l) is
 * called once per quartet with a dense row-major [a][b][c][d] block at
 * out[out_off[q]], threads over quartets, schedule(dynamic, 64). */
#include <stddef.h>
#include INTC_C99_HEADER

int intc_c99_batch(const int *shell_l, const int *shell_nprim, const int *shell_off, const double *shell_cen,
                   const double *exps, const double *coefs, const double *sph, long long n_quartet,
                   const int *quartets, const long long *out_off, double *out, int nthreads)
{
    long long q;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) reduction(| : bad)
    for (q = 0; q < n_quartet; ++q) {
        const int *s = quartets + 4 * q;
        int n[4], k, off_s[4];
        for (k = 0; k < 4; ++k) {
            const int l = shell_l[s[k]];
            int o = 0, j;
            if (l < 0 || l > 7) { bad = 1; break; }
            n[k] = 2 * l + 1;
            for (j = 0; j < l; ++j) o += (j + 1) * (j + 2) / 2 * (2 * j + 1);
            off_s[k] = o;
        }
        if (k < 4) continue;
        INTC_C99_MAIN(exps + shell_off[s[0]], shell_cen + 3 * s[0], shell_l[s[0]], n[0], n[1] * n[2] * n[3],
                      shell_nprim[s[0]], 1, 0, coefs + shell_off[s[0]], sph + off_s[0],
                      exps + shell_off[s[1]], shell_cen + 3 * s[1], shell_l[s[1]], n[1], n[2] * n[3],
                      shell_nprim[s[1]], 1, 0, coefs + shell_off[s[1]], sph + off_s[1],
                      exps + shell_off[s[2]], shell_cen + 3 * s[2], shell_l[s[2]], n[2], n[3],
                      shell_nprim[s[2]], 1, 0, coefs + shell_off[s[2]], sph + off_s[2],
                      exps + shell_off[s[3]], shell_cen + 3 * s[3], shell_l[s[3]], n[3], 1,
                      shell_nprim[s[3]], 1, 0, coefs + shell_off[s[3]], sph + off_s[3],
                      NULL, out + out_off[q]);
    }
    return bad ? -1 : 0;
}
