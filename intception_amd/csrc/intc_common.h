// intc_common.h -- shared device helpers for the gfx950 ERI kernels.
//
// Layout of the data every class kernel reads (all SoA, resident in HBM):
//   quartet list of one class bin: bra pair id, ket pair id, output offset,
//     packed output strides (4 x 16 bit, for the canonical a, b, c, d)
//   shell-pair table: first primitive pair, primitive-pair count, centre A of
//     the pair's first shell and AB = A - B (3 doubles each)
//   primitive-pair arrays: zeta, 1/(2 zeta), P (3 arrays), kappa =
//     sqrt(2 pi^(5/2)) K_AB c_a c_b / zeta (contraction coefficients folded in)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef INTC_BOYS_NINT
#define INTC_BOYS_NINT 36
#define INTC_BOYS_DEG 12
#define INTC_BOYS_MMAX 16
#endif

struct IntcEriArgs {
    const int *q_bra;                 // [nq] bra pair id
    const int *q_ket;                 // [nq] ket pair id
    const long long *q_off;           // [nq] output offset (doubles)
    const unsigned long long *q_str;  // [nq] packed strides s_a | s_b<<16 | s_c<<32 | s_d<<48
    const int *q_idx;                 // optional queue: entry qq works on quartet q_idx[qq]
    const int *nq_dev;                // optional device-side queue length (overrides nq)
    int nq;
    int team_groups;                  // groups of G lanes cooperating on one quartet
    int boys_mode;                    // 0 Chebyshev, 1 reference three-branch scheme
    int pad0;
    const int *pair_off;              // [npair] first primitive pair
    const int *pair_np;               // [npair] Ka*Kb
    const int *pair_nk;               // [npair] Kb (primitive pair p = ia*Kb + ib)
    const double *pair_A;             // [npair*3]
    const double *pair_AB;            // [npair*3]
    const double *pp_zeta;            // [nprimpair]
    const double *pp_o2z;             // 1/(2 zeta)
    const double *pp_px, *pp_py, *pp_pz;
    const double *pp_kappa;
    const double *sph;                // S_l, l = 0..7, column-major, concatenated
    const double *boys_cheb;          // [(MMAX+1)][NINT][DEG+1]
    const double *boys_small;         // [151][33]
    const double *boys_medium;        // [201][7]
    double *out;
    int wg_cls[6];                    // workgroup kernels: la, lb, lc, ld, spherical (canonical class)
    int sph_builtin;                  // 1: sph is the built-in table (register kernels use literals)
    // FP32-recurrence register kernels, routed runs: a quartet whose error
    // estimate exceeds fp32_tol is appended to redo_q (redo_n counts them) for
    // the FP64 kernel that runs after it; fp32_tol <= 0 turns the estimate off
    int *redo_q;
    int *redo_n;
    double fp32_tol;
};

// One-electron integrals (overlap, nuclear attraction) of contracted shell
// pairs: item i is (shell it_a[i] | [centre it_c[i]] | shell it_b[i]) and
// writes out[it_off[i] + ia * s_a + ib * s_b] with s_a = it_str[i] & 0xffffffff,
// s_b = it_str[i] >> 32 over the components of a and b.
struct IntcOneArgs {
    const int *it_a, *it_b, *it_c;
    const long long *it_off;
    const unsigned long long *it_str;
    int n;                            // items of this class bin
    int boys_mode;
    const int *shell_l, *shell_np, *shell_off;
    const double *shell_cen;          // [n_shell * 3]
    const double *exps, *coefs;       // concatenated primitive data (normalised coefficients)
    const double *centres;            // [n_centre * 3] point-charge centres (nuclear attraction)
    const double *sph;                // S_l, l = 0..7 (the IntcEriArgs::sph layout)
    const double *boys_cheb, *boys_small, *boys_medium;
    double *out;
};

// offset of S_l inside IntcEriArgs::sph: sum_{l' < l} ncc(l') (2l' + 1)
__host__ __device__ constexpr int intc_sph_off(int l)
{
    int o = 0;
    for (int k = 0; k < l; ++k) o += (k + 1) * (k + 2) / 2 * (2 * k + 1);
    return o;
}

// cross-lane read of a double: value of v in lane src (byte address src*4)
__device__ __forceinline__ double intc_shfl(double v, int src4)
{
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_ds_bpermute(src4, lo);
    hi = __builtin_amdgcn_ds_bpermute(src4, hi);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ float intc_shflf(float v, int src4)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src4, __float_as_int(v)));
}

// p[i] read through the 32-bit byte offset i sizeof(T) (below 2^32): loads from several
// arrays at one index then share one offset register (scalar base + vector
// offset addressing) instead of forming a 64-bit address per array
template <class T>
__device__ __forceinline__ T intc_ld(const T *p, unsigned byte_off)
{
    return *(const T *)((const char *)p + byte_off);
}

// 1/sqrt(x) for normal x > 0: the hardware estimate (v_rsq_f64, relative
// error within 2^29 ulp = 2^-23) and two Newton steps y <- y + y (1 - x y^2) / 2,
// each squaring the relative error: within 1.3 ulp (checked on the host from
// estimates perturbed by 2^-23), in 9 instructions where a correctly rounded
// division and square root take ~22
__device__ __forceinline__ double intc_rsqrt(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
    y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
    return y;
}

// F_m(T), m = 0..M.  Mode 0: degree-DEG polynomial per unit interval for F_M
// on [0, NINT), downward recursion; asymptotic F_0 and upward recursion with
// the exp(-T) term beyond.  Mode 1: the reference's three-branch scheme
// (src/intception/generator/callables/generator_boys_function.py:98-205) on
// grid tables of the same shape as the reference's.
// Args: any argument struct with the boys_cheb / boys_small / boys_medium
// table pointers (IntcEriArgs, IntcOneArgs).
template <int M, int BM, class Args>
__device__ __forceinline__ void intc_boys(double T, double *F, const Args &a)
{
    // T = rho |PQ|^2 >= 0; the clamp (fmax maps NaN to 0) keeps every table
    // index below in range whatever the input
    T = fmax(T, 0.0);
    if (BM == 0) {
        // exp(-T) once, ahead of the branch: a wave whose lanes take both
        // branches would otherwise evaluate it twice
        const double ex = M > 0 ? exp(-T) : 0.0;
        if (T < (double)INTC_BOYS_NINT) {
            int k = (int)T;
            double t = 2.0 * (T - (double)k) - 1.0;
            const double *c = a.boys_cheb + ((size_t)(M * INTC_BOYS_NINT + k)) * (INTC_BOYS_DEG + 1);
            double v = c[INTC_BOYS_DEG];
#pragma unroll
            for (int j = INTC_BOYS_DEG - 1; j >= 0; --j) {
                v = fma(v, t, c[j]);
                // the upper half of the coefficients is consumed before the
                // lower half is loaded: fewer of them in registers at once
                if (j == INTC_BOYS_DEG / 2) __builtin_amdgcn_sched_barrier(0);
            }
            F[M] = v;
            if (M > 0) {
                double t2 = 2.0 * T;
#pragma unroll
                for (int m = M - 1; m >= 0; --m) F[m] = fma(t2, F[m + 1], ex) * (1.0 / (2 * m + 1));
            }
        } else {
            const double rt = intc_rsqrt(T), it = rt * rt;
            F[0] = 0.88622692545275801365 * rt;  // sqrt(pi) / 2
            if (M > 0) {
                double h = 0.5 * it;
#pragma unroll
                for (int m = 0; m < M; ++m) F[m + 1] = fma((double)(2 * m + 1), F[m], -ex) * h;
            }
        }
    } else {
        const double oof[7] = {1.0, 1.0, 1.0 / 2.0, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0, 1.0 / 720.0};
        if (T <= 15.0) {
            int i = (int)round(T * (1.0 / 0.1));
            double dx = (double)i * 0.1 - T, xn = 1.0;
            const double *row = a.boys_small + i * 33;
            double f = row[M];
#pragma unroll
            for (int ii = 1; ii < 7; ++ii) { xn = xn * dx; f = f + row[M + ii] * xn * oof[ii]; }
            F[M] = f;
            double ex = exp(-T);
#pragma unroll
            for (int m = M - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + ex) / (2.0 * (double)m + 1.0);
        } else if (T < 35.0) {
            int i = (int)round(T * (1.0 / 0.1));
            double dx = (double)i * 0.1 - T;
            i -= 150;
            double oox = 1.0 / T, xn = 1.0;
            const double *row = a.boys_medium + i * 7;
            double f = row[0];
#pragma unroll
            for (int ii = 1; ii < 7; ++ii) { xn = xn * dx; f = f + row[ii] * xn * oof[ii]; }
            F[0] = f;
            double ex = exp(-T);
#pragma unroll
            for (int m = 0; m < M; ++m) F[m + 1] = ((double)m + 0.5) * oox * F[m] - ex * 0.5 * oox;
        } else {
            double oox = 1.0 / T;
            F[0] = 0.5 * sqrt(3.141592653589793238 * oox);
#pragma unroll
            for (int m = 0; m < M; ++m) F[m + 1] = ((double)m + 0.5) * oox * F[m];
        }
    }
}
