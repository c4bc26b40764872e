// intc_eri_wg.hip -- workgroup-per-quartet ERI kernels for classes with an f
// shell (4-centre, every shell l <= 3: bra and ket angular momentum up to 6,
// M up to 12).  The register kernels hold one bra component per lane and all
// ket components per lane; an (ff|ff) quartet would need 84 lanes per group and
// 84 ket components per lane, so these classes run on a whole workgroup with
// the recurrence tables in LDS.
//
// Per contracted quartet (one 256-thread workgroup, persistent over the bin):
//   for each primitive quartet (same sequence as the reference schedule, HGP
//   variant, cf. contracted_vrr_hrr_with_auxiliary.py and
//   integral_wrapper_rr_function.py:88-118):
//     thread 0: Boys F_m(T), m <= M (intc_boys, Chebyshev or reference tables)
//     bra levels n = 1..L:   (e|0)^(m), m <= M - n                 [OS VRR on a]
//     ket levels k = 1..Lk:  (e|f)^(m), |e| <= L, |f| = k, m <= Lk - k
//        (e|f+1_i)^(m) = QC_i (e|f)^(m) + WQ_i (e|f)^(m+1)
//                      + f_i/(2 eta) [(e|f-1_i)^(m) - rho/eta (e|f-1_i)^(m+1)]
//                      + e_i/(2(zeta+eta)) (e-1_i|f)^(m+1)
//        three f-levels live in LDS at a time; m = 0 entries with
//        la <= |e| <= L, |f| >= lc are summed into FP64 accumulators held in
//        registers (primitive contraction, cf. generator_contract_function.py)
//   ket HRR (HGP, on contracted values): (e|c,d+1_i) = (e|c+1_i,d) + CD_i (e|c,d)
//   ket Cartesian -> spherical: FP64 MFMA GEMM with S_c (x) S_d (caller's S)
//   bra HRR:  (a,b+1_i| = (a+1_i,b| + AB_i (a,b|, on the ket-transformed columns
//   bra Cartesian -> spherical: FP64 MFMA GEMM with S_a (x) S_b, written to
//   the output with the quartet's strides (generator_hrr_function.py:265-435,
//   generator_sph_trans_function.py:69-125, contracted_general.py:1808-1875)
//
// VT = double: plain FP64.  VT = float: the VRR tables in FP32 on s^m-scaled
// values (s = max(1, 2T); every m+1 term carries 1/s), Boys, prefactors and
// the contraction sums in FP64, HRR and transforms in FP64.
#include "intc_common.h"

#define INTC_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr int WG_NCOMP = 120;   // components with |e| <= 7 (tables cover e + 1_i for |e| <= 6)
constexpr int WG_NR = 22;       // accumulators per thread: ceil(74 * 74 / 256) for (ff|ff)
constexpr int WG_BLOCK = 256;

typedef double f64x4 __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int ncc(int n) { return n < 0 ? 0 : (n + 1) * (n + 2) / 2; }
__host__ __device__ constexpr int ncc0(int n) { return n < 0 ? 0 : (n + 1) * (n + 2) * (n + 3) / 6; }
__host__ __device__ constexpr int nsph(int n) { return 2 * n + 1; }

}  // namespace

// per-class layout, computed on the host (intc_wg_layout)
struct IntcWgCls {
    int la, lb, lc, ld, L, Lk, M, sph;
    int e0, nE;        // E: bra components la <= |e| <= L, flat [e0, e0 + nE)
    int f0, nF;        // F: ket components lc <= |f| <= Lk, flat [f0, f0 + nF)
    int nAll;          // bra components |e| <= L
    int braoff[8];     // bra level n: ncc(n) * (M - n + 1) entries at braoff[n]
    int lvl_cap;       // entries per ket level buffer
    int nZ;            // ket columns after the ket transform
    int hk_cap, hb_cap;
    int ec, cc;        // row chunks of the ket and bra HRR
    // byte offsets in dynamic LDS
    int o_tab, o_bra, o_lvl, o_Z, o_X, o_hk, o_hb, lds_bytes;
};

namespace {

// LDS component tables: xyz, build direction (first non-zero), e - 1_i, e + 1_i
struct WgTab {
    int *xyz, *dir, *down, *up;
};

__device__ __forceinline__ int cflat(int x, int y, int z)
{
    const int n = x + y + z, yz = y + z;
    return ncc0(n - 1) + yz * (yz + 1) / 2 + z;
}

__device__ void wg_tables(int *base)
{
    for (int t = threadIdx.x; t < WG_NCOMP; t += blockDim.x) {
        int n = 0;
        while (ncc0(n) <= t) ++n;
        int i = t - ncc0(n - 1), cnt = 0, X = 0, Y = 0, Z = 0;
        for (int x = n; x >= 0 && cnt <= i; --x)
            for (int y = n - x; y >= 0; --y) {
                if (cnt == i) { X = x; Y = y; Z = n - x - y; }
                ++cnt;
            }
        int c[3] = {X, Y, Z};
        base[3 * t] = X;
        base[3 * t + 1] = Y;
        base[3 * t + 2] = Z;
        base[3 * WG_NCOMP + t] = X > 0 ? 0 : (Y > 0 ? 1 : (Z > 0 ? 2 : 0));
        for (int d = 0; d < 3; ++d) {
            int e[3] = {c[0], c[1], c[2]};
            e[d] -= 1;
            base[4 * WG_NCOMP + 3 * t + d] = c[d] > 0 ? cflat(e[0], e[1], e[2]) : -1;
            e[d] += 2;
            base[7 * WG_NCOMP + 3 * t + d] = n < 7 ? cflat(e[0], e[1], e[2]) : -1;
        }
    }
}

// C[r][n] = sum_k A(r, k) B(k, n) on v_mfma_f64_16x16x4f64, 16x16 tiles over
// the four waves.  Operand layout (lane l, checked on gfx950): A[i][k]
// i = l % 16, k = l / 16; B[k][j] j = l % 16, k = l / 16; D[i][j] (four
// doubles v = 0..3 per lane) j = l % 16, i = l / 16 + 4 v.
template <class FA, class FB, class FC>
__device__ __forceinline__ void wg_gemm(int R, int K, int N, FA A, FB B, FC C)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nrt = (R + 15) >> 4, nnt = (N + 15) >> 4;
    for (int tile = wave; tile < nrt * nnt; tile += WG_BLOCK / 64) {
        const int rt = tile / nnt, nt = tile - rt * nnt;
        const int i = rt * 16 + (lane & 15), j = nt * 16 + (lane & 15);
        f64x4 acc = {0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < K; k0 += 4) {
            const int k = k0 + (lane >> 4);
            const double av = (i < R && k < K) ? A(i, k) : 0.0;
            const double bv = (k < K && j < N) ? B(k, j) : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
        }
        for (int v = 0; v < 4; ++v) {
            const int r = rt * 16 + (lane >> 4) + 4 * v;
            if (r < R && j < N) C(r, j, acc[v]);
        }
    }
}

// HGP HRR for a chunk of R independent rows.  Level 0 is read through
// src(r, ci), ci over the components lo <= |c| <= lo + nmove; level j holds
// H[r][ci][di] for lo <= |c| <= lo + nmove - j, |d| = j:
//   H_j[c][d] = H_{j-1}[c + 1_i][d - 1_i] + Rv_i H_{j-1}[c][d - 1_i], i = dir(d).
// Returns the buffer with the final level [r][c (|c| = lo)][d (|d| = nmove)];
// for nmove = 0 returns nullptr (the source itself is final).
template <class FS>
__device__ double *wg_hrr(const WgTab &tb, int R, int lo, int nmove, const double Rv[3], FS src, double *buf0,
                          double *buf1, int cap)
{
    if (nmove == 0) return nullptr;
    const int c0 = ncc0(lo - 1);
    double *prev = nullptr, *cur = buf0;
    for (int j = 1; j <= nmove; ++j) {
        const int nC = ncc0(lo + nmove - j) - c0, nCp = ncc0(lo + nmove - j + 1) - c0;
        const int nd = ncc(j), ndp = ncc(j - 1), d0 = ncc0(j - 1), dp0 = ncc0(j - 2);
        const int n = R * nC * nd;
        for (int t = threadIdx.x; t < n; t += WG_BLOCK) {
            const int r = t / (nC * nd), rem = t - r * nC * nd;
            const int ci = rem / nd, di = rem - ci * nd;
            const int dfl = d0 + di, i = tb.dir[dfl];
            const int dm = tb.down[3 * dfl + i] - dp0;
            const int cup = tb.up[3 * (c0 + ci) + i] - c0;
            double hi, h0;
            if (j == 1) {
                hi = src(r, cup);
                h0 = src(r, ci);
            } else {
                hi = prev[(r * nCp + cup) * ndp + dm];
                h0 = prev[(r * nCp + ci) * ndp + dm];
            }
            cur[(r * nC + ci) * nd + di] = fma(Rv[i], h0, hi);
        }
        __syncthreads();
        prev = cur;
        cur = (cur == buf0) ? buf1 : buf0;
    }
    (void)cap;
    return prev;
}

template <int BM>
__device__ void wg_boys(int M, double T, double *F, const IntcEriArgs &a)
{
    switch (M) {
#define INTC_WG_BOYS(m) \
    case m: intc_boys<m, BM>(T, F, a); break;
        INTC_WG_BOYS(0) INTC_WG_BOYS(1) INTC_WG_BOYS(2) INTC_WG_BOYS(3) INTC_WG_BOYS(4) INTC_WG_BOYS(5)
        INTC_WG_BOYS(6) INTC_WG_BOYS(7) INTC_WG_BOYS(8) INTC_WG_BOYS(9) INTC_WG_BOYS(10) INTC_WG_BOYS(11)
        INTC_WG_BOYS(12)
#undef INTC_WG_BOYS
        default: break;
    }
}

template <typename VT, int BM>
__global__ void __launch_bounds__(WG_BLOCK) intc_eri_wg_k(const IntcEriArgs a, const IntcWgCls k)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char wg_lds[];
    int *itab = (int *)(wg_lds + k.o_tab);
    wg_tables(itab);
    WgTab tb{itab, itab + 3 * WG_NCOMP, itab + 4 * WG_NCOMP, itab + 7 * WG_NCOMP};
    VT *bra = (VT *)(wg_lds + k.o_bra);
    VT *lvl[3] = {(VT *)(wg_lds + k.o_lvl), (VT *)(wg_lds + k.o_lvl) + k.lvl_cap,
                  (VT *)(wg_lds + k.o_lvl) + 2 * k.lvl_cap};
    double *Zb = (double *)(wg_lds + k.o_Z);
    double *Xb = (double *)(wg_lds + k.o_X);
    double *hk = (double *)(wg_lds + k.o_hk);
    double *hb = (double *)(wg_lds + k.o_hb);
    __syncthreads();

    const int L = k.L, Lk = k.Lk, M = k.M;
    int aoff[WG_NR], alev[WG_NR];

    const int nq = a.nq_dev ? *a.nq_dev : a.nq;
    for (int qq = blockIdx.x; qq < nq; qq += gridDim.x) {
        const int q = a.q_idx ? a.q_idx[qq] : qq;
        if (q < 0 || q >= a.nq) continue;  // uniform over the workgroup
        const int pb = a.q_bra[q], pk = a.q_ket[q];
        const int ob = a.pair_off[pb], Pb = a.pair_np[pb];
        const int okk = a.pair_off[pk], Pk = a.pair_np[pk];
        // Orientation per pair: the VRR builds angular momentum on the centre
        // that the primitive-pair centres P lie closest to (|kappa|-weighted
        // mean of |P - X|^2) and the HRR moves it to the other centre.  Far
        // apart shells of very different exponents put P next to the tight
        // one; building on the diffuse one and transferring across |AB| would
        // cancel away many digits.  Every thread takes the same decision.
        bool fb, fk;
        double Vb[3], Vk[3], ABv[3], CDv[3];
        {
            const double *A = a.pair_A + 3 * pb, *AB = a.pair_AB + 3 * pb;
            const double *C = a.pair_A + 3 * pk, *CD = a.pair_AB + 3 * pk;
            double wa = 0.0, wb = 0.0, wc = 0.0, wd = 0.0;
            for (int i = 0; i < Pb; ++i) {
                const double w = fabs(a.pp_kappa[ob + i]);
                const double dx = a.pp_px[ob + i] - A[0], dy = a.pp_py[ob + i] - A[1], dz = a.pp_pz[ob + i] - A[2];
                wa += w * (dx * dx + dy * dy + dz * dz);
                const double ex = dx + AB[0], ey = dy + AB[1], ez = dz + AB[2];
                wb += w * (ex * ex + ey * ey + ez * ez);
            }
            for (int i = 0; i < Pk; ++i) {
                const double w = fabs(a.pp_kappa[okk + i]);
                const double dx = a.pp_px[okk + i] - C[0], dy = a.pp_py[okk + i] - C[1], dz = a.pp_pz[okk + i] - C[2];
                wc += w * (dx * dx + dy * dy + dz * dz);
                const double ex = dx + CD[0], ey = dy + CD[1], ez = dz + CD[2];
                wd += w * (ex * ex + ey * ey + ez * ez);
            }
            fb = wb < wa;
            fk = wd < wc;
            for (int d = 0; d < 3; ++d) {
                Vb[d] = fb ? A[d] - AB[d] : A[d];
                ABv[d] = fb ? -AB[d] : AB[d];
                Vk[d] = fk ? C[d] - CD[d] : C[d];
                CDv[d] = fk ? -CD[d] : CD[d];
            }
        }
        // oriented class: VRR index first in each pair
        const int la = fb ? k.lb : k.la, lb = fb ? k.la : k.lb;
        const int lc = fk ? k.ld : k.lc, ld = fk ? k.lc : k.ld;
        const int e0 = ncc0(la - 1), nE = ncc0(L) - e0;
        const int f0 = ncc0(lc - 1), nF = ncc0(Lk) - f0;
        const int nEF = nE * nF;
        // this thread's accumulators: entry j = threadIdx.x + 256 r of the E x F
        // table; its ket level and offset (m = 0) within that level's buffer
        for (int r = 0; r < WG_NR; ++r) {
            const int j = threadIdx.x + WG_BLOCK * r;
            alev[r] = -1;
            aoff[r] = 0;
            if (j < nEF) {
                const int ei = j / nF, fj = j - ei * nF;
                const int ef = e0 + ei, ff = f0 + fj;
                int kl = 0;
                while (ncc0(kl) <= ff) ++kl;
                const int fi = ff - ncc0(kl - 1);
                alev[r] = kl;
                aoff[r] = (ef * ncc(kl) + fi) * (Lk - kl + 1);
            }
        }
        const double Ax = Vb[0], Ay = Vb[1], Az = Vb[2];
        const double Cx = Vk[0], Cy = Vk[1], Cz = Vk[2];
        double acc[WG_NR];
        for (int r = 0; r < WG_NR; ++r) acc[r] = 0.0;

        for (int p = 0; p < Pb * Pk; ++p) {
            const int ib = p / Pk, ik = p - ib * Pk;
            const double zeta = a.pp_zeta[ob + ib], o2z = a.pp_o2z[ob + ib];
            const double Px = a.pp_px[ob + ib], Py = a.pp_py[ob + ib], Pz = a.pp_pz[ob + ib];
            const double kab = a.pp_kappa[ob + ib];
            const double eta = a.pp_zeta[okk + ik], o2e = a.pp_o2z[okk + ik];
            const double Qx = a.pp_px[okk + ik], Qy = a.pp_py[okk + ik], Qz = a.pp_pz[okk + ik];
            const double kcd = a.pp_kappa[okk + ik];
            const double os = 1.0 / (zeta + eta);
            const double rho = zeta * eta * os;
            const double PQx = Px - Qx, PQy = Py - Qy, PQz = Pz - Qz;
            const double Tb = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
            const double pref = kab * kcd * sqrt(os);
            const bool f32 = sizeof(VT) == 4;
            const double sc = (f32 && Tb > 0.5) ? 2.0 * Tb : 1.0, isc = 1.0 / sc;
            const double eo = eta * os, zo = zeta * os;
            const VT PA[3] = {(VT)(Px - Ax), (VT)(Py - Ay), (VT)(Pz - Az)};
            const VT WP[3] = {(VT)(-eo * PQx * isc), (VT)(-eo * PQy * isc), (VT)(-eo * PQz * isc)};
            const VT QC[3] = {(VT)(Qx - Cx), (VT)(Qy - Cy), (VT)(Qz - Cz)};
            const VT WQ[3] = {(VT)(zo * PQx * isc), (VT)(zo * PQy * isc), (VT)(zo * PQz * isc)};
            const VT roz = (VT)(eo * isc), roe = (VT)(zo * isc), o2s = (VT)(0.5 * os * isc);
            const VT o2zv = (VT)o2z, o2ev = (VT)o2e;
            // the previous primitive's last readers of bra / level buffers are done
            __syncthreads();
            if (threadIdx.x == 0) {
                double F[13];
                wg_boys<BM>(M, Tb, F, a);
                double s = 1.0;
                for (int m = 0; m <= M; ++m) {
                    bra[m] = f32 ? (VT)(F[m] * s) : (VT)(pref * F[m]);
                    s *= sc;
                }
            }
            __syncthreads();
            // bra levels: (e|0)^(m), e at level n, m <= M - n
            for (int n = 1; n <= L; ++n) {
                const int nm = M - n + 1, cnt = ncc(n) * nm, e0n = ncc0(n - 1);
                const int nm1 = M - n + 2, nm2 = M - n + 3;
                for (int t = threadIdx.x; t < cnt; t += WG_BLOCK) {
                    const int ei = t / nm, m = t - ei * nm;
                    const int ef = e0n + ei, i = tb.dir[ef];
                    const int e1 = tb.down[3 * ef + i];
                    const VT *B1 = bra + k.braoff[n - 1] + (e1 - ncc0(n - 2)) * nm1;
                    VT v = PA[i] * B1[m] + WP[i] * B1[m + 1];
                    const int ci = tb.xyz[3 * ef + i] - 1;
                    if (ci > 0) {
                        const int e2 = tb.down[3 * e1 + i];
                        const VT *B2 = bra + k.braoff[n - 2] + (e2 - ncc0(n - 3)) * nm2;
                        v += (VT)ci * o2zv * (B2[m] - roz * B2[m + 1]);
                    }
                    bra[k.braoff[n] + ei * nm + m] = v;
                }
                __syncthreads();
            }
            // ket level 0: (e|0)^(m) for every e with |e| <= L, m <= Lk
            {
                const int nm = Lk + 1, cnt = k.nAll * nm;
                VT *K0 = lvl[0];
                for (int t = threadIdx.x; t < cnt; t += WG_BLOCK) {
                    const int ef = t / nm, m = t - ef * nm;
                    int n = 0;
                    while (ncc0(n) <= ef) ++n;
                    K0[t] = bra[k.braoff[n] + (ef - ncc0(n - 1)) * (M - n + 1) + m];
                }
                __syncthreads();
            }
            // accumulate level 0 when the VRR shell of the ket is an s shell
            // (an (s, l) ket pair oriented with the VRR on the s shell, or (ss))
            if (lc == 0) {
                for (int r = 0; r < WG_NR; ++r)
                    if (alev[r] == 0) acc[r] += f32 ? pref * (double)lvl[0][aoff[r]] : (double)lvl[0][aoff[r]];
            }
            // ket levels
            for (int kl = 1; kl <= Lk; ++kl) {
                VT *Kc = lvl[kl % 3];
                const VT *K1 = lvl[(kl - 1) % 3];
                const VT *K2 = lvl[(kl + 1) % 3];  // level kl - 2
                const int nm = Lk - kl + 1, nm1 = nm + 1, nm2 = nm + 2;
                const int nf = ncc(kl), nf1 = ncc(kl - 1), nf2 = ncc(kl - 2);
                const int g0 = ncc0(kl - 1), f10 = ncc0(kl - 2), f20 = ncc0(kl - 3);
                const int cnt = k.nAll * nf * nm;
                for (int t = threadIdx.x; t < cnt; t += WG_BLOCK) {
                    const int ef = t / (nf * nm), rem = t - ef * nf * nm;
                    const int gi = rem / nm, m = rem - gi * nm;
                    const int gf = g0 + gi, i = tb.dir[gf];
                    const int f1 = tb.down[3 * gf + i];
                    const VT *R1 = K1 + (ef * nf1 + (f1 - f10)) * nm1;
                    VT v = QC[i] * R1[m] + WQ[i] * R1[m + 1];
                    const int cf = tb.xyz[3 * gf + i] - 1;
                    if (cf > 0) {
                        const int f2 = tb.down[3 * f1 + i];
                        const VT *R2 = K2 + (ef * nf2 + (f2 - f20)) * nm2;
                        v += (VT)cf * o2ev * (R2[m] - roe * R2[m + 1]);
                    }
                    const int ce = tb.xyz[3 * ef + i];
                    if (ce > 0) {
                        const int em = tb.down[3 * ef + i];
                        v += (VT)ce * o2s * K1[(em * nf1 + (f1 - f10)) * nm1 + m + 1];
                    }
                    Kc[t] = v;
                }
                __syncthreads();
                if (kl >= lc) {
                    for (int r = 0; r < WG_NR; ++r)
                        if (alev[r] == kl) acc[r] += f32 ? pref * (double)Kc[aoff[r]] : (double)Kc[aoff[r]];
                }
            }
        }
        __syncthreads();  // VRR buffers are reused below

        // contracted (e|f), E x F, into LDS
        for (int r = 0; r < WG_NR; ++r) {
            const int j = threadIdx.x + WG_BLOCK * r;
            if (j < nEF) Xb[j] = acc[r];
        }
        __syncthreads();

        const int ncd = ncc(lc) * ncc(ld), nab = ncc(la) * ncc(lb);
        const double *Sc = a.sph + intc_sph_off(lc), *Sd = a.sph + intc_sph_off(ld);
        const double *Sa = a.sph + intc_sph_off(la), *Sb = a.sph + intc_sph_off(lb);
        const int ncl = ncc(lc), ndl = ncc(ld), nsd = nsph(ld);
        const int nal = ncc(la), nbl = ncc(lb), nsb = nsph(lb);

        // ket HRR + ket transform, e in chunks of ec rows
        for (int ec0 = 0; ec0 < nE; ec0 += k.ec) {
            const int R = min(k.ec, nE - ec0);
            const double *Xr = Xb + ec0 * nF;
            double *fin = wg_hrr(tb, R, lc, ld, CDv, [&](int r, int ci) { return Xr[r * nF + ci]; }, hk,
                                 hk + k.ec * k.hk_cap, k.hk_cap);
            const double *Y = fin ? fin : Xr;
            const int ys = fin ? ncd : nF;  // row stride of the Cartesian (e|cd) rows
            if (k.sph) {
                wg_gemm(
                    R, ncd, k.nZ, [&](int r, int kk) { return Y[r * ys + kk]; },
                    [&](int kk, int n) {
                        const int kc = kk / ndl, kd = kk - kc * ndl, nc = n / nsd, nd = n - nc * nsd;
                        return Sc[kc + nc * ncl] * Sd[kd + nd * ndl];
                    },
                    [&](int r, int n, double v) { Zb[(ec0 + r) * k.nZ + n] = v; });
            } else {
                for (int t = threadIdx.x; t < R * ncd; t += WG_BLOCK) {
                    const int r = t / ncd, kk = t - r * ncd;
                    Zb[(ec0 + r) * k.nZ + kk] = Y[r * ys + kk];
                }
            }
            __syncthreads();
        }

        // bra HRR + bra transform + output, columns in chunks of cc; indices
        // of a flipped pair are swapped back on the way out
        const long long qoff = a.q_off[q];
        const unsigned long long qs = a.q_str[q];
        const int s0 = (int)(qs & 0xffff), s1 = (int)((qs >> 16) & 0xffff);
        const int s2 = (int)((qs >> 32) & 0xffff), s3 = (int)((qs >> 48) & 0xffff);
        const long long sA = fb ? s1 : s0, sB = fb ? s0 : s1, sC = fk ? s3 : s2, sD = fk ? s2 : s3;
        const int nkd = k.sph ? nsd : ndl, nkb = k.sph ? nsb : nbl;
        double *out = a.out + qoff;
        for (int cc0 = 0; cc0 < k.nZ; cc0 += k.cc) {
            const int R = min(k.cc, k.nZ - cc0);
            const int nZ = k.nZ;
            double *fin = wg_hrr(tb, R, la, lb, ABv, [&](int r, int ai) { return Zb[ai * nZ + cc0 + r]; }, hb,
                                 hb + k.cc * k.hb_cap, k.hb_cap);
            auto Wv = [&](int r, int kk) { return fin ? fin[r * nab + kk] : Zb[kk * nZ + cc0 + r]; };
            auto store = [&](int r, int n, double v) {
                const int col = cc0 + r, ic = col / nkd, id = col - ic * nkd;
                const int ia = n / nkb, ib = n - ia * nkb;
                out[ia * sA + ib * sB + ic * sC + id * sD] = v;
            };
            if (k.sph) {
                wg_gemm(
                    R, nab, nsph(la) * nsb, Wv,
                    [&](int kk, int n) {
                        const int ka = kk / nbl, kb = kk - ka * nbl, na = n / nsb, nb = n - na * nsb;
                        return Sa[ka + na * nal] * Sb[kb + nb * nbl];
                    },
                    store);
            } else {
                for (int t = threadIdx.x; t < R * nab; t += WG_BLOCK) {
                    const int r = t / nab, kk = t - r * nab;
                    store(r, kk, Wv(r, kk));
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace

// Host: the LDS layout of a class (returns -1 if it needs more than max_lds)
int intc_wg_layout(int la, int lb, int lc, int ld, int sph, int vt_bytes, int max_lds, IntcWgCls *k)
{
    IntcWgCls c{};
    c.la = la; c.lb = lb; c.lc = lc; c.ld = ld;
    c.L = la + lb; c.Lk = lc + ld; c.M = c.L + c.Lk; c.sph = sph;
    if (c.L > 6 || c.Lk > 6 || c.M > 12) return -1;
    // the kernel may flip either pair (VRR on the other shell): size every
    // buffer for the larger of the two orientations
    const int nE = std::max(ncc0(c.L) - ncc0(la - 1), ncc0(c.L) - ncc0(lb - 1));
    const int nF = std::max(ncc0(c.Lk) - ncc0(lc - 1), ncc0(c.Lk) - ncc0(ld - 1));
    c.e0 = ncc0(la - 1); c.nE = nE;
    c.f0 = ncc0(lc - 1); c.nF = nF;
    c.nAll = ncc0(c.L);
    if (nE * nF > WG_NR * WG_BLOCK) return -1;
    int o = 0;
    for (int n = 0; n <= c.L; ++n) {
        c.braoff[n] = o;
        o += ncc(n) * (c.M - n + 1);
    }
    const int bra_total = o;
    c.lvl_cap = c.nAll * (c.Lk + 1);
    for (int kl = 1; kl <= c.Lk; ++kl) c.lvl_cap = std::max(c.lvl_cap, c.nAll * ncc(kl) * (c.Lk - kl + 1));
    c.nZ = sph ? nsph(lc) * nsph(ld) : ncc(lc) * ncc(ld);
    auto hcap = [](int L, int lo, int mv) {
        int cap = 1;
        for (int j = 1; j <= mv; ++j) cap = std::max(cap, (ncc0(L - j) - ncc0(lo - 1)) * ncc(j));
        return cap;
    };
    c.hk_cap = std::max(hcap(c.Lk, lc, ld), hcap(c.Lk, ld, lc));
    c.hb_cap = std::max(hcap(c.L, la, lb), hcap(c.L, lb, la));
    c.ec = 16;
    c.cc = 16;
    auto al = [](int x) { return (x + 15) & ~15; };
    c.o_tab = 0;
    const int base = al(10 * WG_NCOMP * 4);
    c.o_bra = base;
    c.o_lvl = al(c.o_bra + bra_total * vt_bytes);
    const int end1 = al(c.o_lvl + 3 * c.lvl_cap * vt_bytes);
    // after the primitive loop: Z, then X and the ket HRR buffers; the bra
    // HRR buffers reuse the space of X once the ket phase is done
    c.o_Z = base;
    c.o_X = al(c.o_Z + nE * c.nZ * 8);
    c.o_hk = al(c.o_X + nE * nF * 8);
    const int end2 = al(c.o_hk + 2 * c.ec * c.hk_cap * 8);
    c.o_hb = c.o_X;
    const int end3 = al(c.o_hb + 2 * c.cc * c.hb_cap * 8);
    c.lds_bytes = std::max(end1, std::max(end2, end3));
    *k = c;
    return c.lds_bytes <= max_lds ? 0 : -1;
}

int intc_wg_prepare()
{
    const void *fns[] = {(const void *)intc_eri_wg_k<double, 0>, (const void *)intc_eri_wg_k<double, 1>,
                         (const void *)intc_eri_wg_k<float, 0>, (const void *)intc_eri_wg_k<float, 1>};
    for (const void *f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess) return -1;
    return 0;
}

template <typename VT>
static hipError_t wg_launch(const IntcEriArgs *a, int grid, int block, size_t lds, hipStream_t st)
{
    (void)block;
    (void)lds;
    IntcWgCls k;
    if (intc_wg_layout(a->wg_cls[0], a->wg_cls[1], a->wg_cls[2], a->wg_cls[3], a->wg_cls[4], (int)sizeof(VT),
                       160 * 1024, &k))
        return hipErrorInvalidValue;
    if (a->boys_mode == 0)
        hipLaunchKernelGGL((intc_eri_wg_k<VT, 0>), dim3(grid), dim3(WG_BLOCK), k.lds_bytes, st, *a, k);
    else
        hipLaunchKernelGGL((intc_eri_wg_k<VT, 1>), dim3(grid), dim3(WG_BLOCK), k.lds_bytes, st, *a, k);
    return hipGetLastError();
}

extern "C" hipError_t intc_eri_wg_d_launch(const IntcEriArgs *a, int grid, int block, size_t lds, hipStream_t st)
{
    return wg_launch<double>(a, grid, block, lds, st);
}

extern "C" hipError_t intc_eri_wg_f_launch(const IntcEriArgs *a, int grid, int block, size_t lds, hipStream_t st)
{
    return wg_launch<float>(a, grid, block, lds, st);
}
