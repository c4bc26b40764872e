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
#include <utility>

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


// Compile-time layout of the recurrences of canonical class (LA LB | LC LD)
// (every size below is a constant of the kernel instantiation):
//   ket level kl stores (e|f)^(m) for |f| = kl, m <= Lk - kl and |e| >= elev(kl),
//   the lowest |e| a later level or the accumulation can read:
//   elev(kl) = max(0, min(LA, LB) - (Lk - kl)) (either orientation of the pair);
//   layout [e - e0(kl)][f - ncc0(kl - 1)][m]; level 0 is (e|0)^(m), m <= Lk.
//   Work items of level kl are its (e, f) pairs, item j = e_rel * nf + f_rel,
//   each computing its m-vector; thread t takes items t + 256 r.
template <int LA, int LB, int LC, int LD>
struct Wg2 {
    static constexpr int L = LA + LB, Lk = LC + LD, M = L + Lk;
    static constexpr int nAll = ncc0(L);
    static constexpr int lminb = LA < LB ? LA : LB, lmink = LC < LD ? LC : LD;
    static constexpr int elev(int kl) { return lminb - (Lk - kl) > 0 ? lminb - (Lk - kl) : 0; }
    static constexpr int e0(int kl) { return ncc0(elev(kl) - 1); }
    static constexpr int ne(int kl) { return nAll - e0(kl); }
    static constexpr int nf(int kl) { return ncc(kl); }
    static constexpr int nm(int kl) { return Lk - kl + 1; }
    static constexpr int items(int kl) { return ne(kl) * nf(kl); }
    static constexpr int size(int kl) { return items(kl) * nm(kl); }
    static constexpr int reps(int kl) { return (items(kl) + WG_BLOCK - 1) / WG_BLOCK; }
    static constexpr int cap()
    {
        int c = 0;
        for (int kl = 0; kl <= Lk; ++kl) c = size(kl) > c ? size(kl) : c;
        return c;
    }
    // accumulator slots of levels >= max(1, lmink), in level order
    static constexpr int acc_lo = lmink > 1 ? lmink : 1;
    static constexpr int acc_base(int kl)
    {
        int b = 0;
        for (int k2 = acc_lo; k2 < kl; ++k2) b += reps(k2);
        return b;
    }
    static constexpr int nacc() { return acc_base(Lk + 1) > 0 ? acc_base(Lk + 1) : 1; }
    // setup slot: c0_m (m <= M), PA, WP, QC, WQ (3 each), roz, roe, o2s, o2z, o2e, pref
    static constexpr int NSLOT = M + 1 + 18;
    static constexpr int CH = 64;   // primitive quartets per setup chunk
};

// per primitive quartet values the levels read, converted once per quartet
template <typename VT>
struct Wg2Prim {
    VT QC[3], WQ[3], roe, o2e, o2s;
    double pref;
};

// One ket level: items j = threadIdx.x + 256 r (r unrolled) compute their
// m-vector from levels KL-1 and KL-2 (same e) and level KL-1 at e - 1_i:
//   (e|f+1_i)^(m) = QC_i (e|f)^(m) + WQ_i (e|f)^(m+1)
//                 + f_i/(2 eta) [(e|f-1_i)^(m) - rho/eta (e|f-1_i)^(m+1)]
//                 + e_i/(2(zeta+eta)) (e-1_i|f)^(m+1)
// and, from level max(1, lmin_k) on, add m = 0 to the thread's accumulators
// (only items with la' <= |e| and lc' <= KL count for the oriented class).
template <class W, int KL, typename VT>
__device__ __forceinline__ void wg2_level(const WgTab &tb, VT *const lv[3], const Wg2Prim<VT> &pq, double *acc,
                                          int e0o, int lco)
{
    constexpr int NF = W::nf(KL), NM = W::nm(KL);
    constexpr int NF1 = W::nf(KL - 1), NM1 = W::nm(KL - 1);
    constexpr int NF2 = KL >= 2 ? W::nf(KL - 2) : 1, NM2 = KL >= 2 ? W::nm(KL - 2) : 1;
    constexpr int E0 = W::e0(KL), E01 = W::e0(KL - 1), E02 = KL >= 2 ? W::e0(KL - 2) : 0;
    constexpr int G0 = ncc0(KL - 1), F10 = ncc0(KL - 2), F20 = ncc0(KL - 3);
    VT *Kc = lv[KL % 3];
    const VT *K1 = lv[(KL - 1) % 3];
    const VT *K2 = lv[(KL + 1) % 3];   // level KL - 2
#pragma unroll
    for (int r = 0; r < W::reps(KL); ++r) {
        const int j = threadIdx.x + WG_BLOCK * r;
        if (j < W::items(KL)) {
            const int ei = j / NF, fi = j - ei * NF;
            const int ef = E0 + ei, gf = G0 + fi;
            const int i = tb.dir[gf];
            const int f1 = tb.down[3 * gf + i];
            const int cf = tb.xyz[3 * gf + i] - 1;
            const int ce = tb.xyz[3 * ef + i];
            const VT *R1 = K1 + ((ef - E01) * NF1 + (f1 - F10)) * NM1;
            const VT qc = i == 0 ? pq.QC[0] : (i == 1 ? pq.QC[1] : pq.QC[2]);
            const VT wq = i == 0 ? pq.WQ[0] : (i == 1 ? pq.WQ[1] : pq.WQ[2]);
            VT v[NM];
#pragma unroll
            for (int m = 0; m < NM; ++m) v[m] = qc * R1[m] + wq * R1[m + 1];
            if (KL >= 2 && cf > 0) {
                const int f2 = tb.down[3 * f1 + i];
                const VT *R2 = K2 + ((ef - E02) * NF2 + (f2 - F20)) * NM2;
                const VT c = (VT)cf * pq.o2e;
#pragma unroll
                for (int m = 0; m < NM; ++m) v[m] += c * (R2[m] - pq.roe * R2[m + 1]);
            }
            if (ce > 0) {
                const int em = tb.down[3 * ef + i];
                const VT *R3 = K1 + ((em - E01) * NF1 + (f1 - F10)) * NM1 + 1;
                const VT c = (VT)ce * pq.o2s;
#pragma unroll
                for (int m = 0; m < NM; ++m) v[m] += c * R3[m];
            }
            VT *o = Kc + j * NM;
#pragma unroll
            for (int m = 0; m < NM; ++m) o[m] = v[m];
            if constexpr (KL >= W::acc_lo) {
                if (KL >= lco && ef >= e0o) {
                    if (sizeof(VT) == 4)
                        acc[W::acc_base(KL) + r] += pq.pref * (double)v[0];
                    else
                        acc[W::acc_base(KL) + r] += (double)v[0];
                }
            }
        }
    }
    __syncthreads();
}

template <class W, typename VT, int... KL>
__device__ __forceinline__ void wg2_levels(std::integer_sequence<int, KL...>, const WgTab &tb, VT *const lv[3],
                                           const Wg2Prim<VT> &pq, double *acc, int e0o, int lco)
{
    (wg2_level<W, KL + 1, VT>(tb, lv, pq, acc, e0o, lco), ...);
}

// the accumulated (e|f) of this thread's items into X [e - e0o][f - f0o]
template <class W, int KL>
__device__ __forceinline__ void wg2_store_level(const double *acc, double *Xb, int e0o, int f0o, int nFo, int lco)
{
    if constexpr (KL >= W::acc_lo) {
        constexpr int NF = W::nf(KL), E0 = W::e0(KL), G0 = ncc0(KL - 1);
#pragma unroll
        for (int r = 0; r < W::reps(KL); ++r) {
            const int j = threadIdx.x + WG_BLOCK * r;
            if (j < W::items(KL)) {
                const int ei = j / NF, fi = j - ei * NF;
                const int ef = E0 + ei, gf = G0 + fi;
                if (KL >= lco && ef >= e0o) Xb[(ef - e0o) * nFo + (gf - f0o)] = acc[W::acc_base(KL) + r];
            }
        }
    }
}

template <class W, int... KL>
__device__ __forceinline__ void wg2_store(std::integer_sequence<int, KL...>, const double *acc, double *Xb, int e0o,
                                          int f0o, int nFo, int lco)
{
    (wg2_store_level<W, KL + 1>(acc, Xb, e0o, f0o, nFo, lco), ...);
}

template <typename VT, int BM, int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(WG_BLOCK) intc_eri_wg_k(const IntcEriArgs a, const IntcWgCls k)
{
    using W = Wg2<LA, LB, LC, LD>;
    extern __shared__ __attribute__((aligned(16))) unsigned char wg_lds[];
    int *itab = (int *)(wg_lds + k.o_tab);
    wg_tables(itab);
    WgTab tb{itab, itab + 3 * WG_NCOMP, itab + 4 * WG_NCOMP, itab + 7 * WG_NCOMP};
    double *slots = (double *)(wg_lds + k.o_bra);
    VT *const lv[3] = {(VT *)(wg_lds + k.o_lvl), (VT *)(wg_lds + k.o_lvl) + W::cap(),
                       (VT *)(wg_lds + k.o_lvl) + 2 * W::cap()};
    double *Zb = (double *)(wg_lds + k.o_Z);
    double *Xb = (double *)(wg_lds + k.o_X);
    double *hk = (double *)(wg_lds + k.o_hk);
    double *hb = (double *)(wg_lds + k.o_hb);
    __syncthreads();

    constexpr int L = W::L, Lk = W::Lk;

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
        const double Ax = Vb[0], Ay = Vb[1], Az = Vb[2];
        const double Cx = Vk[0], Cy = Vk[1], Cz = Vk[2];
        double acc[W::nacc()];
#pragma unroll
        for (int r = 0; r < W::nacc(); ++r) acc[r] = 0.0;
        double acc0 = 0.0;   // level 0 (an s-shell ket VRR index), by the chain thread of e
        // this thread's bra component for the chains: e = flat index threadIdx.x
        const int te = threadIdx.x;
        int cx = 0, cy = 0, cz = 0;
        if (te < W::nAll) { cx = tb.xyz[3 * te]; cy = tb.xyz[3 * te + 1]; cz = tb.xyz[3 * te + 2]; }
        const int P = Pb * Pk;
        for (int c0p = 0; c0p < P; c0p += W::CH) {
            const int np = min(W::CH, P - c0p);
            __syncthreads();   // the previous chunk's slots are read
            // setup of every primitive quartet of the chunk in parallel: Boys,
            // prefactor, recurrence coefficients (VRR centres: the oriented A', C')
            for (int t = threadIdx.x; t < np; t += WG_BLOCK) {
                const int pq = c0p + t;
                const int ib = pq / Pk, ik = pq - ib * Pk;
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
                double F[W::M + 1];
                intc_boys<W::M, BM>(Tb, F, a);
                double *u = slots + (size_t)t * W::NSLOT;
                double s = 1.0;
#pragma unroll
                for (int m = 0; m <= W::M; ++m) {
                    u[m] = f32 ? F[m] * s : pref * F[m];
                    s *= sc;
                }
                double *w = u + W::M + 1;
                w[0] = Px - Ax; w[1] = Py - Ay; w[2] = Pz - Az;
                w[3] = -eo * PQx * isc; w[4] = -eo * PQy * isc; w[5] = -eo * PQz * isc;
                w[6] = Qx - Cx; w[7] = Qy - Cy; w[8] = Qz - Cz;
                w[9] = zo * PQx * isc; w[10] = zo * PQy * isc; w[11] = zo * PQz * isc;
                w[12] = eo * isc; w[13] = zo * isc; w[14] = 0.5 * os * isc; w[15] = o2z; w[16] = o2e;
                // the prefactor the FP32 route multiplies at accumulation, after the 17 entries
                w[17] = pref;
            }
            __syncthreads();
            for (int pl = 0; pl < np; ++pl) {
                const double *u = slots + (size_t)pl * W::NSLOT;
                const double *w = u + W::M + 1;
                Wg2Prim<VT> pq;
#pragma unroll
                for (int d = 0; d < 3; ++d) { pq.QC[d] = (VT)w[6 + d]; pq.WQ[d] = (VT)w[9 + d]; }
                pq.roe = (VT)w[13];
                pq.o2s = (VT)w[14];
                pq.o2e = (VT)w[16];
                pq.pref = w[17];
                // bra chains: thread e builds (e|0)^(m), m <= M - |e|, along the
                // direction rule's chain (z steps, then y, then x), in registers
                if (te < W::nAll) {
                    VT cm2[W::M + 1], cm1[W::M + 1], cc[W::M + 1];
#pragma unroll
                    for (int m = 0; m <= W::M; ++m) { cm1[m] = (VT)u[m]; cm2[m] = (VT)0; }
                    const VT o2zv = (VT)w[15], roz = (VT)w[12];
                    int s = 0;
                    for (int d = 2; d >= 0; --d) {
                        const int nd = d == 0 ? cx : (d == 1 ? cy : cz);
                        const VT pa = (VT)(d == 0 ? w[0] : (d == 1 ? w[1] : w[2]));
                        const VT wp = (VT)(d == 0 ? w[3] : (d == 1 ? w[4] : w[5]));
                        for (int k2 = 0; k2 < nd; ++k2) {
                            ++s;
                            const VT kk = (VT)k2 * o2zv;
#pragma unroll
                            for (int m = 0; m < W::M; ++m)
                                cc[m] = pa * cm1[m] + wp * cm1[m + 1] + kk * (cm2[m] - roz * cm2[m + 1]);
                            cc[W::M] = (VT)0;
#pragma unroll
                            for (int m = 0; m <= W::M; ++m) { cm2[m] = cm1[m]; cm1[m] = cc[m]; }
                        }
                    }
                    (void)s;
                    if (te >= W::e0(0)) {
                        VT *o = lv[0] + (te - W::e0(0)) * (W::Lk + 1);
#pragma unroll
                        for (int m = 0; m <= W::Lk; ++m) o[m] = cm1[m];
                    }
                    if (lc == 0 && te >= e0) acc0 += sizeof(VT) == 4 ? pq.pref * (double)cm1[0] : (double)cm1[0];
                }
                __syncthreads();
                wg2_levels<W, VT>(std::make_integer_sequence<int, W::Lk>{}, tb, lv, pq, acc, e0, lc);
            }
        }
        __syncthreads();  // VRR buffers are reused below

        // contracted (e|f), E x F, into LDS
        wg2_store<W>(std::make_integer_sequence<int, W::Lk>{}, acc, Xb, e0, f0, nF, lc);
        if (lc == 0 && te < W::nAll && te >= e0) Xb[(te - e0) * nF] = acc0;
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

// Host: the LDS layout of class (LA LB | LC LD), from the kernel's own
// compile-time sizes (returns -1 if it needs more than max_lds).  Loop phase
// after the component tables: CH setup slots, three ket-level buffers;
// epilogue phase (reusing that space): Z, X and the HRR buffers.
template <int LA, int LB, int LC, int LD>
static int wg2_layout(int sph, int vt_bytes, int max_lds, IntcWgCls *k)
{
    using W = Wg2<LA, LB, LC, LD>;
    IntcWgCls c{};
    c.la = LA; c.lb = LB; c.lc = LC; c.ld = LD;
    c.L = W::L; c.Lk = W::Lk; c.M = W::M; c.sph = sph;
    // the kernel may flip either pair (VRR on the other shell): size the
    // epilogue buffers for the larger of the two orientations
    const int nE = std::max(ncc0(c.L) - ncc0(LA - 1), ncc0(c.L) - ncc0(LB - 1));
    const int nF = std::max(ncc0(c.Lk) - ncc0(LC - 1), ncc0(c.Lk) - ncc0(LD - 1));
    c.e0 = ncc0(LA - 1); c.nE = nE;
    c.f0 = ncc0(LC - 1); c.nF = nF;
    c.nAll = W::nAll;
    c.lvl_cap = W::cap();
    c.nZ = sph ? nsph(LC) * nsph(LD) : ncc(LC) * ncc(LD);
    auto hcap = [](int L, int lo, int mv) {
        int cap = 1;
        for (int j = 1; j <= mv; ++j) cap = std::max(cap, (ncc0(L - j) - ncc0(lo - 1)) * ncc(j));
        return cap;
    };
    c.hk_cap = std::max(hcap(c.Lk, LC, LD), hcap(c.Lk, LD, LC));
    c.hb_cap = std::max(hcap(c.L, LA, LB), hcap(c.L, LB, LA));
    c.ec = 16;
    c.cc = 16;
    auto al = [](int x) { return (x + 15) & ~15; };
    c.o_tab = 0;
    const int base = al(10 * WG_NCOMP * 4);
    c.o_bra = base;   // setup slots
    c.o_lvl = al(c.o_bra + W::CH * W::NSLOT * 8);
    const int end1 = al(c.o_lvl + 3 * W::cap() * vt_bytes);
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

template <typename VT, int LA, int LB, int LC, int LD>
static hipError_t wg2_launch(const IntcEriArgs *a, int grid, hipStream_t st)
{
    IntcWgCls k;
    if (wg2_layout<LA, LB, LC, LD>(a->wg_cls[4], (int)sizeof(VT), 160 * 1024, &k)) return hipErrorInvalidValue;
    if (a->boys_mode == 0)
        hipLaunchKernelGGL((intc_eri_wg_k<VT, 0, LA, LB, LC, LD>), dim3(grid), dim3(WG_BLOCK), k.lds_bytes, st, *a, k);
    else
        hipLaunchKernelGGL((intc_eri_wg_k<VT, 1, LA, LB, LC, LD>), dim3(grid), dim3(WG_BLOCK), k.lds_bytes, st, *a, k);
    return hipGetLastError();
}

int intc_wg_prepare()
{
#define INTC_WG_CLASS(A, B, C, D)                                                                            \
    {                                                                                                        \
        const void *fns[] = {(const void *)intc_eri_wg_k<double, 0, A, B, C, D>,                             \
                             (const void *)intc_eri_wg_k<double, 1, A, B, C, D>,                             \
                             (const void *)intc_eri_wg_k<float, 0, A, B, C, D>,                              \
                             (const void *)intc_eri_wg_k<float, 1, A, B, C, D>};                             \
        for (const void *f : fns)                                                                            \
            if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess) \
                return -1;                                                                                   \
    }
#include "gen/wg_classes.inc"
#undef INTC_WG_CLASS
    return 0;
}

// the workgroup kernel of the bin's canonical class (a->wg_cls)
template <typename VT>
static hipError_t wg_launch(const IntcEriArgs *a, int grid, int block, size_t lds, hipStream_t st)
{
    (void)block;
    (void)lds;
    const int la = a->wg_cls[0], lb = a->wg_cls[1], lc = a->wg_cls[2], ld = a->wg_cls[3];
#define INTC_WG_CLASS(A, B, C, D) \
    if (la == A && lb == B && lc == C && ld == D) return wg2_launch<VT, A, B, C, D>(a, grid, st);
#include "gen/wg_classes.inc"
#undef INTC_WG_CLASS
    return hipErrorInvalidValue;
}

extern "C" hipError_t intc_eri_wg_d_launch(const IntcEriArgs *a, int grid, int block, size_t lds, hipStream_t st)
{
    return wg_launch<double>(a, grid, block, lds, st);
}

extern "C" hipError_t intc_eri_wg_f_launch(const IntcEriArgs *a, int grid, int block, size_t lds, hipStream_t st)
{
    return wg_launch<float>(a, grid, block, lds, st);
}
