// intc_eri_wg.hip -- workgroup-per-quartet ERI kernels for classes with an f
// shell (4-centre, every shell l <= 3: bra and ket angular momentum up to 6,
// M up to 12) and (g s|g s).  The register kernels hold one bra component per
// lane and all ket components per lane; an (ff|ff) quartet would need 84 lanes
// per group and 84 ket components per lane, so these classes run on a whole
// workgroup with the recurrence tables in LDS.  One template instantiation per
// canonical class: every level size and item count is a compile-time constant.
//
// Per contracted quartet (one 256-thread workgroup; the grid strides over the
// bin), primitive quartets in the reference schedule's order (HGP variant, cf.
// contracted_vrr_hrr_with_auxiliary.py, integral_wrapper_rr_function.py:88-118):
//   setup, 64 primitive quartets at a time, one per thread: Boys F_m(T)
//     (intc_boys, Chebyshev or reference tables), prefactor, coefficients
//   bra chains: thread e builds (e|0)^(m) in registers            [OS VRR on a]
//   ket levels k = 1..Lk:  (e|f)^(m), |f| = k, m <= Lk - k
//        (e|f+1_i)^(m) = QC_i (e|f)^(m) + WQ_i (e|f)^(m+1)
//                      + f_i/(2 eta) [(e|f-1_i)^(m) - rho/eta (e|f-1_i)^(m+1)]
//                      + e_i/(2(zeta+eta)) (e-1_i|f)^(m+1)
//        three levels live in LDS at a time; each thread's items have fixed
//        descriptors (offsets of their three sources); m = 0 entries with
//        la <= |e| <= L, |f| >= lc are summed into FP64 accumulators in
//        registers (primitive contraction, cf. generator_contract_function.py)
//   epilogue on the contracted X[e][f] (HGP HRR in its closed binomial form,
//   generator_hrr_function.py:265-435, and the Cartesian -> spherical
//   transforms, generator_sph_trans_function.py:69-125, contracted_general.py:
//   1808-1875):
//        (e|c,d) = sum_{p <= d} prod_x C(d_x, p_x) CD_x^(d_x - p_x) (e|c+p, 0)
//        T[f][c'd'] = sum_{c <= f} S_c[c][c'] G_d[f - c][d'],
//        G_d[p][d'] = sum_d H_d[p][d] S_d[d][d'] (H_d the weights above)
//        Z = X T (FP64 MFMA), U[e][a'b'] likewise from AB, out = U^T Z
//        (FP64 MFMA) written with the quartet's strides
//
// VT = double: plain FP64.  VT = float: the VRR tables in FP32 on s^m-scaled
// values (s = max(1, 2T); every m+1 term carries 1/s), Boys, prefactors and
// the contraction sums in FP64, HRR and transforms in FP64.
#include <utility>

#include "intc_common.h"

#define INTC_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr int WG_NCOMP = 84;    // Cartesian components with |e| <= 6
constexpr int WG_BLOCK = 256;

typedef double f64x4 __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int ncc(int n) { return n < 0 ? 0 : (n + 1) * (n + 2) / 2; }
__host__ __device__ constexpr int ncc0(int n) { return n < 0 ? 0 : (n + 1) * (n + 2) * (n + 3) / 6; }
__host__ __device__ constexpr int nsph(int n) { return 2 * n + 1; }
__host__ __device__ constexpr int imax(int a, int b) { return a > b ? a : b; }

}  // namespace

// per-class LDS layout (byte offsets in dynamic LDS), computed on the host by
// wg2_layout from the kernel's compile-time sizes
struct IntcWgCls {
    int la, lb, lc, ld, sph;
    int o_tab;        // component tables
    int o_S;          // S_l, l <= max l of the class (the IntcEriArgs::sph layout)
    int o_bra, o_lvl; // loop phase: setup slots, three ket-level buffers
    int o_X, o_T, o_Z, o_H;   // epilogue phase: X, T (then U), Z, HRR weights
    int lds_bytes;
};

namespace {

// LDS component tables: xyz, build direction (first non-zero), e - 1_i
constexpr int WG_TAB_INTS = 7 * WG_NCOMP;
struct WgTab {
    int *xyz, *dir, *down;
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
        }
    }
}

// Cartesian components of shell l in the table order (x descending, then y)
template <int l>
struct WgComps {
    int x[ncc(l)], y[ncc(l)], z[ncc(l)];
    __host__ __device__ constexpr WgComps() : x{}, y{}, z{}
    {
        int c = 0;
        for (int a = l; a >= 0; --a)
            for (int b = l - a; b >= 0; --b) {
                x[c] = a;
                y[c] = b;
                z[c] = l - a - b;
                ++c;
            }
    }
};

// C[i][j] = sum_k A(i, k) B(k, j) for compile-time M x N x K, A(i, k) =
// A[i AI + k AK], B(k, j) = B[k BK + j BJ] in LDS, on v_mfma_f64_16x16x4f64,
// 16x16 tiles over the four waves.  Operand layout (lane l, checked on
// gfx950): A[i][k] i = l % 16, k = l / 16; B[k][j] j = l % 16, k = l / 16;
// D[i][j] (four doubles v = 0..3 per lane) j = l % 16, i = l / 16 + 4 v.
template <int M, int N, int K, int AI, int AK, int BK, int BJ, class FC>
__device__ __forceinline__ void wg3_gemm(const double *A, const double *B, FC store)
{
    constexpr int MT = (M + 15) / 16, NT = (N + 15) / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    for (int tile = wave; tile < MT * NT; tile += WG_BLOCK / 64) {
        const int rt = tile / NT, nt = tile - rt * NT;
        const int i = rt * 16 + li, j = nt * 16 + li;
        f64x4 acc = {0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < K; k0 += 4) {
            const int kk = k0 + lk;
            const double av = (i < M && kk < K) ? A[i * AI + kk * AK] : 0.0;
            const double bv = (kk < K && j < N) ? B[kk * BK + j * BJ] : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = rt * 16 + lk + 4 * v;
            if (r < M && j < N) store(r, j, acc[v]);
        }
    }
}

// Sizes of the epilogue of the oriented class (LA LB | LC LD): the VRR built
// a and c, the HRR moves b and d.  SPH: spherical output (else Cartesian).
template <int LA, int LB, int LC, int LD, bool SPH>
struct WgEpi {
    static constexpr int L = LA + LB, Lk = LC + LD;
    static constexpr int NSA = SPH ? nsph(LA) : ncc(LA), NSB = SPH ? nsph(LB) : ncc(LB);
    static constexpr int NSC = SPH ? nsph(LC) : ncc(LC), NSD = SPH ? nsph(LD) : ncc(LD);
    static constexpr int e0 = ncc0(LA - 1), nE = ncc0(L) - e0;    // X rows: la <= |e| <= L
    static constexpr int f0 = ncc0(LC - 1), nF = ncc0(Lk) - f0;   // X columns: lc <= |f| <= Lk
    static constexpr int nZ = NSC * NSD, nU = NSA * NSB;
    // HRR weights H[p][d], |p| <= l, |d| = l; with SPH also G = H S_l
    static constexpr int nHk = ncc0(LD) * ncc(LD), nHb = ncc0(LB) * ncc(LB);
    static constexpr int nGk = SPH ? ncc0(LD) * NSD : 0, nGb = SPH ? ncc0(LB) * NSB : 0;
    static constexpr int nH = nHk + nHb + nGk + nGb;
    // T and U in KCH column chunks (c' and a' split evenly) where X, T and Z
    // together would not fit beside the rest of the LDS
    static constexpr int KCH = (nE * nF + nF * nZ + nE * nZ) * 8 > 136 * 1024 ? 2 : 1;
    static_assert(NSC % KCH == 0 && NSA % KCH == 0, "even column chunks");
    static constexpr int NCPC = NSC / KCH, NCPA = NSA / KCH;
    static constexpr int nT = imax(nF * NCPC * NSD, nE * NCPA * NSB);   // T, later U
};

// prod_x C(d_x, p_x) v_x^(d_x - p_x), 0 unless p <= d: the weight of
// (e|c+p, 0) in (e|c, d) once the HRR has moved d to its centre
__device__ __forceinline__ double wg3_hrr_weight(const int *p, const int *d, double vx, double vy, double vz)
{
    double h = 1.0;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        const int px = p[x], dx = d[x];
        if (px > dx) return 0.0;
        const double v = x == 0 ? vx : (x == 1 ? vy : vz);
        int b = 1;
        for (int i = 0; i < px; ++i) b = b * (dx - i) / (i + 1);
        double pw = 1.0;
        for (int i = 0; i < dx - px; ++i) pw *= v;
        h *= (double)b * pw;
    }
    return h;
}

// Out[r][c' - cp0][d'] = sum_{c: |c| = L1, c <= f} S1[c][c'] G[f - c][d'] for
// the rows f = B0 + r, r < NR, and NCP columns c' from cp0 (G rows: the flat
// components p of |p| <= L2); for Cartesian output (no S) the term c = c'.
template <int L1, int L2, int B0, int NR, int NCP, int NS2, bool SPH>
__device__ __forceinline__ void wg3_build(const int *xyz, const double *S1, const double *G, double *Out, int cp0)
{
    constexpr WgComps<L1> cm{};
    for (int t = threadIdx.x; t < NR * NCP; t += WG_BLOCK) {
        const int r = t / NCP, cpl = t - r * NCP, cp = cp0 + cpl, f = B0 + r;
        const int fx = xyz[3 * f], fy = xyz[3 * f + 1], fz = xyz[3 * f + 2];
        const int gb = ncc0(fx + fy + fz - L1 - 1);
        double acc[NS2];
#pragma unroll
        for (int dp = 0; dp < NS2; ++dp) acc[dp] = 0.0;
        if constexpr (SPH) {
#pragma unroll
            for (int c = 0; c < ncc(L1); ++c) {
                const int dx = fx - cm.x[c], dy = fy - cm.y[c], dz = fz - cm.z[c];
                if (dx >= 0 && dy >= 0 && dz >= 0) {
                    const int g = gb + (dy + dz) * (dy + dz + 1) / 2 + dz;
                    const double s = S1[c + cp * ncc(L1)];
#pragma unroll
                    for (int dp = 0; dp < NS2; ++dp) acc[dp] = fma(s, G[g * NS2 + dp], acc[dp]);
                }
            }
        } else {
            const int c = ncc0(L1 - 1) + cp;
            const int dx = fx - xyz[3 * c], dy = fy - xyz[3 * c + 1], dz = fz - xyz[3 * c + 2];
            if (dx >= 0 && dy >= 0 && dz >= 0) {
                const int g = gb + (dy + dz) * (dy + dz + 1) / 2 + dz;
#pragma unroll
                for (int dp = 0; dp < NS2; ++dp) acc[dp] = G[g * NS2 + dp];
            }
        }
#pragma unroll
        for (int dp = 0; dp < NS2; ++dp) Out[r * (NCP * NS2) + cpl * NS2 + dp] = acc[dp];
    }
}

// HRR weights H[p][d] = wg3_hrr_weight(p, d) for |p| <= L2, |d| = L2
template <int L2>
__device__ __forceinline__ void wg3_weights(const int *xyz, double vx, double vy, double vz, double *H)
{
    for (int t = threadIdx.x; t < ncc0(L2) * ncc(L2); t += WG_BLOCK) {
        const int p = t / ncc(L2), d = ncc0(L2 - 1) + (t - p * ncc(L2));
        H[t] = wg3_hrr_weight(xyz + 3 * p, xyz + 3 * d, vx, vy, vz);
    }
}

// G[p][d'] = sum_d H[p][d] S[d][d'] (S_l column-major)
template <int L2, int NS2>
__device__ __forceinline__ void wg3_hs(const double *H, const double *S, double *G)
{
    for (int t = threadIdx.x; t < ncc0(L2) * NS2; t += WG_BLOCK) {
        const int p = t / NS2, dp = t - p * NS2;
        double g = 0.0;
#pragma unroll
        for (int q = 0; q < ncc(L2); ++q) g = fma(H[p * ncc(L2) + q], S[q + dp * ncc(L2)], g);
        G[t] = g;
    }
}

// The epilogue of one oriented class on the contracted X (LDS, k.o_X, row
// stride nF): ket HRR and transform as T, Z = X T; bra as U, out = U^T Z with
// the output strides of the oriented indices.  Shared by the kernel's FP64
// and FP32 forms; ends with a barrier (the LDS is reused after it).
template <int LA, int LB, int LC, int LD, bool SPH>
__device__ __noinline__ void wg3_epilogue(const IntcWgCls k, double abx, double aby, double abz, double cdx,
                                          double cdy, double cdz, double *out, long long sA, long long sB,
                                          long long sC, long long sD)
{
    using E = WgEpi<LA, LB, LC, LD, SPH>;
    extern __shared__ __attribute__((aligned(16))) unsigned char wg_lds[];
    const int *xyz = (const int *)(wg_lds + k.o_tab);
    const double *S = (const double *)(wg_lds + k.o_S);
    const double *X = (const double *)(wg_lds + k.o_X);
    double *T = (double *)(wg_lds + k.o_T);
    double *Z = (double *)(wg_lds + k.o_Z);
    double *Hk = (double *)(wg_lds + k.o_H), *Hb = Hk + E::nHk;
    double *Gk = SPH ? Hb + E::nHb : Hk;
    double *Gb = SPH ? Gk + E::nGk : Hb;
    wg3_weights<LD>(xyz, cdx, cdy, cdz, Hk);
    wg3_weights<LB>(xyz, abx, aby, abz, Hb);
    __syncthreads();
    if constexpr (SPH) {
        wg3_hs<LD, E::NSD>(Hk, S + intc_sph_off(LD), Gk);
        wg3_hs<LB, E::NSB>(Hb, S + intc_sph_off(LB), Gb);
        __syncthreads();
    }
    constexpr int NTC = E::NCPC * E::NSD, NUC = E::NCPA * E::NSB;   // columns per chunk
    for (int c0 = 0; c0 < E::NSC; c0 += E::NCPC) {
        wg3_build<LC, LD, E::f0, E::nF, E::NCPC, E::NSD, SPH>(xyz, S + intc_sph_off(LC), Gk, T, c0);
        __syncthreads();
        wg3_gemm<E::nE, NTC, E::nF, E::nF, 1, NTC, 1>(
            X, T, [&](int i, int j, double v) { Z[i * E::nZ + c0 * E::NSD + j] = v; });
        __syncthreads();
    }
    for (int a0 = 0; a0 < E::NSA; a0 += E::NCPA) {
        wg3_build<LA, LB, E::e0, E::nE, E::NCPA, E::NSB, SPH>(xyz, S + intc_sph_off(LA), Gb, T, a0);
        __syncthreads();
        wg3_gemm<NUC, E::nZ, E::nE, 1, NUC, E::nZ, 1>(T, Z, [&](int i, int j, double v) {
            const int ia = a0 + i / E::NSB, ib = i - (i / E::NSB) * E::NSB, ic = j / E::NSD, id = j - ic * E::NSD;
            out[ia * sA + ib * sB + ic * sC + id * sD] = v;
        });
        __syncthreads();
    }
}

// the epilogue of the orientation the quartet took (fb: bra pair flipped,
// fk: ket pair flipped); a pair of equal shells has one
template <int LA, int LB, int LC, int LD, bool SPH>
__device__ __forceinline__ void wg3_epi_any(bool fb, bool fk, const IntcWgCls &k, const double *ab, const double *cd,
                                            double *out, long long sA, long long sB, long long sC, long long sD)
{
    const bool b = fb && LA != LB, c = fk && LC != LD;
    if (!b && !c)
        wg3_epilogue<LA, LB, LC, LD, SPH>(k, ab[0], ab[1], ab[2], cd[0], cd[1], cd[2], out, sA, sB, sC, sD);
    else if (b && !c)
        wg3_epilogue<LB, LA, LC, LD, SPH>(k, ab[0], ab[1], ab[2], cd[0], cd[1], cd[2], out, sA, sB, sC, sD);
    else if (!b && c)
        wg3_epilogue<LA, LB, LD, LC, SPH>(k, ab[0], ab[1], ab[2], cd[0], cd[1], cd[2], out, sA, sB, sC, sD);
    else
        wg3_epilogue<LB, LA, LD, LC, SPH>(k, ab[0], ab[1], ab[2], cd[0], cd[1], cd[2], out, sA, sB, sC, sD);
}

// Compile-time layout of the recurrences of canonical class (LA LB | LC LD)
// (every size below is a constant of the kernel instantiation):
//   ket level kl stores (e|f)^(m) for |f| = kl, m <= Lk - kl and |e| >= elev(kl),
//   the lowest |e| a later level or the accumulation can read:
//   elev(kl) = max(0, min(LA, LB) - (Lk - kl)) (either orientation of the pair);
//   item j = (e - e0(kl)) * nf + (f - ncc0(kl - 1)), stored plane-major,
//   value (j, m) at m * items(kl) + j, so the 64 lanes of a wave that handle
//   consecutive items touch consecutive addresses (no bank conflicts on the
//   stores, few on the reads); level 0 is (e|0)^(m), m <= Lk.  Thread t
//   computes the m-vectors of items t + 256 r.
template <int LA, int LB, int LC, int LD>
struct Wg2 {
    static constexpr int L = LA + LB, Lk = LC + LD, M = L + Lk;
    static constexpr int nAll = ncc0(L);
    static_assert(ncc0(L) <= WG_NCOMP && ncc0(Lk) <= WG_NCOMP, "component tables cover |e|, |f| <= 6");
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
    // item descriptors: one per (ket level >= 1, rep), levels in order
    static constexpr int rep_base(int kl)
    {
        int b = 0;
        for (int k2 = 1; k2 < kl; ++k2) b += reps(k2);
        return b;
    }
    static constexpr int nrep() { return rep_base(Lk + 1) > 0 ? rep_base(Lk + 1) : 1; }
    // setup slot (doubles): the ket coefficients QC_x WQ_x QC_y WQ_y QC_z WQ_z
    // in the recurrence type (8 doubles reserved), c0_m (m <= M), then
    // PA (3), WP (3), roz, roe, o2s, o2z, o2e, pref; an even length keeps
    // every slot 16-byte aligned
    static constexpr int S_U = 8, S_W = 8 + M + 1;
    static constexpr int NSLOT = (S_W + 12 + 1) & ~1;
    static constexpr int CH = 64;   // primitive quartets per setup chunk
};

// Item descriptor of ket level KL, rep r of this thread (computed once per
// workgroup from the component tables), item indices within the levels:
//   x: R1 = (e|f-1_i) in level KL-1 (12 bits), R3 = (e-1_i|f-1_i)^(m+1) in
//      level KL-1 (12 bits), i = dir(f) (2 bits), f_i - 1 (3 bits), e_i (3 bits)
//   y: R2 = (e|f-2_i) in level KL-2 (12 bits), e (7 bits), f - ncc0(KL-1)
//      (5 bits)
// A term whose integer factor is zero reads item 0 of its level instead (a
// value of this primitive quartet, finite), so every item runs the same code.
template <class W, typename VT, int KL>
__device__ __forceinline__ void wg3_desc_level(const WgTab &tb, uint2 *d)
{
    constexpr int NF = W::nf(KL);
    constexpr int NF1 = W::nf(KL - 1), NM1 = W::nm(KL - 1);
    constexpr int NF2 = KL >= 2 ? W::nf(KL - 2) : 1;
    constexpr int E0 = W::e0(KL), E01 = W::e0(KL - 1), E02 = KL >= 2 ? W::e0(KL - 2) : 0;
    constexpr int G0 = ncc0(KL - 1), F10 = ncc0(KL - 2), F20 = ncc0(KL - 3);
    static_assert(W::cap() < 4096 && W::nAll <= 128 && NF <= 32, "descriptor fields");
#pragma unroll
    for (int r = 0; r < W::reps(KL); ++r) {
        const int j = threadIdx.x + WG_BLOCK * r;
        unsigned x = 0, y = 0;
        if (j < W::items(KL)) {
            const int ei = j / NF, fi = j - ei * NF;
            const int ef = E0 + ei, gf = G0 + fi;
            const int i = tb.dir[gf];
            const int f1 = tb.down[3 * gf + i];
            const int cf = tb.xyz[3 * gf + i] - 1;
            const int ce = tb.xyz[3 * ef + i];
            const int o1 = (ef - E01) * NF1 + (f1 - F10);
            const int o3 = ce > 0 ? (tb.down[3 * ef + i] - E01) * NF1 + (f1 - F10) : 0;
            const int o2 = (KL >= 2 && cf > 0) ? (ef - E02) * NF2 + (tb.down[3 * f1 + i] - F20) : 0;
            x = (unsigned)o1 | ((unsigned)o3 << 12) | ((unsigned)i << 24) | ((unsigned)cf << 26) | ((unsigned)ce << 29);
            y = (unsigned)o2 | ((unsigned)ef << 12) | ((unsigned)fi << 19);
        }
        d[W::rep_base(KL) + r] = make_uint2(x, y);
    }
}

template <class W, typename VT, int... KL>
__device__ __forceinline__ void wg3_desc(std::integer_sequence<int, KL...>, const WgTab &tb, uint2 *d)
{
    (wg3_desc_level<W, VT, KL + 1>(tb, d), ...);
}

// One ket level: items j = threadIdx.x + 256 r (r unrolled) compute their
// m-vector from levels KL-1 and KL-2 (same e) and level KL-1 at e - 1_i:
//   (e|f+1_i)^(m) = QC_i (e|f)^(m) + WQ_i (e|f)^(m+1)
//                 + f_i/(2 eta) [(e|f-1_i)^(m) - rho/eta (e|f-1_i)^(m+1)]
//                 + e_i/(2(zeta+eta)) (e-1_i|f)^(m+1)
// and, from level max(1, lmin_k) on, add m = 0 to the thread's accumulators
// (only items with la' <= |e| and lc' <= KL count for the oriented class).
// kc: this primitive quartet's QC_x WQ_x QC_y WQ_y QC_z WQ_z.
template <class W, int KL, typename VT>
__device__ __forceinline__ void wg3_level(VT *const lv[3], const VT *kc, VT roe, VT o2e, VT o2s, double pref,
                                          const uint2 *d, double *acc, int e0o, int lco)
{
    constexpr int NM = W::nm(KL);
    // plane-major levels: (item, m) at m * items + item
    constexpr int I0 = W::items(KL), I1 = W::items(KL - 1), I2 = KL >= 2 ? W::items(KL - 2) : 1;
    VT *Kc = lv[KL % 3];
    const VT *K1 = lv[(KL - 1) % 3];
    const VT *K2 = lv[(KL + 1) % 3];   // level KL - 2
#pragma unroll
    for (int r = 0; r < W::reps(KL); ++r) {
        const int j = threadIdx.x + WG_BLOCK * r;
        if (WG_BLOCK * (r + 1) <= W::items(KL) || j < W::items(KL)) {
            const unsigned x = d[W::rep_base(KL) + r].x, y = d[W::rep_base(KL) + r].y;
            const VT *R1 = K1 + (x & 0xfffu);
            const VT *R3 = K1 + ((x >> 12) & 0xfffu) + I1;   // (m+1) planes
            const int i = (int)((x >> 24) & 3u);
            const VT qc = kc[2 * i], wq = kc[2 * i + 1];
            const VT c3 = (VT)(int)(x >> 29) * o2s;
            VT v[NM];
#pragma unroll
            for (int m = 0; m < NM; ++m) v[m] = fma(wq, R1[(m + 1) * I1], qc * R1[m * I1]);
            if constexpr (KL >= 2) {
                const VT *R2 = K2 + (y & 0xfffu);
                const VT c2 = (VT)(int)((x >> 26) & 7u) * o2e;
#pragma unroll
                for (int m = 0; m < NM; ++m) v[m] = fma(c2, fma(-roe, R2[(m + 1) * I2], R2[m * I2]), v[m]);
            }
#pragma unroll
            for (int m = 0; m < NM; ++m) v[m] = fma(c3, R3[m * I1], v[m]);
            VT *o = Kc + j;
#pragma unroll
            for (int m = 0; m < NM; ++m) o[m * I0] = v[m];
            if constexpr (KL >= W::acc_lo) {
                const int ef = (int)((y >> 12) & 127u);
                const double t = sizeof(VT) == 4 ? pref * (double)v[0] : (double)v[0];
                acc[W::acc_base(KL) + r] += (KL >= lco && ef >= e0o) ? t : 0.0;
            }
        }
    }
    __syncthreads();
}

template <class W, typename VT, int... KL>
__device__ __forceinline__ void wg3_levels(std::integer_sequence<int, KL...>, VT *const lv[3], const VT *kc, VT roe,
                                           VT o2e, VT o2s, double pref, const uint2 *d, double *acc, int e0o, int lco)
{
    (wg3_level<W, KL + 1, VT>(lv, kc, roe, o2e, o2s, pref, d, acc, e0o, lco), ...);
}

// the accumulated (e|f) of this thread's items into X [e - e0o][f - f0o]
template <class W, int KL>
__device__ __forceinline__ void wg2_store_level(const double *acc, const uint2 *d, double *Xb, int e0o, int f0o,
                                                int nFo, int lco)
{
    if constexpr (KL >= W::acc_lo) {
        constexpr int G0 = ncc0(KL - 1);
#pragma unroll
        for (int r = 0; r < W::reps(KL); ++r) {
            const int j = threadIdx.x + WG_BLOCK * r;
            if (j < W::items(KL)) {
                const unsigned y = d[W::rep_base(KL) + r].y;
                const int ef = (int)((y >> 12) & 127u), gf = G0 + (int)(y >> 19);
                if (KL >= lco && ef >= e0o) Xb[(ef - e0o) * nFo + (gf - f0o)] = acc[W::acc_base(KL) + r];
            }
        }
    }
}

template <class W, int... KL>
__device__ __forceinline__ void wg2_store(std::integer_sequence<int, KL...>, const double *acc, const uint2 *d,
                                          double *Xb, int e0o, int f0o, int nFo, int lco)
{
    (wg2_store_level<W, KL + 1>(acc, d, Xb, e0o, f0o, nFo, lco), ...);
}

// Bra chain step S (|e| = S + 1) of a thread whose component e has ne =
// |e| steps, the first cz along z, then cy along y, the rest along x:
//   (e+1_d|0)^(m) = PA_d (e|0)^(m) + WP_d (e|0)^(m+1)
//                 + k/(2 zeta) [(e-1_d|0)^(m) - rho/zeta (e-1_d|0)^(m+1)]
// (k: earlier steps along d), m <= M - S - 1.  Even steps read cur = a,
// prev = b and write b; odd steps the other way round.  w: PA (3), WP (3).
template <int M, int S, int NS, typename VT>
__device__ __forceinline__ void wg3_chain(VT *a, VT *b, const double *w, int ne, int cz, int cy, VT o2z, VT roz)
{
    if constexpr (S < NS) {
        VT *cur = (S & 1) ? b : a, *prv = (S & 1) ? a : b;
        if (S < ne) {
            const int d = S < cz ? 2 : (S < cz + cy ? 1 : 0);
            const int k = d == 2 ? S : (d == 1 ? S - cz : S - cz - cy);
            const VT pa = (VT)w[d], wp = (VT)w[3 + d], kk = (VT)k * o2z;
#pragma unroll
            for (int m = 0; m < M - S; ++m)
                prv[m] = fma(pa, cur[m], fma(wp, cur[m + 1], kk * fma(-roz, prv[m + 1], prv[m])));
        }
        wg3_chain<M, S + 1, NS>(a, b, w, ne, cz, cy, o2z, roz);
    }
}

template <typename VT, int BM, int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(WG_BLOCK) intc_eri_wg_k(const IntcEriArgs a, const IntcWgCls k)
{
    using W = Wg2<LA, LB, LC, LD>;
    extern __shared__ __attribute__((aligned(16))) unsigned char wg_lds[];
    int *itab = (int *)(wg_lds + k.o_tab);
    wg_tables(itab);
    WgTab tb{itab, itab + 3 * WG_NCOMP, itab + 4 * WG_NCOMP};
    double *slots = (double *)(wg_lds + k.o_bra);
    VT *const lv[3] = {(VT *)(wg_lds + k.o_lvl), (VT *)(wg_lds + k.o_lvl) + W::cap(),
                       (VT *)(wg_lds + k.o_lvl) + 2 * W::cap()};
    double *Xb = (double *)(wg_lds + k.o_X);
    // the caller's S_l for the shells of this class, staged once
    constexpr int LMAX = imax(imax(LA, LB), imax(LC, LD));
    if (k.sph) {
        double *Sl = (double *)(wg_lds + k.o_S);
        for (int t = threadIdx.x; t < intc_sph_off(LMAX + 1); t += WG_BLOCK) Sl[t] = a.sph[t];
    }
    __syncthreads();

    constexpr int Lk = W::Lk;
    // this thread's ket-level item descriptors, for the whole launch
    uint2 desc[W::nrep()];
    wg3_desc<W, VT>(std::make_integer_sequence<int, W::Lk>{}, tb, desc);

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
        const int e0 = ncc0(la - 1);
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
                const double ros = intc_rsqrt(zeta + eta), os = ros * ros;
                const double rho = zeta * eta * os;
                const double PQx = Px - Qx, PQy = Py - Qy, PQz = Pz - Qz;
                const double Tb = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
                const double pref = kab * kcd * ros;
                const bool f32 = sizeof(VT) == 4;
                const double sc = (f32 && Tb > 0.5) ? 2.0 * Tb : 1.0, isc = 1.0 / sc;
                const double eo = eta * os, zo = zeta * os;
                double F[W::M + 1];
                intc_boys<W::M, BM>(Tb, F, a);
                double *slot = slots + (size_t)t * W::NSLOT;
                VT *kcw = (VT *)slot;
                kcw[0] = (VT)(Qx - Cx); kcw[1] = (VT)(zo * PQx * isc);
                kcw[2] = (VT)(Qy - Cy); kcw[3] = (VT)(zo * PQy * isc);
                kcw[4] = (VT)(Qz - Cz); kcw[5] = (VT)(zo * PQz * isc);
                double *u = slot + W::S_U;
                double s = 1.0;
#pragma unroll
                for (int m = 0; m <= W::M; ++m) {
                    u[m] = f32 ? F[m] * s : pref * F[m];
                    s *= sc;
                }
                double *w = slot + W::S_W;
                w[0] = Px - Ax; w[1] = Py - Ay; w[2] = Pz - Az;
                w[3] = -eo * PQx * isc; w[4] = -eo * PQy * isc; w[5] = -eo * PQz * isc;
                w[6] = eo * isc; w[7] = zo * isc; w[8] = 0.5 * os * isc; w[9] = o2z; w[10] = o2e;
                // the prefactor the FP32 route multiplies at accumulation
                w[11] = pref;
            }
            __syncthreads();
            for (int pl = 0; pl < np; ++pl) {
                const double *slot = slots + (size_t)pl * W::NSLOT;
                const double *u = slot + W::S_U;
                const double *w = slot + W::S_W;
                const double pref = w[11];
                // bra chains: thread e builds (e|0)^(m) along the direction
                // rule's chain (z steps, then y, then x) in registers, step s
                // building |e| = s + 1 for m <= M - s - 1 from the two previous
                // steps (alternating buffers); lanes past their |e| hold still
                if (te < W::nAll) {
                    VT ca[W::M + 1], cb[W::M + 1];
#pragma unroll
                    for (int m = 0; m <= W::M; ++m) { ca[m] = (VT)u[m]; cb[m] = (VT)0; }
                    const VT o2zv = (VT)w[9], roz = (VT)w[6];
                    const int ne = cx + cy + cz;
                    wg3_chain<W::M, 0, W::L>(ca, cb, w, ne, cz, cy, o2zv, roz);
                    if (te >= W::e0(0)) {
                        VT *o = lv[0] + (te - W::e0(0));
#pragma unroll
                        for (int m = 0; m <= W::Lk; ++m) o[m * W::items(0)] = (ne & 1) ? cb[m] : ca[m];
                    }
                    if (lc == 0 && te >= e0)
                        acc0 += sizeof(VT) == 4 ? pref * (double)((ne & 1) ? cb[0] : ca[0])
                                                : (double)((ne & 1) ? cb[0] : ca[0]);
                }
                __syncthreads();
                wg3_levels<W, VT>(std::make_integer_sequence<int, W::Lk>{}, lv, (const VT *)slot, (VT)w[7], (VT)w[10],
                                  (VT)w[8], pref, desc, acc, e0, lc);
            }
        }
        __syncthreads();  // VRR buffers are reused below

        // contracted (e|f), E x F, into LDS
        wg2_store<W>(std::make_integer_sequence<int, W::Lk>{}, acc, desc, Xb, e0, f0, nF, lc);
        if (lc == 0 && te < W::nAll && te >= e0) Xb[(te - e0) * nF] = acc0;
        __syncthreads();

        // HRR and transforms of the oriented class; indices of a flipped pair
        // are swapped back on the way out
        const unsigned long long qs = a.q_str[q];
        const int s0 = (int)(qs & 0xffff), s1 = (int)((qs >> 16) & 0xffff);
        const int s2 = (int)((qs >> 32) & 0xffff), s3 = (int)((qs >> 48) & 0xffff);
        const long long sA = fb ? s1 : s0, sB = fb ? s0 : s1, sC = fk ? s3 : s2, sD = fk ? s2 : s3;
        double *out = a.out + a.q_off[q];
        if (k.sph)
            wg3_epi_any<LA, LB, LC, LD, true>(fb, fk, k, ABv, CDv, out, sA, sB, sC, sD);
        else
            wg3_epi_any<LA, LB, LC, LD, false>(fb, fk, k, ABv, CDv, out, sA, sB, sC, sD);
    }
}

}  // namespace

// Host: the LDS layout of class (LA LB | LC LD), from the kernel's own
// compile-time sizes (returns -1 if it needs more than max_lds).  After the
// component tables and the staged S_l: loop phase, CH setup slots and three
// ket-level buffers; epilogue phase (reusing that space), X, T/U, Z and the
// HRR weights, each sized for the largest of the four orientations.
template <int A, int B, int C, int D, bool SPH>
static void wg3_epi_max(int *nX, int *nT, int *nZ, int *nH)
{
    using E = WgEpi<A, B, C, D, SPH>;
    *nX = std::max(*nX, E::nE * E::nF);
    *nT = std::max(*nT, E::nT);
    *nZ = std::max(*nZ, E::nE * E::nZ);
    *nH = std::max(*nH, E::nH);
}

template <int LA, int LB, int LC, int LD, bool SPH>
static void wg3_epi_sizes(int *nX, int *nT, int *nZ, int *nH)
{
    wg3_epi_max<LA, LB, LC, LD, SPH>(nX, nT, nZ, nH);
    wg3_epi_max<LB, LA, LC, LD, SPH>(nX, nT, nZ, nH);
    wg3_epi_max<LA, LB, LD, LC, SPH>(nX, nT, nZ, nH);
    wg3_epi_max<LB, LA, LD, LC, SPH>(nX, nT, nZ, nH);
}

template <int LA, int LB, int LC, int LD>
static int wg2_layout(int sph, int vt_bytes, int max_lds, IntcWgCls *k)
{
    using W = Wg2<LA, LB, LC, LD>;
    constexpr int LMAX = imax(imax(LA, LB), imax(LC, LD));
    IntcWgCls c{};
    c.la = LA; c.lb = LB; c.lc = LC; c.ld = LD; c.sph = sph;
    int nX = 0, nT = 0, nZ = 0, nH = 0;
    if (sph)
        wg3_epi_sizes<LA, LB, LC, LD, true>(&nX, &nT, &nZ, &nH);
    else
        wg3_epi_sizes<LA, LB, LC, LD, false>(&nX, &nT, &nZ, &nH);
    auto al = [](int x) { return (x + 15) & ~15; };
    c.o_tab = 0;
    c.o_S = al(WG_TAB_INTS * 4);
    const int base = al(c.o_S + intc_sph_off(LMAX + 1) * 8);
    c.o_bra = base;
    c.o_lvl = al(c.o_bra + W::CH * W::NSLOT * 8);
    const int end1 = al(c.o_lvl + 3 * W::cap() * vt_bytes);
    c.o_X = base;
    c.o_T = al(c.o_X + nX * 8);
    c.o_Z = al(c.o_T + nT * 8);
    c.o_H = al(c.o_Z + nZ * 8);
    const int end2 = al(c.o_H + nH * 8);
    c.lds_bytes = std::max(end1, end2);
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
