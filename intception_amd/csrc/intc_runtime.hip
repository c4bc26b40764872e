// intc_runtime.hip -- C-ABI of the MI355X contracted-ERI library (libintc.so).
//
// Entry points (declared in include/intc.h):
//   * per-quartet twins of the reference's generated functions, same argument
//     layout: intc_eri_cont_{sph,cart}_main and ..._work_array_size
//   * batched path: a plan (canonical class bins, shell-pair list) built once
//     from host index arrays, then executed on device-resident primitive data:
//     pair builder kernel + one kernel per class bin.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "intc_common.h"
#include "boys_tables.inc"
#include "gen/class_table.inc"

#define INTC_API extern "C" __attribute__((visibility("default")))

namespace {

thread_local std::string g_err;
int g_boys_mode = 0;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                    \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess)                                                        \
            return fail(-3, std::string(#x) + ": " + hipGetErrorString(e_));         \
    } while (0)

struct DevTables {
    double *cheb = nullptr, *small = nullptr, *medium = nullptr;
    double *sph = nullptr;  // built-in S_l, l = 0..7
};
std::mutex g_mu;
std::vector<DevTables> g_tables;  // per device

}  // namespace

// workgroup kernels (intc_eri_wg.hip): allow their dynamic LDS on this device
int intc_wg_prepare();
// persistent multi-class kernel (gen/intc_eri_multi.hip)
extern "C" int intc_multi_prepare();
extern "C" hipError_t intc_multi_launch(const IntcEriArgs *args, const int4 *tiles, int n_tiles, int *next, int reg,
                                        int grid, hipStream_t st);

namespace {

int ensure_tables(int dev, DevTables **out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if ((int)g_tables.size() <= dev) g_tables.resize(dev + 1);
    DevTables &t = g_tables[dev];
    if (!t.cheb) {
        HIPCHK(hipMalloc(&t.cheb, sizeof(intc_boys_cheb_host)));
        HIPCHK(hipMalloc(&t.small, sizeof(intc_boys_smallx_host)));
        HIPCHK(hipMalloc(&t.medium, sizeof(intc_boys_mediumx_host)));
        HIPCHK(hipMemcpy(t.cheb, intc_boys_cheb_host, sizeof(intc_boys_cheb_host), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(t.small, intc_boys_smallx_host, sizeof(intc_boys_smallx_host), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(t.medium, intc_boys_mediumx_host, sizeof(intc_boys_mediumx_host), hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&t.sph, sizeof(intc_builtin_sph)));
        HIPCHK(hipMemcpy(t.sph, intc_builtin_sph, sizeof(intc_builtin_sph), hipMemcpyHostToDevice));
        if (intc_wg_prepare()) return fail(-3, "workgroup kernels: cannot enable their LDS size");
        if (intc_multi_prepare()) return fail(-3, "persistent kernel: cannot enable its LDS size");
    }
    *out = &t;
    return 0;
}

inline int ncc(int l) { return (l + 1) * (l + 2) / 2; }
inline int ncc0(int l) { return l < 0 ? 0 : (l + 1) * (l + 2) * (l + 3) / 6; }
inline int nsph(int l) { return 2 * l + 1; }

const IntcClassEntry *find_class(int la, int lb, int lc, int ld, int sph, int prec)
{
    for (int i = 0; i < intc_n_class; ++i) {
        const IntcClassEntry &c = intc_class_table[i];
        if (c.la == la && c.lb == lb && c.lc == lc && c.ld == ld && c.sph == sph && c.prec == prec) return &c;
    }
    return nullptr;
}

// canonical order: la >= lb, lc >= ld, smaller pair (by (L, la)) in the bra.
// perm[k] = original position of canonical position k.
void canonical(const int l[4], int cls[4], int perm[4])
{
    int b0 = 0, b1 = 1, k0 = 2, k1 = 3;
    if (l[b0] < l[b1]) std::swap(b0, b1);
    if (l[k0] < l[k1]) std::swap(k0, k1);
    int kb0 = l[b0] + l[b1], kb1 = l[b0], kk0 = l[k0] + l[k1], kk1 = l[k0];
    if (kb0 > kk0 || (kb0 == kk0 && kb1 > kk1)) {
        std::swap(b0, k0);
        std::swap(b1, k1);
    }
    perm[0] = b0; perm[1] = b1; perm[2] = k0; perm[3] = k1;
    for (int k = 0; k < 4; ++k) cls[k] = l[perm[k]];
}

// ---------------------------------------------------------------------------
// pair builder: one thread per primitive pair
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) intc_build_pairs(
    long long n_pp, const int *__restrict__ pp_pair, const int *__restrict__ pair_sa,
    const int *__restrict__ pair_sb, const int *__restrict__ pair_off, const int *__restrict__ pair_nk,
    const int *__restrict__ exp_off, const int *__restrict__ coef_off, const int *__restrict__ cen_idx,
    const double *__restrict__ shell_cen, const double *__restrict__ exps, const double *__restrict__ coefs,
    double *__restrict__ pz, double *__restrict__ po2z, double *__restrict__ ppx, double *__restrict__ ppy,
    double *__restrict__ ppz, double *__restrict__ pkap, double *__restrict__ pairA, double *__restrict__ pairAB)
{
    // plan shell s: exponents at exps[exp_off[s]], coefficients (one
    // contraction) at coefs[coef_off[s]], centre shell_cen[3 cen_idx[s]]; a
    // plan over general contractions has one plan shell per (shell,
    // contraction), all of them reading the caller's arrays
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_pp) return;
    const int pr = pp_pair[t];
    const int sa = pair_sa[pr], sb = pair_sb[pr];
    const int local = (int)(t - pair_off[pr]);
    const int kb = pair_nk[pr];
    const int ia = local / kb, ib = local - ia * kb;
    const double al = exps[exp_off[sa] + ia], be = exps[exp_off[sb] + ib];
    const double ca = coefs[coef_off[sa] + ia], cb = coefs[coef_off[sb] + ib];
    const int ca3 = 3 * cen_idx[sa], cb3 = 3 * cen_idx[sb];
    const double Ax = shell_cen[ca3], Ay = shell_cen[ca3 + 1], Az = shell_cen[ca3 + 2];
    const double Bx = shell_cen[cb3], By = shell_cen[cb3 + 1], Bz = shell_cen[cb3 + 2];
    const double ABx = Ax - Bx, ABy = Ay - By, ABz = Az - Bz;
    const double z = al + be, oz = 1.0 / z;
    const double r2 = ABx * ABx + ABy * ABy + ABz * ABz;
    pz[t] = z;
    po2z[t] = 0.5 * oz;
    ppx[t] = (al * Ax + be * Bx) * oz;
    ppy[t] = (al * Ay + be * By) * oz;
    ppz[t] = (al * Az + be * Bz) * oz;
    // sqrt(2 pi^(5/2)) K_AB c_a c_b / zeta: the product of a bra and a ket
    // kappa is the base prefactor 2 pi^(5/2) K_AB K_CD c_a c_b c_c c_d / (zeta eta)
    pkap[t] = 5.914967172795612878 * exp(-al * be * oz * r2) * ca * cb * oz;
    if (local == 0) {
        pairA[3 * pr] = Ax; pairA[3 * pr + 1] = Ay; pairA[3 * pr + 2] = Az;
        pairAB[3 * pr] = ABx; pairAB[3 * pr + 1] = ABy; pairAB[3 * pr + 2] = ABz;
    }
}

// Primitive-pair screen: one thread per shell pair keeps, in order, the
// primitive pairs whose s-type Schwarz factor w = |kappa| (2 zeta)^(-1/4)
// reaches `cut` and moves them to the front of the pair's range (in place: the
// write position never passes the read position); the kept count goes to
// np_out.  |(ij|kl)| <= w_ij w_kl for s shells (1/sqrt(zeta + eta) <=
// (4 zeta eta)^(-1/4)), so dropping primitive pairs with w < cut changes an
// integral by at most about cut * sum(w) of the other pair times an angular
// factor.
__global__ void __launch_bounds__(256) intc_screen_pairs(int n_pair, const int *__restrict__ pair_off,
                                                          const int *__restrict__ pair_np, double cut,
                                                          double *__restrict__ pz, double *__restrict__ po2z,
                                                          double *__restrict__ ppx, double *__restrict__ ppy,
                                                          double *__restrict__ ppz, double *__restrict__ pkap,
                                                          int *__restrict__ np_out)
{
    const int pr = blockIdx.x * blockDim.x + threadIdx.x;
    if (pr >= n_pair) return;
    const int o = pair_off[pr], n = pair_np[pr];
    int kept = 0;
    for (int k = 0; k < n; ++k) {
        const double z = pz[o + k], kap = pkap[o + k];
        if (!(fabs(kap) >= cut * sqrt(sqrt(2.0 * z)))) continue;
        if (kept != k) {
            pz[o + kept] = z;
            po2z[o + kept] = po2z[o + k];
            ppx[o + kept] = ppx[o + k];
            ppy[o + kept] = ppy[o + k];
            ppz[o + kept] = ppz[o + k];
            pkap[o + kept] = kap;
        }
        ++kept;
    }
    np_out[pr] = kept;
}

template <class T>
int dev_upload(T **d, const std::vector<T> &h)
{
    *d = nullptr;
    if (h.empty()) return 0;
    HIPCHK(hipMalloc(d, sizeof(T) * h.size()));
    HIPCHK(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return 0;
}

}  // namespace

struct IntcBin {
    const IntcClassEntry *cls[2];   // [0] FP64 kernel, [1] FP32-recurrence kernel
    int q0, nq;
    int team_groups[2];
    int grid[2];
    int block;
};

struct intc_plan {
    int spherical = 0;
    int device = 0;
    long long n_quartet = 0;
    long long n_pp = 0;
    int n_pair = 0;
    std::vector<IntcBin> bins;
    // device index data, in plan (sorted) order
    int *q_bra = nullptr, *q_ket = nullptr;
    long long *q_off = nullptr;
    unsigned long long *q_str = nullptr;
    int *q_orig = nullptr;        // plan position -> caller's quartet index
    int *q_bin = nullptr;         // plan position -> bin
    int *q_size = nullptr;        // integrals in the quartet's block
    int *pair_sa = nullptr, *pair_sb = nullptr, *pair_off = nullptr, *pair_np = nullptr, *pair_nk = nullptr;
    int *pp_pair = nullptr;
    int *shell_off = nullptr;     // per plan shell: exponent offset
    int *coef_off = nullptr;      // per plan shell: coefficient offset
    int *cen_idx = nullptr;       // per plan shell: centre index
    bool strided = false;         // caller strides or contractions: blocks are not dense
    // routed queues: per bin, FP64, FP32 and skip index lists (segments at q0)
    // and their lengths (3 per bin)
    int *idx64 = nullptr, *idx32 = nullptr, *idxs = nullptr, *counts = nullptr;
    int *bin_q0 = nullptr;
    // route-compaction tiles (consecutive plan positions inside one bin)
    int n_tile = 0;
    int *tile_q0 = nullptr, *tile_n = nullptr, *tile_bin = nullptr, *bin_tile0 = nullptr;
    int *tile_cnt = nullptr, *tile_base = nullptr;
    bool identity_order = false;  // caller's quartet order is the plan order
    // device pair data (filled by execute)
    double *pz = nullptr, *po2z = nullptr, *ppx = nullptr, *ppy = nullptr, *ppz = nullptr, *pkap = nullptr;
    double *pairA = nullptr, *pairAB = nullptr;
    bool pairs_built = false;   // primitive-pair arrays written at least once
    // primitive-pair screen (intc_plan_set_prim_screen): cut on the s-type
    // Schwarz factor of each primitive pair, 0 = keep every pair; the kept
    // count of each shell pair after the last build
    double prim_cut = 0.0;
    int *pair_np_scr = nullptr;
    bool route_fp32 = true;     // routed runs launch the FP32-recurrence queues
    // FP32 error check: the FP32 register kernels append quartets whose error
    // estimate exceeds fp32_tol to redo_q (bin-local positions, at bin_q0;
    // redo_n per bin) and each bin's FP64 kernel then recomputes them
    double fp32_tol = 0.0;
    int *redo_q = nullptr, *redo_n = nullptr;
    // persistent multi-class launch (intc_plan_set_persistent): the bins whose
    // class the multi kernel holds run as tiles of one launch; one buffer set
    // for eager execution and one bound by the captured graph
    bool persistent = false;
    int tile_iters = 64;                     // iterations of each team per tile (at most)
    int multi_grid = 2 * 256;                // blocks of 256 (two per CU)
    std::vector<long long> prim_prefix;      // primitive quartets before each plan position
    std::vector<char> multi_bin;             // bins the multi kernel may hold (empty: all)
    struct MultiBufs {
        IntcEriArgs *args = nullptr;         // [2 * bins]: bin b, precision r at 2b + r
        int4 *tiles = nullptr;
        int *next = nullptr;
        int n_tiles = 0, cap_tiles = 0, reg = 0;
        std::vector<IntcEriArgs> h_args;
        std::vector<int4> h_tiles;
    } m_eager, m_graph;
    // captured execution
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> events;
};

namespace {

void plan_free(intc_plan *p)
{
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (hipEvent_t e : p->events) (void)hipEventDestroy(e);
    for (hipStream_t st : p->side) (void)hipStreamDestroy(st);
    p->graph_exec = nullptr;
    p->graph = nullptr;
    p->events.clear();
    p->side.clear();
    void *ptrs[] = {p->q_bra, p->q_ket, p->q_off, p->q_str, p->q_orig, p->q_bin, p->q_size, p->pair_sa,
                    p->pair_sb, p->pair_off, p->pair_np, p->pair_nk, p->pp_pair, p->shell_off, p->coef_off, p->cen_idx, p->idx64,
                    p->idx32, p->idxs, p->counts, p->tile_q0, p->tile_n, p->tile_bin, p->bin_tile0,
                    p->tile_cnt, p->tile_base, p->bin_q0, p->pz, p->po2z, p->ppx, p->ppy, p->ppz, p->pkap,
                    p->pairA, p->pairAB, p->pair_np_scr, p->redo_q, p->redo_n};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    for (intc_plan::MultiBufs *m : {&p->m_eager, &p->m_graph}) {
        void *mp[] = {m->args, m->tiles, m->next};
        for (void *q : mp)
            if (q) (void)hipFree(q);
        *m = intc_plan::MultiBufs();
    }
}

// lanes of a wave doing useful work when teams of G*T lanes tile it
double lane_use(int G, int T)
{
    int tl = G * T;
    return tl > 64 ? 0.0 : (double)((64 / tl) * tl) / 64.0;
}

// Team size: among the T that keep the wave's lane use within 10% of the
// best, the smallest that still gives >= 8 waves per SIMD over the bin
// (teams of one group finish their epilogue with every lane busy; larger
// teams only pay off when a bin is too small to fill 1024 SIMDs).
int default_team_groups(const IntcClassEntry *c, int nq, double mean_prims)
{
    int G = c->G;
    double best = 0.0;
    for (int T = 1; G * T <= 64; ++T) best = std::max(best, lane_use(G, T));
    int pick = 1;
    for (int T = 1; G * T <= 64; ++T) {
        if (lane_use(G, T) < 0.9 * best) continue;
        if (T > 1 && mean_prims < 2.0 * T) break;
        double waves = (double)nq * G * T / 64.0;
        pick = T;
        if (waves >= 8.0 * 1024.0) break;
    }
    return pick;
}

int grid_for(const IntcBin &b, int prec)
{
    // a workgroup kernel takes one quartet per workgroup at a time
    if (b.cls[prec]->wg) return std::max(1, std::min(b.nq, 256 * 32));
    int tpw = 64 / (b.cls[prec]->G * b.team_groups[prec]);
    int teams_per_block = (b.block / 64) * tpw;
    return std::max(1, std::min((b.nq + teams_per_block - 1) / teams_per_block, 256 * 32));
}

// Route compaction, stable and without atomics: the plan's quartets are cut
// into tiles of RT_TILE consecutive plan positions that never straddle a bin.
//   intc_route_count    per tile: how many quartets take each route code
//   intc_route_scan     per bin: exclusive scan of the tile counts -> each
//                       tile's first queue slot per code, and the bin totals
//   intc_route_scatter  per tile: every quartet's position within its bin,
//                       appended to the bin's FP64, FP32 or skip queue
// The queues list positions in plan order (primitive count descending), the
// same order the unrouted launch walks.  Route codes: 0 FP64, 1 FP32 (read as
// 0 when use32 is off), 2 skip; anything else reads as 0.  When the caller's
// quartet order is the plan order (q_orig is the identity) route[] is read
// directly, otherwise through q_orig.
constexpr int RT_TILE = 2048, RT_PER_THREAD = RT_TILE / 256;

__device__ __forceinline__ int route_code(const signed char *__restrict__ route, const int *__restrict__ q_orig,
                                          int identity, int i, int use32)
{
    int r = route[identity ? i : q_orig[i]];
    if (r < 0 || r > 2) r = 0;
    if (r == 1 && !use32) r = 0;
    return r;
}

// block-wide exclusive scan of three counters per thread (256 threads);
// returns the block totals in tot[]
__device__ __forceinline__ void block_scan3(int v[3], int ex[3], int tot[3])
{
    __shared__ int wsum[4][3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc[3] = {v[0], v[1], v[2]};
    for (int off = 1; off < 64; off <<= 1)
        for (int c = 0; c < 3; ++c) {
            int u = __shfl_up(inc[c], off);
            if (lane >= off) inc[c] += u;
        }
    if (lane == 63)
        for (int c = 0; c < 3; ++c) wsum[wave][c] = inc[c];
    __syncthreads();
    for (int c = 0; c < 3; ++c) {
        int before = 0, all = 0;
        for (int w = 0; w < 4; ++w) {
            if (w < wave) before += wsum[w][c];
            all += wsum[w][c];
        }
        ex[c] = before + inc[c] - v[c];
        tot[c] = all;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(256) intc_route_count(const signed char *__restrict__ route,
                                                       const int *__restrict__ q_orig, int identity,
                                                       const int *__restrict__ tile_q0,
                                                       const int *__restrict__ tile_n, int use32,
                                                       int *__restrict__ tile_cnt)
{
    const int t = blockIdx.x, q0 = tile_q0[t], n = tile_n[t];
    int v[3] = {0, 0, 0};
    for (int k = 0; k < RT_PER_THREAD; ++k) {
        const int j = threadIdx.x * RT_PER_THREAD + k;
        if (j < n) ++v[route_code(route, q_orig, identity, q0 + j, use32)];
    }
    int ex[3], tot[3];
    block_scan3(v, ex, tot);
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) tile_cnt[3 * t + c] = tot[c];
}

__global__ void __launch_bounds__(256) intc_route_scan(const int *__restrict__ bin_tile0,
                                                      const int *__restrict__ tile_cnt,
                                                      int *__restrict__ tile_base, int *__restrict__ counts,
                                                      int *__restrict__ redo_n)
{
    const int b = blockIdx.x, t0 = bin_tile0[b], nt = bin_tile0[b + 1] - t0;
    int run[3] = {0, 0, 0};
    for (int base = 0; base < nt; base += 256) {
        const int i = base + threadIdx.x;
        int v[3], ex[3], tot[3];
        for (int c = 0; c < 3; ++c) v[c] = i < nt ? tile_cnt[3 * (t0 + i) + c] : 0;
        block_scan3(v, ex, tot);
        if (i < nt)
            for (int c = 0; c < 3; ++c) tile_base[3 * (t0 + i) + c] = run[c] + ex[c];
        for (int c = 0; c < 3; ++c) run[c] += tot[c];
    }
    if (threadIdx.x == 0) {
        for (int c = 0; c < 3; ++c) counts[3 * b + c] = run[c];
        if (redo_n) redo_n[b] = 0;   // the FP32 kernels of this run append from 0
    }
}

__global__ void __launch_bounds__(256) intc_route_scatter(const signed char *__restrict__ route,
                                                         const int *__restrict__ q_orig, int identity,
                                                         const int *__restrict__ tile_q0,
                                                         const int *__restrict__ tile_n,
                                                         const int *__restrict__ tile_bin,
                                                         const int *__restrict__ bin_q0,
                                                         const int *__restrict__ tile_base, int use32,
                                                         int *__restrict__ idx64, int *__restrict__ idx32,
                                                         int *__restrict__ idxs)
{
    const int t = blockIdx.x, q0 = tile_q0[t], n = tile_n[t];
    const int bq0 = bin_q0[tile_bin[t]];
    int r[RT_PER_THREAD];
    int v[3] = {0, 0, 0};
    for (int k = 0; k < RT_PER_THREAD; ++k) {
        const int j = threadIdx.x * RT_PER_THREAD + k;
        r[k] = j < n ? route_code(route, q_orig, identity, q0 + j, use32) : 3;
        if (r[k] < 3) ++v[r[k]];
    }
    int ex[3], tot[3];
    block_scan3(v, ex, tot);
    int pos[3];
    for (int c = 0; c < 3; ++c) pos[c] = bq0 + tile_base[3 * t + c] + ex[c];
    for (int k = 0; k < RT_PER_THREAD; ++k) {
        const int j = threadIdx.x * RT_PER_THREAD + k;
        const int c = r[k];
        if (c < 3) {
            // queue entries are positions relative to the bin's first quartet
            (c == 0 ? idx64 : c == 1 ? idx32 : idxs)[pos[c]++] = q0 + j - bq0;
        }
    }
}

// Zero the output blocks of skipped quartets.  blockIdx.y = bin (all blocks
// of a bin have the same size sz).  Each wave takes 64 skip entries at a time:
// lane j loads entry j's output offset once, then the wave walks the 64 * sz
// flattened elements, one double per lane, fetching the entry's offset by a
// lane shuffle.  Skipped quartets that are adjacent in the output give
// contiguous 512-byte stores.
__global__ void __launch_bounds__(256) intc_zero_skipped(const int *__restrict__ bin_q0,
                                                        const int *__restrict__ counts,
                                                        const int *__restrict__ idxs,
                                                        const long long *__restrict__ q_off,
                                                        const int *__restrict__ q_size, double *__restrict__ out)
{
    const int b = blockIdx.y;
    const int n = counts[3 * b + 2];
    if (n == 0) return;
    const int q0 = bin_q0[b];
    const int sz = q_size[q0];
    const float inv = 1.0f / (float)sz;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (long long c0 = ((long long)blockIdx.x * 4 + wave) * 64; c0 < n; c0 += (long long)gridDim.x * 4 * 64) {
        const int m = (int)min((long long)64, n - c0);
        const long long off = lane < m ? q_off[q0 + idxs[q0 + c0 + lane]] : 0;
        const int ne = m * sz;  // < 2^24: exact in float
        for (int base = 0; base < ne; base += 64) {
            const int e = base + lane;
            int j = (int)((float)e * inv);
            if (j * sz > e) --j;
            if ((j + 1) * sz <= e) ++j;
            j = min(j, m - 1);
            // every lane takes part in the shuffle; only in-range lanes store
            const long long oj = __shfl(off, j);
            if (e < ne) out[oj + (e - j * sz)] = 0.0;
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
INTC_API const char *intc_last_error(void) { return g_err.c_str(); }

INTC_API int intc_version(void) { return 2; }

INTC_API int intc_set_boys_mode(int mode)
{
    if (mode != 0 && mode != 1) return fail(-1, "boys mode must be 0 (Chebyshev) or 1 (reference scheme)");
    g_boys_mode = mode;
    return 0;
}

// shared with the one-electron plans (intc_onee.hip): error message and the
// device tables of the current device
int intc_internal_fail(int code, const char *msg) { return fail(code, msg); }

int intc_internal_tables(const double **cheb, const double **small, const double **medium, const double **sph)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    DevTables *t;
    int rc = ensure_tables(dev, &t);
    if (rc) return rc;
    *cheb = t->cheb;
    *small = t->small;
    *medium = t->medium;
    *sph = t->sph;
    return 0;
}

int intc_internal_boys_mode() { return g_boys_mode; }

// Scratch of the allocation-free per-quartet twin (intc_eri_cont_sph_main_noalloc):
// one static device array per GPU, its device address looked up once per
// device by intc_init.  Layout: NoAllocHead (inputs and index tables, copied
// from the host), the primitive-pair arrays, then the output blocks.
namespace {
constexpr int NA_PRIM = 8;                        // primitives per shell
constexpr int NA_CONT = 2;                        // contractions per shell
constexpr int NA_SHELL = 4 * NA_CONT;
constexpr int NA_COMBO = NA_CONT * NA_CONT * NA_CONT * NA_CONT;
constexpr int NA_PAIR = 2 * NA_COMBO;             // one bra and one ket pair per combination
constexpr int NA_PP = NA_PAIR * NA_PRIM * NA_PRIM;
constexpr size_t NA_SCRATCH_DOUBLES = (size_t)1 << 20;   // 8 MiB per device
constexpr int NA_MAX_DEV = 64;
struct NoAllocHead {
    double exps[NA_SHELL * NA_PRIM], coefs[NA_SHELL * NA_PRIM], cen[NA_SHELL * 3];
    double sph[intc_sph_off(8)];
    long long q_off[NA_COMBO];
    unsigned long long q_str[NA_COMBO];
    int shell_off[NA_SHELL], vid[NA_SHELL];
    int pair_sa[NA_PAIR], pair_sb[NA_PAIR], pair_off[NA_PAIR], pair_np[NA_PAIR], pair_nk[NA_PAIR];
    int pp_pair[NA_PP];
    int q_bra[NA_COMBO], q_ket[NA_COMBO];
};
constexpr size_t NA_HEAD_DOUBLES = (sizeof(NoAllocHead) + 255) / 256 * 256 / sizeof(double);
constexpr size_t NA_OUT_OFFSET = NA_HEAD_DOUBLES + 6 * (size_t)NA_PP + 6 * (size_t)NA_PAIR;
constexpr size_t NA_OUT_DOUBLES = NA_SCRATCH_DOUBLES - NA_OUT_OFFSET;
double *g_scratch_ptr[NA_MAX_DEV] = {};
}  // namespace
static __device__ __align__(256) double scratch[NA_SCRATCH_DOUBLES];

INTC_API int intc_init(void)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    DevTables *t;
    int rc = ensure_tables(dev, &t);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    if (dev < NA_MAX_DEV && !g_scratch_ptr[dev]) {
        // the host-side name `scratch` is a stand-in; kernels get this address
        double *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = ptr;
    }
    return 0;
}

INTC_API int intc_class_supported(int la, int lb, int lc, int ld, int spherical)
{
    int l[4] = {la, lb, lc, ld}, cls[4], perm[4];
    canonical(l, cls, perm);
    return find_class(cls[0], cls[1], cls[2], cls[3], spherical ? 1 : 0, 0) != nullptr;
}

// Shared tail of plan creation: quartets already in plan order with their
// pair ids, class entries, offsets, strides and sizes; pairs in kernel
// orientation.  Builds bins, primitive-pair layout and uploads everything.
static int finish_plan(intc_plan *p, const std::vector<const IntcClassEntry *> &qcls, const std::vector<int> &prims,
                       std::vector<int> &q_bra, std::vector<int> &q_ket, std::vector<long long> &q_off,
                       std::vector<unsigned long long> &q_str, std::vector<int> &q_size, std::vector<int> &q_orig,
                       const std::vector<int> &pair_sa, const std::vector<int> &pair_sb, const int *shell_nprim,
                       const int *shell_off, int n_shell, const int *coef_off = nullptr,
                       const int *cen_idx = nullptr)
{
    const long long nq = (long long)q_bra.size();
    const size_t npair = pair_sa.size();
    std::vector<int> pair_off(npair), pair_np(npair), pair_nk(npair);
    long long npp = 0;
    for (size_t pr = 0; pr < npair; ++pr) {
        pair_off[pr] = (int)npp;
        pair_np[pr] = shell_nprim[pair_sa[pr]] * shell_nprim[pair_sb[pr]];
        pair_nk[pr] = shell_nprim[pair_sb[pr]];
        npp += pair_np[pr];
        if (npp > 0x7fffffffLL) return fail(-1, "too many primitive pairs for one plan");
    }
    std::vector<int> pp_pair(npp);
    for (size_t pr = 0; pr < npair; ++pr)
        for (int k = 0; k < pair_np[pr]; ++k) pp_pair[pair_off[pr] + k] = (int)pr;
    p->prim_prefix.assign(nq + 1, 0);
    for (long long i = 0; i < nq; ++i) p->prim_prefix[i + 1] = p->prim_prefix[i] + prims[i];
    std::vector<int> bin_q0, q_bin(nq);
    for (long long i = 0; i < nq;) {
        long long j = i;
        const IntcClassEntry *c = qcls[i];
        double sum = 0.0;
        while (j < nq && qcls[j] == c) sum += prims[j++];
        IntcBin b;
        b.cls[0] = c;
        // FP32-recurrence kernel; register classes with an f or g shell (the
        // three-centre (ab|P s) classes) have none, and their FP32-routed
        // quartets run the FP64 kernel
        b.cls[1] = find_class(c->la, c->lb, c->lc, c->ld, p->spherical, 1);
        if (!b.cls[1]) b.cls[1] = c;
        b.q0 = (int)i;
        b.nq = (int)(j - i);
        b.block = 256;
        for (int pr = 0; pr < 2; ++pr) {
            b.team_groups[pr] = default_team_groups(b.cls[pr], b.nq, sum / b.nq);
            b.grid[pr] = grid_for(b, pr);
        }
        for (long long k = i; k < j; ++k) q_bin[k] = (int)p->bins.size();
        bin_q0.push_back((int)i);
        p->bins.push_back(b);
        i = j;
    }
    std::vector<int> tile_q0, tile_n, tile_bin, bin_tile0;
    for (size_t b = 0; b < p->bins.size(); ++b) {
        bin_tile0.push_back((int)tile_q0.size());
        for (int s0 = 0; s0 < p->bins[b].nq; s0 += RT_TILE) {
            tile_q0.push_back(p->bins[b].q0 + s0);
            tile_n.push_back(std::min(RT_TILE, p->bins[b].nq - s0));
            tile_bin.push_back((int)b);
        }
    }
    bin_tile0.push_back((int)tile_q0.size());
    p->n_tile = (int)tile_q0.size();
    p->identity_order = true;
    for (long long i = 0; i < nq && p->identity_order; ++i) p->identity_order = q_orig[i] == i;
    std::vector<int> soff(shell_off, shell_off + n_shell);
    std::vector<int> coff(coef_off ? coef_off : shell_off, (coef_off ? coef_off : shell_off) + n_shell);
    std::vector<int> cidx(n_shell);
    for (int s2 = 0; s2 < n_shell; ++s2) cidx[s2] = cen_idx ? cen_idx[s2] : s2;
    int rc = 0;
    rc |= dev_upload(&p->tile_q0, tile_q0);
    rc |= dev_upload(&p->tile_n, tile_n);
    rc |= dev_upload(&p->tile_bin, tile_bin);
    rc |= dev_upload(&p->bin_tile0, bin_tile0);
    rc |= dev_upload(&p->q_bra, q_bra);
    rc |= dev_upload(&p->q_ket, q_ket);
    rc |= dev_upload(&p->q_off, q_off);
    rc |= dev_upload(&p->q_str, q_str);
    rc |= dev_upload(&p->q_orig, q_orig);
    rc |= dev_upload(&p->q_bin, q_bin);
    rc |= dev_upload(&p->q_size, q_size);
    rc |= dev_upload(&p->bin_q0, bin_q0);
    rc |= dev_upload(&p->pair_sa, pair_sa);
    rc |= dev_upload(&p->pair_sb, pair_sb);
    rc |= dev_upload(&p->pair_off, pair_off);
    rc |= dev_upload(&p->pair_np, pair_np);
    rc |= dev_upload(&p->pair_nk, pair_nk);
    rc |= dev_upload(&p->pp_pair, pp_pair);
    rc |= dev_upload(&p->shell_off, soff);
    rc |= dev_upload(&p->coef_off, coff);
    rc |= dev_upload(&p->cen_idx, cidx);
    if (rc) return rc;
    p->n_pair = (int)npair;
    p->n_pp = npp;
    p->n_quartet = nq;
    if (nq > 0) {
        if (hipMalloc(&p->idx64, sizeof(int) * nq) != hipSuccess ||
            hipMalloc(&p->idx32, sizeof(int) * nq) != hipSuccess || hipMalloc(&p->idxs, sizeof(int) * nq) != hipSuccess ||
            hipMalloc(&p->counts, sizeof(int) * 3 * std::max<size_t>(1, p->bins.size())) != hipSuccess ||
            hipMalloc(&p->tile_cnt, sizeof(int) * 3 * std::max(1, p->n_tile)) != hipSuccess ||
            hipMalloc(&p->tile_base, sizeof(int) * 3 * std::max(1, p->n_tile)) != hipSuccess)
            return fail(-3, "hipMalloc route queues");
    }
    if (npp > 0) {
        double **arrs[] = {&p->pz, &p->po2z, &p->ppx, &p->ppy, &p->ppz, &p->pkap};
        for (double **a : arrs)
            if (hipMalloc(a, sizeof(double) * npp) != hipSuccess) return fail(-3, "hipMalloc pair data");
        if (hipMalloc(&p->pairA, sizeof(double) * 3 * npair) != hipSuccess ||
            hipMalloc(&p->pairAB, sizeof(double) * 3 * npair) != hipSuccess)
            return fail(-3, "hipMalloc pair geometry");
    }
    return 0;
}

static int check_shells(int n_shell, const int *shell_l, const int *shell_nprim)
{
    if (n_shell < 0) return fail(-1, "negative size");
    for (int s = 0; s < n_shell; ++s) {
        if (shell_l[s] < 0 || shell_l[s] > 7) return fail(-1, "shell angular momentum out of range");
        if (shell_nprim[s] < 1 || shell_nprim[s] > 64) return fail(-1, "shell primitive count out of range");
    }
    return 0;
}

static int class_error(int la, int lb, int lc, int ld, int spherical)
{
    char buf[128];
    snprintf(buf, sizeof buf, "class (%d %d|%d %d) %s not built into this library", la, lb, lc, ld,
             spherical ? "spherical" : "cartesian");
    return fail(-2, buf);
}

// plan order: class (table order), then primitive-quartet count descending,
// then caller index -- one 64-bit key per quartet
static std::vector<int> plan_order(const std::vector<const IntcClassEntry *> &cls, const std::vector<int> &prims)
{
    const size_t nq = cls.size();
    std::vector<unsigned long long> key(nq);
    for (size_t q = 0; q < nq; ++q) {
        unsigned long long c = (unsigned long long)(cls[q] - intc_class_table);
        unsigned long long pr = 0xfffffull - (unsigned long long)std::min(prims[q], 0xfffff);
        key[q] = (c << 52) | (pr << 32) | (unsigned long long)q;
    }
    std::sort(key.begin(), key.end());
    std::vector<int> order(nq);
    for (size_t i = 0; i < nq; ++i) order[i] = (int)(key[i] & 0xffffffffull);
    return order;
}

// Orientation of a pair of two shells with the same l (the class, and so the
// kernel, is the same either way round).  The kernels build angular momentum
// on the pair's first centre and move it to the second with the HRR over
// A - B; when the primitive-pair centres P lie far from the first centre (a
// diffuse shell next to a tight one on another atom) that transfer cancels
// many digits.  True when the second shell should come first: the primitive
// pairs, weighted by |c_a c_b| exp(-mu R^2) / (alpha + beta), lie closer to
// it (|P - A| = beta/(alpha+beta) R, |P - B| = alpha/(alpha+beta) R).
static bool vrr_on_second(int a, int b, const int *shell_nprim, const int *shell_off, const double *shell_cen,
                          const double *exps, const double *coefs)
{
    if (!shell_cen || !exps || !coefs || a == b) return false;
    double r2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double x = shell_cen[3 * a + d] - shell_cen[3 * b + d];
        r2 += x * x;
    }
    if (r2 == 0.0) return false;
    double wa = 0.0, wb = 0.0;
    for (int i = 0; i < shell_nprim[a]; ++i)
        for (int j = 0; j < shell_nprim[b]; ++j) {
            const double al = exps[shell_off[a] + i], be = exps[shell_off[b] + j], oz = 1.0 / (al + be);
            const double w = std::fabs(coefs[shell_off[a] + i] * coefs[shell_off[b] + j]) *
                             std::exp(-al * be * oz * r2) * oz;
            wa += w * be * be * oz * oz;
            wb += w * al * al * oz * oz;
        }
    return wb < wa;
}

// Plan over plan shells (shell_l, shell_nprim; host exps/coefs at shell_off
// for the orientation choice).  strides: 4 per quartet (caller order a, b,
// c, d) or NULL for dense row-major blocks.  dev_exp_off / dev_coef_off /
// dev_cen: where the pair builder finds each plan shell's exponents,
// coefficients and centre in the device arrays execute receives (NULL: at
// shell_off and the shell's own index).
static int plan_create_impl(intc_plan **out, int spherical, int n_shell, const int *shell_l, const int *shell_nprim,
                            const int *shell_off, const double *shell_cen, const double *exps, const double *coefs,
                            long long n_quartet, const int *quartets, const long long *out_off, const int *strides,
                            const int *dev_exp_off, const int *dev_coef_off, const int *dev_cen)
{
    *out = nullptr;
    if (n_quartet < 0) return fail(-1, "negative size");
    if (n_quartet > 0x3fffffffLL) return fail(-1, "too many quartets for one plan");
    int rc = check_shells(n_shell, shell_l, shell_nprim);
    if (rc) return rc;
    const int nq = (int)n_quartet;
    std::vector<const IntcClassEntry *> cls(nq);
    std::vector<int> prims(nq);
    std::vector<std::array<int, 4>> perms(nq);
    for (int q = 0; q < nq; ++q) {
        int l[4], c[4], perm[4];
        for (int k = 0; k < 4; ++k) {
            int s = quartets[4 * (size_t)q + k];
            if (s < 0 || s >= n_shell) return fail(-1, "quartet shell id out of range");
            l[k] = shell_l[s];
        }
        canonical(l, c, perm);
        cls[q] = find_class(c[0], c[1], c[2], c[3], spherical ? 1 : 0, 0);
        if (!cls[q]) return class_error(l[0], l[1], l[2], l[3], spherical);
        const int *qs = quartets + 4 * (size_t)q;
        for (int h = 0; h < 4; h += 2)
            if (c[h] == c[h + 1] &&
                vrr_on_second(qs[perm[h]], qs[perm[h + 1]], shell_nprim, shell_off, shell_cen, exps, coefs))
                std::swap(perm[h], perm[h + 1]);
        for (int k = 0; k < 4; ++k) perms[q][k] = perm[k];
        int pr = 1;
        for (int k = 0; k < 4; ++k) pr *= shell_nprim[quartets[4 * (size_t)q + k]];
        prims[q] = pr;
    }
    std::vector<int> order = plan_order(cls, prims);
    intc_plan *p = new intc_plan();
    p->spherical = spherical ? 1 : 0;
    if (hipGetDevice(&p->device) != hipSuccess) { delete p; return fail(-3, "hipGetDevice"); }
    // one bra and one ket pair per quartet, in canonical shell order
    std::vector<int> q_bra(nq), q_ket(nq), q_size(nq), pair_sa(2 * (size_t)nq), pair_sb(2 * (size_t)nq);
    std::vector<long long> q_off(nq);
    std::vector<unsigned long long> q_str(nq);
    std::vector<const IntcClassEntry *> qcls(nq);
    std::vector<int> qprims(nq);
    for (int i = 0; i < nq; ++i) {
        int q = order[i];
        const int *s = quartets + 4 * (size_t)q;
        const int *pm = perms[q].data();
        int sh[4] = {s[pm[0]], s[pm[1]], s[pm[2]], s[pm[3]]};
        pair_sa[2 * i] = sh[0];
        pair_sb[2 * i] = sh[1];
        pair_sa[2 * i + 1] = sh[2];
        pair_sb[2 * i + 1] = sh[3];
        q_bra[i] = 2 * i;
        q_ket[i] = 2 * i + 1;
        q_off[i] = out_off[q];
        int n[4];
        for (int k = 0; k < 4; ++k) {
            int l = shell_l[s[k]];
            n[k] = p->spherical ? nsph(l) : ncc(l);
        }
        q_size[i] = n[0] * n[1] * n[2] * n[3];
        int st[4] = {n[1] * n[2] * n[3], n[2] * n[3], n[3], 1};
        if (strides)
            for (int k = 0; k < 4; ++k) st[k] = strides[4 * (size_t)q + k];
        unsigned long long pk = 0;
        for (int k = 0; k < 4; ++k) pk |= (unsigned long long)(st[pm[k]] & 0xffff) << (16 * k);
        q_str[i] = pk;
        qcls[i] = cls[q];
        qprims[i] = prims[q];
    }
    p->strided = strides != nullptr;
    rc = finish_plan(p, qcls, qprims, q_bra, q_ket, q_off, q_str, q_size, order, pair_sa, pair_sb, shell_nprim,
                     dev_exp_off ? dev_exp_off : shell_off, n_shell, dev_coef_off, dev_cen);
    if (rc) { plan_free(p); delete p; return rc; }
    *out = p;
    return 0;
}

INTC_API int intc_plan_create(intc_plan **out, int spherical, int n_shell, const int *shell_l,
                              const int *shell_nprim, const int *shell_off, const double *shell_cen,
                              const double *exps, const double *coefs, long long n_quartet, const int *quartets,
                              const long long *out_off)
{
    return plan_create_impl(out, spherical, n_shell, shell_l, shell_nprim, shell_off, shell_cen, exps, coefs,
                            n_quartet, quartets, out_off, nullptr, nullptr, nullptr, nullptr);
}

// General contraction and caller strides: shell s has shell_ncont[s]
// contractions, column k of its coefficients at coefs[coef_off[s] + k nprim];
// quartet q writes element (ia, ib, ic, id) of contraction (ka, kb, kc, kd) at
// out[out_off[q] + sum_X (iX iskip[4q+X] + kX iskip_cont[4q+X])], the
// reference's per-quartet addressing (contracted_general.py:1808-1875).  Each
// (shell, contraction) becomes one plan shell and each quartet one plan
// quartet per combination of contractions.
INTC_API int intc_plan_create_general(intc_plan **out, int spherical, int n_shell, const int *shell_l,
                                      const int *shell_nprim, const int *shell_ncont, const int *shell_off,
                                      const int *coef_off, const double *shell_cen, const double *exps,
                                      const double *coefs, long long n_quartet, const int *quartets,
                                      const long long *out_off, const int *iskip, const int *iskip_cont)
{
    *out = nullptr;
    if (n_quartet < 0 || n_shell < 0) return fail(-1, "negative size");
    if (!shell_ncont || !coef_off || !iskip || !iskip_cont) return fail(-1, "shell_ncont, coef_off, iskip and "
                                                                           "iskip_cont are required");
    std::vector<int> base(n_shell), v_l, v_np, v_eoff, v_coff, v_cen, v_hoff;
    std::vector<double> h_ex, h_co, h_cen;
    for (int s2 = 0; s2 < n_shell; ++s2) {
        if (shell_ncont[s2] < 1) return fail(-1, "shell_ncont must be >= 1");
        if (shell_nprim[s2] < 1) return fail(-1, "shell_nprim must be >= 1");
        base[s2] = (int)v_l.size();
        for (int k = 0; k < shell_ncont[s2]; ++k) {
            v_l.push_back(shell_l[s2]);
            v_np.push_back(shell_nprim[s2]);
            v_eoff.push_back(shell_off[s2]);
            v_coff.push_back(coef_off[s2] + k * shell_nprim[s2]);
            v_cen.push_back(s2);
            v_hoff.push_back((int)h_ex.size());
            for (int i = 0; i < shell_nprim[s2]; ++i) {
                h_ex.push_back(exps ? exps[shell_off[s2] + i] : 0.0);
                h_co.push_back(coefs ? coefs[coef_off[s2] + k * shell_nprim[s2] + i] : 0.0);
            }
            for (int d = 0; d < 3; ++d) h_cen.push_back(shell_cen ? shell_cen[3 * s2 + d] : 0.0);
        }
    }
    std::vector<int> vq;
    std::vector<long long> voff;
    std::vector<int> vst;
    for (long long q = 0; q < n_quartet; ++q) {
        const int *sq = quartets + 4 * q;
        int nc[4];
        for (int k = 0; k < 4; ++k) {
            if (sq[k] < 0 || sq[k] >= n_shell) return fail(-1, "quartet shell id out of range");
            nc[k] = shell_ncont[sq[k]];
            const int st = iskip[4 * q + k];
            if (st < 0 || st > 0xffff) return fail(-1, "iskip must be in [0, 65535] (16-bit strides)");
            if (iskip_cont[4 * q + k] < 0) return fail(-1, "iskip_cont must be >= 0");
        }
        for (int ka = 0; ka < nc[0]; ++ka)
            for (int kb = 0; kb < nc[1]; ++kb)
                for (int kc = 0; kc < nc[2]; ++kc)
                    for (int kd = 0; kd < nc[3]; ++kd) {
                        const int kk[4] = {ka, kb, kc, kd};
                        long long o = out_off[q];
                        for (int k = 0; k < 4; ++k) {
                            vq.push_back(base[sq[k]] + kk[k]);
                            vst.push_back(iskip[4 * q + k]);
                            o += (long long)kk[k] * iskip_cont[4 * q + k];
                        }
                        voff.push_back(o);
                    }
    }
    const bool have_host = shell_cen && exps && coefs;
    int rc = plan_create_impl(out, spherical, (int)v_l.size(), v_l.data(), v_np.data(), v_hoff.data(),
                              have_host ? h_cen.data() : nullptr, have_host ? h_ex.data() : nullptr,
                              have_host ? h_co.data() : nullptr, (long long)voff.size(), vq.data(), voff.data(),
                              vst.data(), v_eoff.data(), v_coff.data(), v_cen.data());
    return rc;
}

// Quartets given as pairs of shell pairs: pair P = (pair_shells[2P], pair_shells[2P+1]),
// quartet q = (pair quartet_pairs[2q] | pair quartet_pairs[2q+1]), block
// row-major over the four shells in that order.  Pair data are built once per
// pair and shared by every quartet that uses it.
INTC_API int intc_plan_create_pairs(intc_plan **out, int spherical, int n_shell, const int *shell_l,
                                    const int *shell_nprim, const int *shell_off, const double *shell_cen,
                                    const double *exps, const double *coefs, int n_pair, const int *pair_shells,
                                    long long n_quartet, const int *quartet_pairs, const long long *out_off)
{
    *out = nullptr;
    if (n_quartet < 0 || n_pair < 0) return fail(-1, "negative size");
    if (n_quartet > 0x3fffffffLL) return fail(-1, "too many quartets for one plan");
    int rc = check_shells(n_shell, shell_l, shell_nprim);
    if (rc) return rc;
    // pairs in kernel orientation: l(first) >= l(second)
    std::vector<int> pair_sa(n_pair), pair_sb(n_pair);
    std::vector<char> swapped(n_pair);
    for (int k = 0; k < n_pair; ++k) {
        int a = pair_shells[2 * (size_t)k], b = pair_shells[2 * (size_t)k + 1];
        if (a < 0 || a >= n_shell || b < 0 || b >= n_shell) return fail(-1, "pair shell id out of range");
        swapped[k] = shell_l[a] < shell_l[b] ||
                     (shell_l[a] == shell_l[b] &&
                      vrr_on_second(a, b, shell_nprim, shell_off, shell_cen, exps, coefs));
        pair_sa[k] = swapped[k] ? b : a;
        pair_sb[k] = swapped[k] ? a : b;
    }
    const long long nq = n_quartet;
    std::vector<const IntcClassEntry *> cls(nq);
    std::vector<int> prims(nq);
    std::vector<char> braket(nq);
    for (long long q = 0; q < nq; ++q) {
        int P = quartet_pairs[2 * q], R = quartet_pairs[2 * q + 1];
        if (P < 0 || P >= n_pair || R < 0 || R >= n_pair) return fail(-1, "quartet pair id out of range");
        int la = shell_l[pair_sa[P]], lb = shell_l[pair_sb[P]], lc = shell_l[pair_sa[R]], ld = shell_l[pair_sb[R]];
        bool sw = (la + lb > lc + ld) || (la + lb == lc + ld && la > lc);
        braket[q] = sw;
        int c[4] = {sw ? lc : la, sw ? ld : lb, sw ? la : lc, sw ? lb : ld};
        cls[q] = find_class(c[0], c[1], c[2], c[3], spherical ? 1 : 0, 0);
        if (!cls[q]) return class_error(la, lb, lc, ld, spherical);
        prims[q] = shell_nprim[pair_sa[P]] * shell_nprim[pair_sb[P]] * shell_nprim[pair_sa[R]] *
                   shell_nprim[pair_sb[R]];
    }
    std::vector<int> order = plan_order(cls, prims);
    intc_plan *p = new intc_plan();
    p->spherical = spherical ? 1 : 0;
    if (hipGetDevice(&p->device) != hipSuccess) { delete p; return fail(-3, "hipGetDevice"); }
    std::vector<int> q_bra(nq), q_ket(nq), q_size(nq), qprims(nq);
    std::vector<long long> q_off(nq);
    std::vector<unsigned long long> q_str(nq);
    std::vector<const IntcClassEntry *> qcls(nq);
    for (long long i = 0; i < nq; ++i) {
        int q = order[i];
        int P = quartet_pairs[2 * (size_t)q], R = quartet_pairs[2 * (size_t)q + 1];
        // caller layout [a][b][c][d] with (a,b) = pair P as given, (c,d) = pair R
        int sh[4] = {pair_shells[2 * (size_t)P], pair_shells[2 * (size_t)P + 1], pair_shells[2 * (size_t)R],
                     pair_shells[2 * (size_t)R + 1]};
        int n[4];
        for (int k = 0; k < 4; ++k) n[k] = p->spherical ? nsph(shell_l[sh[k]]) : ncc(shell_l[sh[k]]);
        int st[4] = {n[1] * n[2] * n[3], n[2] * n[3], n[3], 1};
        // canonical position -> caller position
        int posP[2] = {swapped[P] ? 1 : 0, swapped[P] ? 0 : 1};
        int posR[2] = {swapped[R] ? 3 : 2, swapped[R] ? 2 : 3};
        int perm[4];
        if (braket[q]) {
            perm[0] = posR[0]; perm[1] = posR[1]; perm[2] = posP[0]; perm[3] = posP[1];
            q_bra[i] = R;
            q_ket[i] = P;
        } else {
            perm[0] = posP[0]; perm[1] = posP[1]; perm[2] = posR[0]; perm[3] = posR[1];
            q_bra[i] = P;
            q_ket[i] = R;
        }
        unsigned long long pk = 0;
        for (int k = 0; k < 4; ++k) pk |= (unsigned long long)(st[perm[k]] & 0xffff) << (16 * k);
        q_str[i] = pk;
        q_off[i] = out_off[q];
        q_size[i] = n[0] * n[1] * n[2] * n[3];
        qcls[i] = cls[q];
        qprims[i] = prims[q];
    }
    rc = finish_plan(p, qcls, qprims, q_bra, q_ket, q_off, q_str, q_size, order, pair_sa, pair_sb, shell_nprim,
                     shell_off, n_shell);
    if (rc) { plan_free(p); delete p; return rc; }
    *out = p;
    return 0;
}

INTC_API int intc_plan_num_bins(const intc_plan *p) { return (int)p->bins.size(); }

INTC_API int intc_plan_bin_info(const intc_plan *p, int i, int prec, int *cls4, int *nq, int *team_groups,
                                int *grid, int *G)
{
    if (i < 0 || i >= (int)p->bins.size()) return fail(-1, "bin index out of range");
    if (prec != 0 && prec != 1) return fail(-1, "prec must be 0 (FP64) or 1 (FP32 recurrences)");
    const IntcBin &b = p->bins[i];
    cls4[0] = b.cls[0]->la; cls4[1] = b.cls[0]->lb; cls4[2] = b.cls[0]->lc; cls4[3] = b.cls[0]->ld;
    *nq = b.nq;
    *team_groups = b.team_groups[prec];
    *grid = b.grid[prec];
    *G = b.cls[prec]->G;
    return 0;
}

INTC_API int intc_plan_set_bin_tuning(intc_plan *p, int i, int prec, int team_groups, int grid)
{
    if (i < 0 || i >= (int)p->bins.size()) return fail(-1, "bin index out of range");
    if (prec != 0 && prec != 1) return fail(-1, "prec must be 0 (FP64) or 1 (FP32 recurrences)");
    IntcBin &b = p->bins[i];
    if (team_groups < 1 || team_groups * b.cls[prec]->G > 64) return fail(-1, "team_groups * G must be in [1, 64]");
    if (grid < 1) return fail(-1, "grid must be positive");
    b.team_groups[prec] = team_groups;
    b.grid[prec] = grid;
    return 0;
}

// launch one bin in one precision; q_idx/nq_dev select a routed queue
static int launch_bin(const intc_plan *p, const IntcBin &b, int prec, const IntcEriArgs &base, const int *q_idx,
                      const int *nq_dev, hipStream_t st)
{
    IntcEriArgs a = base;
    a.q_bra = p->q_bra + b.q0;
    a.q_ket = p->q_ket + b.q0;
    a.q_off = p->q_off + b.q0;
    a.q_str = p->q_str + b.q0;
    a.q_idx = q_idx;
    a.nq_dev = nq_dev;
    a.nq = b.nq;
    a.team_groups = b.team_groups[prec];
    const IntcClassEntry *c = b.cls[prec];
    a.wg_cls[0] = c->la;
    a.wg_cls[1] = c->lb;
    a.wg_cls[2] = c->lc;
    a.wg_cls[3] = c->ld;
    a.wg_cls[4] = c->sph;
    int tpw = 64 / (c->G * a.team_groups);
    // workgroup kernels size their own LDS
    size_t lds = c->wg ? 0 : (size_t)(b.block / 64) * tpw * c->NE * c->NCD * sizeof(double);
    hipError_t e = c->fn(&a, b.grid[prec], b.block, lds, st);
    if (e != hipSuccess) return fail(-3, std::string("class kernel launch: ") + hipGetErrorString(e));
    return 0;
}

static int fill_args(const intc_plan *p, const double *sph, double *out, IntcEriArgs *a)
{
    DevTables *t;
    int rc = ensure_tables(p->device, &t);
    if (rc) return rc;
    memset(a, 0, sizeof(*a));
    a->boys_mode = g_boys_mode;
    a->pair_off = p->pair_off;
    a->pair_np = p->prim_cut > 0.0 ? p->pair_np_scr : p->pair_np;
    a->pair_nk = p->pair_nk;
    a->pair_A = p->pairA;
    a->pair_AB = p->pairAB;
    a->pp_zeta = p->pz;
    a->pp_o2z = p->po2z;
    a->pp_px = p->ppx;
    a->pp_py = p->ppy;
    a->pp_pz = p->ppz;
    a->pp_kappa = p->pkap;
    // no caller matrices: the built-in real solid harmonics (the register
    // kernels then use their compiled-in copy)
    a->sph = sph ? sph : t->sph;
    a->sph_builtin = sph ? 0 : 1;
    a->boys_cheb = t->cheb;
    a->boys_small = t->small;
    a->boys_medium = t->medium;
    a->out = out;
    return 0;
}

INTC_API int intc_plan_build_pairs(intc_plan *p, const double *exps, const double *coefs, const double *shell_cen,
                                   void *stream)
{
    if (p->n_pp == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    int blocks = (int)((p->n_pp + 255) / 256);
    hipLaunchKernelGGL(intc_build_pairs, dim3(blocks), dim3(256), 0, st, p->n_pp, p->pp_pair, p->pair_sa,
                       p->pair_sb, p->pair_off, p->pair_nk, p->shell_off, p->coef_off, p->cen_idx, shell_cen, exps,
                       coefs, p->pz, p->po2z,
                       p->ppx, p->ppy, p->ppz, p->pkap, p->pairA, p->pairAB);
    HIPCHK(hipGetLastError());
    if (p->prim_cut > 0.0) {
        hipLaunchKernelGGL(intc_screen_pairs, dim3((p->n_pair + 255) / 256), dim3(256), 0, st, p->n_pair,
                           p->pair_off, p->pair_np, p->prim_cut, p->pz, p->po2z, p->ppx, p->ppy, p->ppz, p->pkap,
                           p->pair_np_scr);
        HIPCHK(hipGetLastError());
    }
    p->pairs_built = true;
    return 0;
}

// Primitive-pair screen for the builds (execute, routed execute, graph
// capture, build_pairs) after this call; cut = 0 keeps every pair.  The pair
// data must be built again before intc_plan_run_bin.
INTC_API int intc_plan_set_prim_screen(intc_plan *p, double cut)
{
    if (!(cut >= 0.0) || !std::isfinite(cut)) return fail(-1, "prim_cut must be finite and >= 0");
    if (cut > 0.0 && !p->pair_np_scr && p->n_pair > 0) {
        if (hipMalloc(&p->pair_np_scr, sizeof(int) * p->n_pair) != hipSuccess)
            return fail(-3, "hipMalloc screened pair counts");
    }
    p->prim_cut = cut;
    p->pairs_built = false;
    return 0;
}

// Primitive pairs of every shell pair (plan pair order): all of them
// (screened = 0) or those the last build kept (screened = 1).
INTC_API int intc_plan_pair_prims(const intc_plan *p, int screened, int *out)
{
    if (p->n_pair == 0) return 0;
    if (screened && (p->prim_cut <= 0.0 || !p->pairs_built))
        return fail(-1, "no screened build: set a primitive cut and build the pair data first");
    HIPCHK(hipMemcpy(out, screened ? p->pair_np_scr : p->pair_np, sizeof(int) * p->n_pair, hipMemcpyDeviceToHost));
    return 0;
}

INTC_API int intc_plan_run_bin(intc_plan *p, int i, int prec, const double *sph, double *out, void *stream)
{
    if (i < 0 || i >= (int)p->bins.size()) return fail(-1, "bin index out of range");
    if (prec != 0 && prec != 1) return fail(-1, "prec must be 0 (FP64) or 1 (FP32 recurrences)");
    if (!p->pairs_built && p->n_pp > 0)
        return fail(-1, "pair data not built: call intc_plan_build_pairs (or execute) first");
    IntcEriArgs a;
    int rc = fill_args(p, sph, out, &a);
    if (rc) return rc;
    return launch_bin(p, p->bins[i], prec, a, nullptr, nullptr, (hipStream_t)stream);
}

// device route: one int8 per caller quartet (0 FP64, 1 FP32, 2 skip)
static int route_queues(intc_plan *p, const signed char *route, hipStream_t st)
{
    int nb = (int)p->bins.size();
    if (p->n_quartet == 0) return 0;
    const int use32 = p->route_fp32 ? 1 : 0, ident = p->identity_order ? 1 : 0;
    hipLaunchKernelGGL(intc_route_count, dim3(p->n_tile), dim3(256), 0, st, route, p->q_orig, ident, p->tile_q0,
                       p->tile_n, use32, p->tile_cnt);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(intc_route_scan, dim3(nb), dim3(256), 0, st, p->bin_tile0, p->tile_cnt, p->tile_base,
                       p->counts, p->redo_n);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(intc_route_scatter, dim3(p->n_tile), dim3(256), 0, st, route, p->q_orig, ident, p->tile_q0,
                       p->tile_n, p->tile_bin, p->bin_q0, p->tile_base, use32, p->idx64, p->idx32, p->idxs);
    HIPCHK(hipGetLastError());
    return 0;
}

// zeros the skipped blocks (after route_queues on the same or a joined stream)
static int zero_skipped(intc_plan *p, double *out, hipStream_t st)
{
    const int nb = (int)p->bins.size();
    if (p->n_quartet == 0 || nb == 0) return 0;
    hipLaunchKernelGGL(intc_zero_skipped, dim3(1024, nb), dim3(256), 0, st, p->bin_q0, p->counts, p->idxs,
                       p->q_off, p->q_size, out);
    HIPCHK(hipGetLastError());
    return 0;
}

// Kernel arguments of one bin in one precision (the per-class launch and the
// persistent kernel's argument array use the same fields).
static IntcEriArgs bin_args(const intc_plan *p, const IntcBin &b, int prec, const IntcEriArgs &base, bool routed,
                            int bi)
{
    IntcEriArgs a = base;
    a.q_bra = p->q_bra + b.q0;
    a.q_ket = p->q_ket + b.q0;
    a.q_off = p->q_off + b.q0;
    a.q_str = p->q_str + b.q0;
    a.q_idx = routed ? (prec == 0 ? p->idx64 : p->idx32) + b.q0 : nullptr;
    a.nq_dev = routed ? p->counts + 3 * bi + prec : nullptr;
    a.nq = b.nq;
    if (routed && prec == 1 && p->fp32_tol > 0.0) {
        a.fp32_tol = p->fp32_tol;
        a.redo_q = p->redo_q + b.q0;
        a.redo_n = p->redo_n + bi;
    }
    a.team_groups = b.team_groups[prec];
    const IntcClassEntry *c = b.cls[prec];
    a.wg_cls[0] = c->la;
    a.wg_cls[1] = c->lb;
    a.wg_cls[2] = c->lc;
    a.wg_cls[3] = c->ld;
    a.wg_cls[4] = c->sph;
    return a;
}

// LDS doubles per wave of a register class kernel at team size T
static int reg_doubles(const IntcClassEntry *c, int T)
{
    const int tpw = 64 / (c->G * T);
    return std::max(tpw * c->NE * c->NCD, 64 * c->NS);
}

static bool multi_covers(const intc_plan *p, const IntcBin &b, int prec)
{
    if (!p->persistent || b.cls[prec]->multi < 0) return false;
    if (p->multi_bin.empty()) return true;
    const ptrdiff_t i = &b - p->bins.data();
    return i >= 0 && i < (ptrdiff_t)p->multi_bin.size() && p->multi_bin[i];
}

// Argument array and tile list of the persistent launch.  A tile is a run of
// one covered bin's queue positions, whole team iterations of (teams per
// wave) quartets, at most tile_iters of them, closed early once its estimated
// cost reaches 1/8 of an average wave's share of the launch (so that no tile
// holds several of the heaviest quartets, which the per-class grids spread
// over separate waves).  Routed tiles cover the whole bin for each queue
// (the queue lengths are known on the device only; positions past a queue's
// length end at once).  Device buffers are rewritten only when their contents
// change, after the device has finished with them.
static int prepare_multi(intc_plan *p, const double *sph, double *out, bool routed, intc_plan::MultiBufs &mb)
{
    if (sph) return fail(-1, "the persistent launch uses the built-in transforms: pass sph = NULL");
    if (g_boys_mode != 0) return fail(-1, "the persistent launch uses the Chebyshev Boys function (mode 0)");
    IntcEriArgs base;
    int rc = fill_args(p, sph, out, &base);
    if (rc) return rc;
    const int nb = (int)p->bins.size();
    std::vector<IntcEriArgs> args(2 * std::max(1, nb));
    memset(args.data(), 0, sizeof(IntcEriArgs) * args.size());
    const int nprec = routed ? (p->route_fp32 ? 2 : 1) : 1;
    // cost model per quartet: per primitive quartet ~ G lanes x ket VRR size,
    // plus the contracted epilogue ~ NE x NCD
    auto per_prim = [](const IntcClassEntry *c) { return (double)c->G * ncc0(c->lc + c->ld) * (c->M + 1); };
    auto per_q = [](const IntcClassEntry *c) { return 4.0 * c->NE * c->NCD; };
    auto range_cost = [&](const IntcBin &b, const IntcClassEntry *c, long long lo, long long hi) {
        return (double)(p->prim_prefix[b.q0 + hi] - p->prim_prefix[b.q0 + lo]) * per_prim(c) +
               (double)(hi - lo) * per_q(c);
    };
    double total = 0.0;
    for (int bi = 0; bi < nb; ++bi)
        for (int prec = 0; prec < nprec; ++prec)
            if (multi_covers(p, p->bins[bi], prec)) total += range_cost(p->bins[bi], p->bins[bi].cls[prec], 0,
                                                                        p->bins[bi].nq);
    const double cap = total / (8.0 * 4.0 * p->multi_grid);
    std::vector<std::pair<double, int4>> tl;
    int reg = 0;
    for (int bi = 0; bi < nb; ++bi) {
        const IntcBin &b = p->bins[bi];
        for (int prec = 0; prec < nprec; ++prec) {
            if (!multi_covers(p, b, prec)) continue;
            const IntcClassEntry *c = b.cls[prec];
            args[2 * bi + prec] = bin_args(p, b, prec, base, routed, bi);
            const int T = b.team_groups[prec], tpw = 64 / (c->G * T);
            reg = std::max(reg, reg_doubles(c, T));
            for (long long lo = 0; lo < b.nq;) {
                long long hi = lo;
                double cost = 0.0;
                for (int it = 0; it < p->tile_iters && hi < b.nq && (it == 0 || cost < cap); ++it) {
                    const long long h2 = std::min<long long>(b.nq, hi + tpw);
                    cost += range_cost(b, c, hi, h2);
                    hi = h2;
                }
                tl.push_back({cost, make_int4(2 * bi + prec, c->multi, (int)lo, (int)hi)});
                lo = hi;
            }
        }
    }
    // Tile order: bins interleaved (the k-th tile of every bin before any
    // bin's (k+1)-th), so that the waves work on a mix of classes at any
    // time; measured on MI355X against largest-cost-first and a random
    // order (config 3: 49.0 vs 50.6 and 54.0 ms; config 1: 21.7 vs 27.7 and
    // 23.7 ms, 64 iterations per tile).
    {
        std::vector<std::pair<long long, size_t>> key(tl.size());
        std::vector<long long> cnt(args.size(), 0);
        for (size_t i = 0; i < tl.size(); ++i) key[i] = {cnt[tl[i].second.x]++, i};
        std::stable_sort(key.begin(), key.end());
        std::vector<std::pair<double, int4>> t2(tl.size());
        for (size_t i = 0; i < tl.size(); ++i) t2[i] = tl[key[i].second];
        tl.swap(t2);
    }
    std::vector<int4> tiles(tl.size());
    for (size_t i = 0; i < tl.size(); ++i) tiles[i] = tl[i].second;
    if (reg > 160 * 1024 / 8 / 4) return fail(-1, "persistent launch: LDS per block exceeds 160 KB");
    auto same4 = [](const std::vector<int4> &x, const std::vector<int4> &y) {
        if (x.size() != y.size()) return false;
        for (size_t i = 0; i < x.size(); ++i)
            if (x[i].x != y[i].x || x[i].y != y[i].y || x[i].z != y[i].z || x[i].w != y[i].w) return false;
        return true;
    };
    const bool args_same = mb.h_args.size() == args.size() &&
                           memcmp(mb.h_args.data(), args.data(), sizeof(IntcEriArgs) * args.size()) == 0;
    const bool tiles_same = same4(mb.h_tiles, tiles);
    if (args_same && tiles_same && mb.next) return 0;
    // the previous launches that read these buffers must have finished
    HIPCHK(hipDeviceSynchronize());
    if (!mb.args || mb.h_args.size() != args.size()) {
        if (mb.args) HIPCHK(hipFree(mb.args));
        mb.args = nullptr;
        HIPCHK(hipMalloc(&mb.args, sizeof(IntcEriArgs) * args.size()));
    }
    if (!mb.tiles || mb.cap_tiles < (int)std::max<size_t>(1, tiles.size())) {
        if (mb.tiles) HIPCHK(hipFree(mb.tiles));
        mb.tiles = nullptr;
        mb.cap_tiles = (int)std::max<size_t>(1, tiles.size());
        HIPCHK(hipMalloc(&mb.tiles, sizeof(int4) * mb.cap_tiles));
    }
    if (!mb.next) HIPCHK(hipMalloc(&mb.next, sizeof(int)));
    HIPCHK(hipMemcpy(mb.args, args.data(), sizeof(IntcEriArgs) * args.size(), hipMemcpyHostToDevice));
    if (!tiles.empty()) HIPCHK(hipMemcpy(mb.tiles, tiles.data(), sizeof(int4) * tiles.size(), hipMemcpyHostToDevice));
    mb.h_args = args;
    mb.h_tiles = tiles;
    mb.n_tiles = (int)tiles.size();
    mb.reg = std::max(reg, 1);
    return 0;
}

// the FP32 error check applies to a bin's FP32 queue when its FP32 kernel is
// a register kernel of its own (FP32-routed quartets of other bins run the
// FP64 kernel, the workgroup kernels carry no estimate)
static bool redo_applies(const intc_plan *p, const IntcBin &b)
{
    return p->fp32_tol > 0.0 && b.cls[1] && b.cls[1]->prec == 1 && !b.cls[1]->wg;
}

// FP64 recomputation of the quartets that a bin's FP32 launch queued (device-
// side count; the class kernels stride over any queue length with this grid).
// Ordered after that launch on the same stream.
static int launch_redo(const intc_plan *p, const IntcBin &b, int bi, const IntcEriArgs &a, hipStream_t st)
{
    IntcBin br = b;
    br.grid[0] = std::min(b.grid[0], 256 * 2);
    return launch_bin(p, br, 0, a, p->redo_q + b.q0, p->redo_n + bi, st);
}

static int run_all(intc_plan *p, const double *sph, double *out, const signed char *route, int n_streams,
                   hipStream_t st, const intc_plan::MultiBufs *mb = nullptr)
{
    IntcEriArgs a;
    int rc = fill_args(p, sph, out, &a);
    if (rc) return rc;
    if (route) {
        rc = route_queues(p, route, st);
        if (rc) return rc;
    }
    if (n_streams > 1) {
        if (hipEventRecord(p->events[0], st) != hipSuccess) return fail(-3, "event record");
        for (int k = 0; k < n_streams - 1; ++k)
            if (hipStreamWaitEvent(p->side[k], p->events[0], 0) != hipSuccess) return fail(-3, "stream wait");
    }
    // Skipped blocks are disjoint from every block a class kernel writes, so
    // the HBM-bound zeroing runs beside the arithmetic-bound bins: first on the
    // last stream (the stream that receives the fewest bins).
    if (route) {
        rc = zero_skipped(p, out, n_streams > 1 ? p->side[n_streams - 2] : st);
        if (rc) return rc;
    }
    // largest bins first, round-robin over the streams
    std::vector<int> ord(p->bins.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = (int)i;
    std::sort(ord.begin(), ord.end(), [&](int x, int y) {
        return (double)p->bins[x].nq * p->bins[x].cls[0]->NE * p->bins[x].cls[0]->NCD >
               (double)p->bins[y].nq * p->bins[y].cls[0]->NE * p->bins[y].cls[0]->NCD;
    });
    int k = 0;
    if (mb) {
        // the persistent kernel on the launch stream; uncovered bins on the others
        if (mb->n_tiles > 0) {
            hipError_t e = intc_multi_launch(mb->args, mb->tiles, mb->n_tiles, mb->next, mb->reg, p->multi_grid, st);
            if (e != hipSuccess) return fail(-3, std::string("persistent kernel launch: ") + hipGetErrorString(e));
            if (route && p->route_fp32)
                for (size_t bi = 0; bi < p->bins.size(); ++bi)
                    if (multi_covers(p, p->bins[bi], 1) && redo_applies(p, p->bins[bi])) {
                        rc = launch_redo(p, p->bins[bi], (int)bi, a, st);
                        if (rc) return rc;
                    }
        }
        k = n_streams > 1 ? 1 : 0;
    }
    const int nprec = route ? (p->route_fp32 ? 2 : 1) : 1;
    for (int bi : ord) {
        const IntcBin &b = p->bins[bi];
        for (int prec = 0; prec < nprec; ++prec) {
            if (mb && multi_covers(p, b, prec)) continue;
            hipStream_t s2 = (k % n_streams) == 0 ? st : p->side[(k % n_streams) - 1];
            ++k;
            const int *qi = route ? (prec == 0 ? p->idx64 : p->idx32) + b.q0 : nullptr;
            const int *nd = route ? p->counts + 3 * bi + prec : nullptr;
            if (route) {
                // queue lengths are only known on the device: persistent grid
                IntcBin br = b;
                br.grid[prec] = std::min(b.grid[prec], 256 * 16);
                const bool redo = prec == 1 && redo_applies(p, b);
                IntcEriArgs a2 = a;
                if (redo) {
                    a2.fp32_tol = p->fp32_tol;
                    a2.redo_q = p->redo_q + b.q0;
                    a2.redo_n = p->redo_n + bi;
                }
                rc = launch_bin(p, br, prec, a2, qi, nd, s2);
                if (!rc && redo) rc = launch_redo(p, b, bi, a, s2);
            } else {
                rc = launch_bin(p, b, prec, a, qi, nd, s2);
            }
            if (rc) return rc;
        }
    }
    if (n_streams > 1) {
        for (int j = 0; j < n_streams - 1; ++j) {
            if (hipEventRecord(p->events[j + 1], p->side[j]) != hipSuccess) return fail(-3, "event record");
            if (hipStreamWaitEvent(st, p->events[j + 1], 0) != hipSuccess) return fail(-3, "stream wait");
        }
    }
    return 0;
}

static int ensure_streams(intc_plan *p, int n_streams)
{
    while ((int)p->side.size() < n_streams - 1) {
        hipStream_t s2;
        HIPCHK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        p->side.push_back(s2);
    }
    while ((int)p->events.size() < n_streams) {
        hipEvent_t e;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        p->events.push_back(e);
    }
    return 0;
}

INTC_API int intc_plan_execute(intc_plan *p, const double *exps, const double *coefs, const double *shell_cen,
                               const double *sph, double *out, void *stream)
{
    if (p->persistent) {
        int rc = prepare_multi(p, sph, out, false, p->m_eager);
        if (rc) return rc;
    }
    int rc = intc_plan_build_pairs(p, exps, coefs, shell_cen, stream);
    if (rc) return rc;
    return run_all(p, sph, out, nullptr, 1, (hipStream_t)stream, p->persistent ? &p->m_eager : nullptr);
}

// Persistent multi-class launch for execute, execute_routed and graphs
// captured afterwards: the bins whose class the multi kernel holds run as
// tiles of one launch (spherical plans, built-in transforms, Chebyshev Boys).
INTC_API int intc_plan_set_persistent(intc_plan *p, int enable, int tile_iters, int grid)
{
    if (enable && !p->spherical) return fail(-2, "the persistent kernel holds the spherical class kernels only");
    if (tile_iters < 1 || tile_iters > (1 << 20)) return fail(-1, "tile_iters must be in [1, 2^20]");
    if (grid < 0 || grid > 256 * 64) return fail(-1, "grid must be in [0, 16384] (0: two blocks per CU)");
    p->persistent = enable != 0;
    p->tile_iters = tile_iters;
    p->multi_grid = grid ? grid : 2 * 256;
    return 0;
}

// Bins the persistent launch may hold (cover[i] != 0 for bin i); n = 0 clears
// the selection (every bin whose class the multi kernel holds).  The other
// bins keep their own class kernels on the side streams.
INTC_API int intc_plan_set_persistent_bins(intc_plan *p, int n, const int *cover)
{
    if (n != 0 && n != (int)p->bins.size()) return fail(-1, "cover needs one entry per bin (or n = 0)");
    p->multi_bin.assign(n, 0);
    for (int i = 0; i < n; ++i) p->multi_bin[i] = cover[i] != 0;
    return 0;
}

// Tiles and LDS doubles per wave of the last persistent eager execution.
INTC_API int intc_plan_persistent_info(const intc_plan *p, int *n_tiles, int *reg_doubles)
{
    *n_tiles = p->m_eager.n_tiles;
    *reg_doubles = p->m_eager.reg;
    return 0;
}

INTC_API int intc_plan_execute_routed(intc_plan *p, const double *exps, const double *coefs,
                                      const double *shell_cen, const double *sph, const signed char *route,
                                      double *out, void *stream)
{
    if (!route) return fail(-1, "route array is required");
    if (p->strided)
        return fail(-1, "routed execution needs dense output blocks (a plan without caller strides or contractions)");
    if (p->persistent) {
        int rc = prepare_multi(p, sph, out, true, p->m_eager);
        if (rc) return rc;
    }
    int rc = intc_plan_build_pairs(p, exps, coefs, shell_cen, stream);
    if (rc) return rc;
    return run_all(p, sph, out, route, 1, (hipStream_t)stream, p->persistent ? &p->m_eager : nullptr);
}

// Capture pair building, the optional route compaction and all class bins
// into one hipGraph.  With n_streams > 1 the bins are spread over side
// streams forked from `stream`, so the graph has independent bin nodes that
// the device runs concurrently.
INTC_API int intc_plan_graph_create(intc_plan *p, const double *exps, const double *coefs, const double *shell_cen,
                                    const double *sph, const signed char *route, double *out, int n_streams,
                                    void *stream)
{
    if (n_streams < 1 || n_streams > 16) return fail(-1, "n_streams must be in [1, 16]");
    if (route && p->strided)
        return fail(-1, "routed execution needs dense output blocks (a plan without caller strides or contractions)");
    if (p->graph_exec) {
        (void)hipGraphExecDestroy(p->graph_exec);
        p->graph_exec = nullptr;
    }
    if (p->graph) {
        (void)hipGraphDestroy(p->graph);
        p->graph = nullptr;
    }
    int rc = ensure_streams(p, n_streams);
    if (rc) return rc;
    {
        DevTables *t;
        rc = ensure_tables(p->device, &t);
        if (rc) return rc;
    }
    if (p->persistent) {
        // the graph binds its own argument and tile buffers
        rc = prepare_multi(p, sph, out, route != nullptr, p->m_graph);
        if (rc) return rc;
    }
    hipStream_t st = (hipStream_t)stream;
    bool own = false;
    if (!st) {
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        own = true;
    }
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    rc = intc_plan_build_pairs(p, exps, coefs, shell_cen, st);
    if (!rc) rc = run_all(p, sph, out, route, n_streams, st, p->persistent ? &p->m_graph : nullptr);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (own) (void)hipStreamDestroy(st);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(-3, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    p->graph = g;
    HIPCHK(hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0));
    return 0;
}

INTC_API int intc_plan_graph_launch(intc_plan *p, void *stream)
{
    if (!p->graph_exec) return fail(-1, "no graph captured for this plan");
    HIPCHK(hipGraphLaunch(p->graph_exec, (hipStream_t)stream));
    return 0;
}

// routed runs: 1 = FP64 and FP32-recurrence queues (default), 0 = route code
// 1 is treated as FP64 and only the FP64 kernels are launched
INTC_API int intc_plan_set_route_fp32(intc_plan *p, int use_fp32)
{
    p->route_fp32 = use_fp32 != 0;
    return 0;
}

// FP32 error check of routed runs (tol > 0; 0 turns it off).  Set it before
// capturing a graph: graphs keep the setting they were captured with.
INTC_API int intc_plan_set_fp32_check(intc_plan *p, double tol)
{
    if (!(tol >= 0.0) || !std::isfinite(tol)) return fail(-1, "tol must be finite and >= 0 (0: off)");
    if (tol > 0.0 && !p->redo_q) {
        const size_t nb = std::max<size_t>(1, p->bins.size());
        if (hipMalloc(&p->redo_q, sizeof(int) * std::max<long long>(1, p->n_quartet)) != hipSuccess ||
            hipMalloc(&p->redo_n, sizeof(int) * nb) != hipSuccess)
            return fail(-3, "hipMalloc FP32 redo queues");
        HIPCHK(hipMemset(p->redo_n, 0, sizeof(int) * nb));
    }
    p->fp32_tol = tol;
    return 0;
}

// per bin: quartets the last routed run recomputed in FP64 after the FP32
// error check (host copy, synchronising; zeros when the check is off)
INTC_API int intc_plan_redo_counts(intc_plan *p, int *counts_out)
{
    const int nb = (int)p->bins.size();
    if (nb == 0) return 0;
    if (!p->redo_n) {
        std::fill(counts_out, counts_out + nb, 0);
        return 0;
    }
    HIPCHK(hipMemcpy(counts_out, p->redo_n, sizeof(int) * nb, hipMemcpyDeviceToHost));
    return 0;
}

// plan order: order[i] = the caller's index of the quartet at plan position i
INTC_API int intc_plan_quartet_order(intc_plan *p, int *order)
{
    if (p->n_quartet == 0) return 0;
    HIPCHK(hipMemcpy(order, p->q_orig, sizeof(int) * p->n_quartet, hipMemcpyDeviceToHost));
    return 0;
}

// queue lengths of the last routed execution (host copy, synchronising)
INTC_API int intc_plan_route_counts(intc_plan *p, int *counts_out)
{
    int nb = (int)p->bins.size();
    if (nb == 0) return 0;
    HIPCHK(hipMemcpy(counts_out, p->counts, sizeof(int) * 3 * nb, hipMemcpyDeviceToHost));
    return 0;
}

INTC_API int intc_plan_destroy(intc_plan *p)
{
    if (!p) return 0;
    plan_free(p);
    delete p;
    return 0;
}

// ---------------------------------------------------------------------------
// per-quartet twins of the generated functions
// ---------------------------------------------------------------------------
static long hrr_layer(int la, int lb)
{
    long best = 0;
    for (int lq = 0; lq <= la; ++lq) {
        long s = (long)(ncc0(la + lb - lq) - ncc0(lb - 1)) * ncc(lq);
        best = std::max(best, s);
    }
    return best;
}

static int work_size(int spherical, int la, int lb, int lc, int ld, int na, int nb, int nc, int nd, int ka, int kb,
                     int kc, int kd)
{
    int L = la + lb, M = L + lc + ld;
    long nvrr = (long)ncc0(L) * ncc0(lc) * ncc0(ld);
    long nprim = (long)na * nb * nc * nd;
    long mem = 2 * (nvrr * nprim + M + 1);
    int np[4] = {na, nb, nc, nd}, nk[4] = {ka, kb, kc, kd};
    long nmax = nprim;
    for (int dir = 0; dir < 2; ++dir) {
        int cur[4] = {np[0], np[1], np[2], np[3]};
        for (int s = 0; s < 4; ++s) {
            int k = dir == 0 ? s : 3 - s;
            cur[k] = nk[k];
            nmax = std::max(nmax, (long)cur[0] * cur[1] * cur[2] * cur[3]);
        }
    }
    mem = std::max(mem, 2 * nvrr * nmax);
    long ncont = (long)ka * kb * kc * kd;
    int nc_c = spherical ? nsph(lc) : ncc(lc), nc_d = spherical ? nsph(ld) : ncc(ld);
    mem = std::max(mem, 2 * hrr_layer(lb, la) * nc_c * nc_d * ncont);
    long late = spherical ? 2L * ncc(la) * ncc(lb) * ncc(lc) * ncc(ld) * ncont
                          : 2L * ncc(la) * ncc(lb) * nc_c * nc_d * ncont;
    mem = std::max(mem, late);
    return (int)mem;
}

INTC_API int intc_eri_cont_sph_work_array_size(int la, int lb, int lc, int ld, int na_prim, int nb_prim,
                                               int nc_prim, int nd_prim, int na_cont, int nb_cont, int nc_cont,
                                               int nd_cont)
{
    return work_size(1, la, lb, lc, ld, na_prim, nb_prim, nc_prim, nd_prim, na_cont, nb_cont, nc_cont, nd_cont);
}

INTC_API int intc_eri_cont_cart_work_array_size(int la, int lb, int lc, int ld, int na_prim, int nb_prim,
                                                int nc_prim, int nd_prim, int na_cont, int nb_cont, int nc_cont,
                                                int nd_cont)
{
    return work_size(0, la, lb, lc, ld, na_prim, nb_prim, nc_prim, nd_prim, na_cont, nb_cont, nc_cont, nd_cont);
}

struct ShellArg {
    const double *x, *c;
    int l, n, iskip, n_prim, n_cont, iskip_cont;
    const double *contract, *sph;
};

// One contracted quartet (all contraction combinations) through the batched
// device path: upload, plan of ncont_a*..*ncont_d quartets, run, scatter.
static int per_quartet(int spherical, const ShellArg sa[4], double *output)
{
    int rc = 0;
    int nk = 1;
    for (int k = 0; k < 4; ++k) {
        if (sa[k].l < 0 || sa[k].l > 7 || sa[k].n_prim < 1 || sa[k].n_cont < 1)
            return fail(-1, "invalid shell arguments");
        int want = spherical ? nsph(sa[k].l) : ncc(sa[k].l);
        if (sa[k].n != want) return fail(-1, "component count n does not match l");
        nk *= sa[k].n_cont;
    }
    // shells: one per (index, contraction) with that contraction's column
    std::vector<int> sl, snp, soff;
    std::vector<double> cen, ex, co;
    int base[4];
    for (int k = 0; k < 4; ++k) {
        base[k] = (int)sl.size();
        for (int j = 0; j < sa[k].n_cont; ++j) {
            sl.push_back(sa[k].l);
            snp.push_back(sa[k].n_prim);
            soff.push_back((int)ex.size());
            for (int i = 0; i < 3; ++i) cen.push_back(sa[k].c[i]);
            for (int i = 0; i < sa[k].n_prim; ++i) {
                ex.push_back(sa[k].x[i]);
                co.push_back(sa[k].contract[i + j * sa[k].n_prim]);
            }
        }
    }
    int nblk = 1;
    for (int k = 0; k < 4; ++k) nblk *= sa[k].n;
    std::vector<int> quart;
    std::vector<long long> off;
    std::vector<int> combo;
    for (int ka = 0; ka < sa[0].n_cont; ++ka)
        for (int kb = 0; kb < sa[1].n_cont; ++kb)
            for (int kc = 0; kc < sa[2].n_cont; ++kc)
                for (int kd = 0; kd < sa[3].n_cont; ++kd) {
                    quart.push_back(base[0] + ka);
                    quart.push_back(base[1] + kb);
                    quart.push_back(base[2] + kc);
                    quart.push_back(base[3] + kd);
                    off.push_back((long long)(off.size()) * nblk);
                    combo.push_back(ka); combo.push_back(kb); combo.push_back(kc); combo.push_back(kd);
                }
    intc_plan *p = nullptr;
    rc = intc_plan_create(&p, spherical, (int)sl.size(), sl.data(), snp.data(), soff.data(), cen.data(), ex.data(),
                          co.data(), (long long)nk, quart.data(), off.data());
    if (rc) return rc;
    double *d_ex = nullptr, *d_co = nullptr, *d_cen = nullptr, *d_sph = nullptr, *d_out = nullptr;
    std::vector<double> sph_all;
    if (spherical) {
        sph_all.assign(intc_sph_off(8), 0.0);
        for (int k = 0; k < 4; ++k) {
            int l = sa[k].l;
            // one S_l per l reaches the kernels
            for (int k2 = 0; k2 < k; ++k2)
                if (sa[k2].l == l && memcmp(sa[k2].sph, sa[k].sph, sizeof(double) * ncc(l) * nsph(l)) != 0) {
                    intc_plan_destroy(p);
                    return fail(-1, "indices with the same l must pass the same sph_trans_array");
                }
            for (int i = 0; i < ncc(l) * nsph(l); ++i) sph_all[intc_sph_off(l) + i] = sa[k].sph[i];
        }
    }
    std::vector<double> res((size_t)nk * nblk);
    auto cleanup = [&]() {
        if (d_ex) (void)hipFree(d_ex);
        if (d_co) (void)hipFree(d_co);
        if (d_cen) (void)hipFree(d_cen);
        if (d_sph) (void)hipFree(d_sph);
        if (d_out) (void)hipFree(d_out);
        intc_plan_destroy(p);
    };
#define PQCHK(x)                                                                       \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) { cleanup(); return fail(-3, std::string(#x) + ": " + hipGetErrorString(e_)); } \
    } while (0)
    PQCHK(hipMalloc(&d_ex, sizeof(double) * ex.size()));
    PQCHK(hipMalloc(&d_co, sizeof(double) * co.size()));
    PQCHK(hipMalloc(&d_cen, sizeof(double) * cen.size()));
    PQCHK(hipMalloc(&d_out, sizeof(double) * res.size()));
    PQCHK(hipMemcpy(d_ex, ex.data(), sizeof(double) * ex.size(), hipMemcpyHostToDevice));
    PQCHK(hipMemcpy(d_co, co.data(), sizeof(double) * co.size(), hipMemcpyHostToDevice));
    PQCHK(hipMemcpy(d_cen, cen.data(), sizeof(double) * cen.size(), hipMemcpyHostToDevice));
    if (spherical) {
        PQCHK(hipMalloc(&d_sph, sizeof(double) * sph_all.size()));
        PQCHK(hipMemcpy(d_sph, sph_all.data(), sizeof(double) * sph_all.size(), hipMemcpyHostToDevice));
    }
    rc = intc_plan_execute(p, d_ex, d_co, d_cen, d_sph, d_out, nullptr);
    if (rc) { cleanup(); return rc; }
    PQCHK(hipDeviceSynchronize());
    PQCHK(hipMemcpy(res.data(), d_out, sizeof(double) * res.size(), hipMemcpyDeviceToHost));
#undef PQCHK
    cleanup();
    // scatter: out[sum_X(iX*iskipX + icontX*iskipX_cont)]
    for (int k = 0; k < nk; ++k) {
        const int *cb = &combo[4 * k];
        const double *blk = &res[(size_t)k * nblk];
        size_t i = 0;
        for (int a = 0; a < sa[0].n; ++a)
            for (int b = 0; b < sa[1].n; ++b)
                for (int c = 0; c < sa[2].n; ++c)
                    for (int d = 0; d < sa[3].n; ++d, ++i)
                        output[(long)a * sa[0].iskip + (long)cb[0] * sa[0].iskip_cont + (long)b * sa[1].iskip +
                               (long)cb[1] * sa[1].iskip_cont + (long)c * sa[2].iskip +
                               (long)cb[2] * sa[2].iskip_cont + (long)d * sa[3].iskip +
                               (long)cb[3] * sa[3].iskip_cont] = blk[i];
    }
    return 0;
}

// Allocation-free variant of per_quartet for spherical output: the same
// canonical class, pair orientation, team size and kernels, with every device
// buffer in this device's `scratch` (intc_init must have run on the device).
// Host memory: a NoAllocHead and a chunk buffer on the stack.  Limits: n_prim
// <= NA_PRIM, n_cont <= NA_CONT per index, all blocks <= NA_OUT_DOUBLES doubles.
// Calls that overlap on one GPU share the array and race.
static int per_quartet_noalloc(const ShellArg sa[4], double *output)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    double *scr = nullptr;
    const DevTables *tab = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (dev < NA_MAX_DEV) scr = g_scratch_ptr[dev];
        if (dev < (int)g_tables.size() && g_tables[dev].cheb) tab = &g_tables[dev];
    }
    if (!scr || !tab) return fail(-1, "intc_init() has not run on this device");
    int nk = 1, nblk = 1, ls[4];
    for (int k = 0; k < 4; ++k) {
        const ShellArg &x = sa[k];
        if (x.l < 0 || x.l > 7 || x.n_prim < 1 || x.n_cont < 1) return fail(-1, "invalid shell arguments");
        if (x.n_prim > NA_PRIM || x.n_cont > NA_CONT)
            return fail(-1, "noalloc entry point: at most 8 primitives and 2 contractions per shell");
        if (x.n != nsph(x.l)) return fail(-1, "component count n does not match l");
        nk *= x.n_cont;
        nblk *= x.n;
        ls[k] = x.l;
    }
    if ((size_t)nk * nblk > NA_OUT_DOUBLES) return fail(-1, "noalloc entry point: output blocks exceed the scratch size");
    int cls[4], perm0[4];
    canonical(ls, cls, perm0);
    const IntcClassEntry *c = find_class(cls[0], cls[1], cls[2], cls[3], 1, 0);
    if (!c) return class_error(ls[0], ls[1], ls[2], ls[3], 1);
    NoAllocHead h;
    memset(&h, 0, sizeof(h));
    // shells: one per (index, contraction) with that contraction's column
    int base[4], ns = 0, np = 0;
    for (int k = 0; k < 4; ++k) {
        base[k] = ns;
        for (int j = 0; j < sa[k].n_cont; ++j, ++ns) {
            h.shell_off[ns] = np;
            h.vid[ns] = ns;
            for (int i = 0; i < 3; ++i) h.cen[3 * ns + i] = sa[k].c[i];
            for (int i = 0; i < sa[k].n_prim; ++i, ++np) {
                h.exps[np] = sa[k].x[i];
                h.coefs[np] = sa[k].contract[i + j * sa[k].n_prim];
            }
        }
    }
    // the caller's S_l, one per l: indices with the same l must pass the same matrix
    for (int k = 0; k < 4; ++k) {
        const int l = sa[k].l, n = ncc(l) * nsph(l);
        for (int k2 = 0; k2 < k; ++k2)
            if (sa[k2].l == l && memcmp(sa[k2].sph, sa[k].sph, sizeof(double) * n) != 0)
                return fail(-1, "indices with the same l must pass the same sph_trans_array");
        memcpy(h.sph + intc_sph_off(l), sa[k].sph, sizeof(double) * n);
    }
    int snp[NA_SHELL];
    for (int k = 0; k < 4; ++k)
        for (int j = 0; j < sa[k].n_cont; ++j) snp[base[k] + j] = sa[k].n_prim;
    const int st_dense[4] = {sa[1].n * sa[2].n * sa[3].n, sa[2].n * sa[3].n, sa[3].n, 1};
    int nq = 0, npair = 0, npp = 0, prims = 0;
    for (int ka = 0; ka < sa[0].n_cont; ++ka)
        for (int kb = 0; kb < sa[1].n_cont; ++kb)
            for (int kc = 0; kc < sa[2].n_cont; ++kc)
                for (int kd = 0; kd < sa[3].n_cont; ++kd, ++nq) {
                    const int qs[4] = {base[0] + ka, base[1] + kb, base[2] + kc, base[3] + kd};
                    int perm[4] = {perm0[0], perm0[1], perm0[2], perm0[3]};
                    for (int hh = 0; hh < 4; hh += 2)
                        if (cls[hh] == cls[hh + 1] &&
                            vrr_on_second(qs[perm[hh]], qs[perm[hh + 1]], snp, h.shell_off, h.cen, h.exps, h.coefs))
                            std::swap(perm[hh], perm[hh + 1]);
                    for (int side = 0; side < 2; ++side, ++npair) {
                        const int s0 = qs[perm[2 * side]], s1 = qs[perm[2 * side + 1]];
                        h.pair_sa[npair] = s0;
                        h.pair_sb[npair] = s1;
                        h.pair_off[npair] = npp;
                        h.pair_np[npair] = snp[s0] * snp[s1];
                        h.pair_nk[npair] = snp[s1];
                        for (int t = 0; t < snp[s0] * snp[s1]; ++t) h.pp_pair[npp++] = npair;
                    }
                    h.q_bra[nq] = npair - 2;
                    h.q_ket[nq] = npair - 1;
                    h.q_off[nq] = (long long)nq * nblk;
                    unsigned long long pk = 0;
                    for (int k = 0; k < 4; ++k) pk |= (unsigned long long)(st_dense[perm[k]] & 0xffff) << (16 * k);
                    h.q_str[nq] = pk;
                    prims += snp[qs[0]] * snp[qs[1]] * snp[qs[2]] * snp[qs[3]];
                }
    // device view of the scratch layout
    const size_t o_head = 0, o_pp = NA_HEAD_DOUBLES;
    NoAllocHead *dh = (NoAllocHead *)(scr + o_head);
    double *pz = scr + o_pp, *po2z = pz + NA_PP, *ppx = po2z + NA_PP, *ppy = ppx + NA_PP, *ppz = ppy + NA_PP,
           *pkap = ppz + NA_PP, *pairA = pkap + NA_PP, *pairAB = pairA + 3 * NA_PAIR;
    double *dout = scr + NA_OUT_OFFSET;
    HIPCHK(hipMemcpy(dh, &h, sizeof(h), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(intc_build_pairs, dim3((npp + 255) / 256), dim3(256), 0, (hipStream_t)0, (long long)npp,
                       dh->pp_pair, dh->pair_sa, dh->pair_sb, dh->pair_off, dh->pair_nk, dh->shell_off, dh->shell_off,
                       dh->vid, dh->cen,
                       dh->exps, dh->coefs, pz, po2z, ppx, ppy, ppz, pkap, pairA, pairAB);
    HIPCHK(hipGetLastError());
    IntcEriArgs a;
    memset(&a, 0, sizeof(a));
    a.boys_mode = g_boys_mode;
    a.pair_off = dh->pair_off;
    a.pair_np = dh->pair_np;
    a.pair_nk = dh->pair_nk;
    a.pair_A = pairA;
    a.pair_AB = pairAB;
    a.pp_zeta = pz;
    a.pp_o2z = po2z;
    a.pp_px = ppx;
    a.pp_py = ppy;
    a.pp_pz = ppz;
    a.pp_kappa = pkap;
    a.sph = dh->sph;
    a.sph_builtin = 0;
    a.boys_cheb = tab->cheb;
    a.boys_small = tab->small;
    a.boys_medium = tab->medium;
    a.out = dout;
    a.q_bra = dh->q_bra;
    a.q_ket = dh->q_ket;
    a.q_off = dh->q_off;
    a.q_str = dh->q_str;
    a.nq = nq;
    // team size and grid as a plan bin of these quartets (finish_plan, grid_for)
    IntcBin b;
    b.cls[0] = b.cls[1] = c;
    b.q0 = 0;
    b.nq = nq;
    b.block = 256;
    b.team_groups[0] = b.team_groups[1] = default_team_groups(c, nq, (double)prims / nq);
    b.grid[0] = b.grid[1] = grid_for(b, 0);
    a.team_groups = b.team_groups[0];
    a.wg_cls[0] = c->la;
    a.wg_cls[1] = c->lb;
    a.wg_cls[2] = c->lc;
    a.wg_cls[3] = c->ld;
    a.wg_cls[4] = c->sph;
    const int tpw = 64 / (c->G * a.team_groups);
    const size_t lds = c->wg ? 0 : (size_t)(b.block / 64) * tpw * c->NE * c->NCD * sizeof(double);
    hipError_t e = c->fn(&a, b.grid[0], b.block, lds, (hipStream_t)0);
    if (e != hipSuccess) return fail(-3, std::string("class kernel launch: ") + hipGetErrorString(e));
    HIPCHK(hipDeviceSynchronize());
    // results back in chunks, scattered to out[sum_X(iX*iskipX + icontX*iskipX_cont)]
    double buf[4096];
    const long long total = (long long)nq * nblk;
    long long done = 0;
    int q = 0, ia = 0, ib = 0, ic = 0, id = 0, ca = 0, cb = 0, cc = 0, cd = 0;
    while (done < total) {
        const int n = (int)std::min<long long>(4096, total - done);
        HIPCHK(hipMemcpy(buf, dout + done, sizeof(double) * n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) {
            output[(long)ia * sa[0].iskip + (long)ca * sa[0].iskip_cont + (long)ib * sa[1].iskip +
                   (long)cb * sa[1].iskip_cont + (long)ic * sa[2].iskip + (long)cc * sa[2].iskip_cont +
                   (long)id * sa[3].iskip + (long)cd * sa[3].iskip_cont] = buf[i];
            // next element: d fastest within a block, blocks in (ka, kb, kc, kd) order
            if (++id == sa[3].n) {
                id = 0;
                if (++ic == sa[2].n) {
                    ic = 0;
                    if (++ib == sa[1].n) {
                        ib = 0;
                        if (++ia == sa[0].n) {
                            ia = 0;
                            ++q;
                            if (++cd == sa[3].n_cont) {
                                cd = 0;
                                if (++cc == sa[2].n_cont) {
                                    cc = 0;
                                    if (++cb == sa[1].n_cont) {
                                        cb = 0;
                                        ++ca;
                                    }
                                }
                            }
                        }
                    }
                }
            }
        }
        done += n;
    }
    (void)q;
    return 0;
}

#define INTC_SHELL_PARAMS(X)                                                                        \
    const double *x##X, const double c##X[3], const int l##X, const int n##X, const int iskip##X,   \
        const int n##X##_prim, const int n##X##_cont, const int iskip##X##_cont,                    \
        const double *contract_array_##X

#define INTC_MK(X, S) {x##X, c##X, l##X, n##X, iskip##X, n##X##_prim, n##X##_cont, iskip##X##_cont, contract_array_##X, S}

INTC_API int intc_eri_cont_sph_main(INTC_SHELL_PARAMS(a), const double *sph_trans_array_a, INTC_SHELL_PARAMS(b),
                                    const double *sph_trans_array_b, INTC_SHELL_PARAMS(c),
                                    const double *sph_trans_array_c, INTC_SHELL_PARAMS(d),
                                    const double *sph_trans_array_d, double *work_array, double *output_array)
{
    (void)work_array;
    ShellArg s[4] = {INTC_MK(a, sph_trans_array_a), INTC_MK(b, sph_trans_array_b), INTC_MK(c, sph_trans_array_c),
                     INTC_MK(d, sph_trans_array_d)};
    return per_quartet(1, s, output_array);
}

// No counterpart in the reference: intc_eri_cont_sph_main without any device
// allocation (scratch in a static device array; limits above).
INTC_API int intc_eri_cont_sph_main_noalloc(INTC_SHELL_PARAMS(a), const double *sph_trans_array_a,
                                            INTC_SHELL_PARAMS(b), const double *sph_trans_array_b,
                                            INTC_SHELL_PARAMS(c), const double *sph_trans_array_c,
                                            INTC_SHELL_PARAMS(d), const double *sph_trans_array_d,
                                            double *work_array, double *output_array)
{
    (void)work_array;
    ShellArg s[4] = {INTC_MK(a, sph_trans_array_a), INTC_MK(b, sph_trans_array_b), INTC_MK(c, sph_trans_array_c),
                     INTC_MK(d, sph_trans_array_d)};
    return per_quartet_noalloc(s, output_array);
}

INTC_API int intc_eri_cont_cart_main(INTC_SHELL_PARAMS(a), INTC_SHELL_PARAMS(b), INTC_SHELL_PARAMS(c),
                                     INTC_SHELL_PARAMS(d), double *work_array, double *output_array)
{
    (void)work_array;
    ShellArg s[4] = {INTC_MK(a, nullptr), INTC_MK(b, nullptr), INTC_MK(c, nullptr), INTC_MK(d, nullptr)};
    return per_quartet(0, s, output_array);
}
