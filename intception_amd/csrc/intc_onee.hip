// intc_onee.hip -- one-electron integrals of the input-script API's other
// integral kinds: the 2-index overlap (a|b) and nuclear attraction (a|C|b)
// of the reference's example script (examples/example_input.py:79-110).
//
// Entry points (declared in include/intc.h):
//   * a batched plan over contracted shell-pair items (overlap) or
//     (shell pair, point-charge centre) items (nuclear attraction), binned by
//     (la, lb) and run by the generated class kernels (gen/onee_kernels.inc);
//   * per-call twins of the reference's generated <full>_main for these
//     kinds: primitive Cartesian (the example's options, contracted=False) and
//     contracted Cartesian / spherical.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "intc_common.h"
#include "gen/onee_kernels.inc"

#define INTC_API extern "C" __attribute__((visibility("default")))

int intc_internal_fail(int code, const char *msg);
int intc_internal_tables(const double **cheb, const double **small, const double **medium, const double **sph);
int intc_internal_boys_mode();

namespace {

inline int ncc(int l) { return (l + 1) * (l + 2) / 2; }
inline int nsph(int l) { return 2 * l + 1; }

int hip_fail(const char *what, hipError_t e)
{
    return intc_internal_fail(-3, (std::string(what) + ": " + hipGetErrorString(e)).c_str());
}

#define OCHK(x)                                                   \
    do {                                                          \
        hipError_t e_ = (x);                                      \
        if (e_ != hipSuccess) return hip_fail(#x, e_);            \
    } while (0)

const IntcOneEntry *find_one(int kind, int la, int lb, int sph)
{
    for (int i = 0; i < intc_n_one; ++i) {
        const IntcOneEntry &e = intc_one_table[i];
        if (e.kind == kind && e.la == la && e.lb == lb && e.sph == sph) return &e;
    }
    return nullptr;
}

template <class T>
int upload(T **d, const std::vector<T> &h)
{
    if (h.empty()) return 0;
    OCHK(hipMalloc(d, sizeof(T) * h.size()));
    OCHK(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return 0;
}

}  // namespace

struct intc_onee_plan {
    int kind = 0, spherical = 0, n_centre = 0;
    long long n = 0;
    struct Bin { const IntcOneEntry *e; long long first, count; };
    std::vector<Bin> bins;
    int *it_a = nullptr, *it_b = nullptr, *it_c = nullptr;
    long long *it_off = nullptr;
    unsigned long long *it_str = nullptr;
    int *shell_l = nullptr, *shell_np = nullptr, *shell_off = nullptr;
};

INTC_API int intc_onee_plan_destroy(intc_onee_plan *p)
{
    if (!p) return 0;
    void *ptrs[] = {p->it_a, p->it_b, p->it_c, p->it_off, p->it_str, p->shell_l, p->shell_np, p->shell_off};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    delete p;
    return 0;
}

// kind 0: overlap, items = n_item pairs (a, b); kind 1: nuclear attraction,
// items = n_item triples (a, b, c) with c indexing n_centre point-charge
// centres.  Item i writes a row-major [a][b] block at out[out_off[i]].
// Orientation of one pair: true when the primitive-pair centres P lie closer
// to b than to a (|c_a c_b| exp(-mu R^2) / zeta weighted mean of |P - X|^2);
// the class kernel then builds on b and the HRR moves angular momentum to a.
// Building on the far shell and transferring across |AB| cancels digits.
static bool build_on_b(int a, int b, const int *shell_nprim, const int *shell_off, const double *shell_cen,
                       const double *exps, const double *coefs)
{
    if (!shell_cen || !exps || !coefs) return false;
    double r2 = 0.0;
    for (int d = 0; d < 3; ++d) r2 += (shell_cen[3 * a + d] - shell_cen[3 * b + d]) *
                                      (shell_cen[3 * a + d] - shell_cen[3 * b + d]);
    if (r2 == 0.0) return false;
    double da = 0.0, db = 0.0;
    for (int i = 0; i < shell_nprim[a]; ++i)
        for (int j = 0; j < shell_nprim[b]; ++j) {
            const double al = exps[shell_off[a] + i], be = exps[shell_off[b] + j], oz = 1.0 / (al + be);
            const double w = fabs(coefs[shell_off[a] + i] * coefs[shell_off[b] + j]) * exp(-al * be * oz * r2) * oz;
            da += w * be * be * oz * oz;   // |P - A|^2 / R^2
            db += w * al * al * oz * oz;   // |P - B|^2 / R^2
        }
    return db < da;
}

INTC_API int intc_onee_plan_create(intc_onee_plan **plan, int kind, int spherical, int n_shell, const int *shell_l,
                                   const int *shell_nprim, const int *shell_off, const double *shell_cen,
                                   const double *exps, const double *coefs, int n_centre, long long n_item,
                                   const int *items, const long long *out_off)
{
    *plan = nullptr;
    if (kind != 0 && kind != 1) return intc_internal_fail(-1, "kind must be 0 (overlap) or 1 (nuclear attraction)");
    if (n_shell < 0 || n_item < 0 || n_centre < 0) return intc_internal_fail(-1, "negative size");
    if (n_item > 0x7fffffffLL) return intc_internal_fail(-1, "too many items for one plan");
    for (int s = 0; s < n_shell; ++s) {
        if (shell_l[s] < 0 || shell_l[s] > intc_one_lmax)
            return intc_internal_fail(-2, "one-electron kernels are built for l <= 3");
        if (shell_nprim[s] < 1 || shell_nprim[s] > 64) return intc_internal_fail(-1, "shell primitive count out of range");
    }
    const int w = kind == 0 ? 2 : 3;
    std::vector<long long> order(n_item);
    for (long long i = 0; i < n_item; ++i) {
        for (int k = 0; k < 2; ++k)
            if (items[w * i + k] < 0 || items[w * i + k] >= n_shell)
                return intc_internal_fail(-1, "item shell index out of range");
        if (kind == 1 && (items[w * i + 2] < 0 || items[w * i + 2] >= n_centre))
            return intc_internal_fail(-1, "item centre index out of range");
        order[i] = i;
    }
    // kernel orientation per item: (first, second) shell of the class kernel
    std::vector<unsigned char> swap(n_item, 0);
    for (long long i = 0; i < n_item; ++i)
        swap[i] = build_on_b(items[w * i], items[w * i + 1], shell_nprim, shell_off, shell_cen, exps, coefs) ? 1 : 0;
    auto first = [&](long long i) { return items[w * i + (swap[i] ? 1 : 0)]; };
    auto second = [&](long long i) { return items[w * i + (swap[i] ? 0 : 1)]; };
    auto cls = [&](long long i) { return shell_l[first(i)] * 16 + shell_l[second(i)]; };
    std::stable_sort(order.begin(), order.end(), [&](long long x, long long y) { return cls(x) < cls(y); });
    intc_onee_plan *p = new intc_onee_plan();
    p->kind = kind;
    p->spherical = spherical ? 1 : 0;
    p->n_centre = n_centre;
    p->n = n_item;
    std::vector<int> ia(n_item), ib(n_item), ic(n_item, 0);
    std::vector<long long> off(n_item);
    std::vector<unsigned long long> str(n_item);
    for (long long k = 0; k < n_item; ++k) {
        const long long i = order[k];
        ia[k] = first(i);
        ib[k] = second(i);
        if (kind == 1) ic[k] = items[w * i + 2];
        off[k] = out_off[i];
        // caller layout [a][b]: a stride n_b, b stride 1; the kernel's first
        // shell is b when the item is swapped
        const unsigned long long nb = (unsigned long long)(spherical ? nsph(shell_l[items[w * i + 1]])
                                                                     : ncc(shell_l[items[w * i + 1]]));
        str[k] = swap[i] ? (1ull | (nb << 32)) : (nb | (1ull << 32));
    }
    for (long long k = 0; k < n_item;) {
        long long j = k;
        while (j < n_item && cls(order[j]) == cls(order[k])) ++j;
        const IntcOneEntry *e = find_one(kind, shell_l[ia[k]], shell_l[ib[k]], p->spherical);
        if (!e) { intc_onee_plan_destroy(p); return intc_internal_fail(-2, "one-electron class not built"); }
        p->bins.push_back({e, k, j - k});
        k = j;
    }
    std::vector<int> sl(shell_l, shell_l + n_shell), snp(shell_nprim, shell_nprim + n_shell),
        soff(shell_off, shell_off + n_shell);
    int rc = 0;
    rc |= upload(&p->it_a, ia);
    rc |= upload(&p->it_b, ib);
    rc |= upload(&p->it_c, ic);
    rc |= upload(&p->it_off, off);
    rc |= upload(&p->it_str, str);
    rc |= upload(&p->shell_l, sl);
    rc |= upload(&p->shell_np, snp);
    rc |= upload(&p->shell_off, soff);
    if (rc) { intc_onee_plan_destroy(p); return rc; }
    *plan = p;
    return 0;
}

// Device pointers: exps/coefs (normalised coefficients), shell_cen [3 n_shell],
// centres [3 n_centre] (nuclear attraction; may be NULL for overlap), sph (S_l
// table in the intc.h layout; NULL = the built-in real solid harmonics).
INTC_API int intc_onee_plan_execute(intc_onee_plan *p, const double *exps, const double *coefs,
                                    const double *shell_cen, const double *centres, const double *sph, double *out,
                                    void *stream)
{
    if (p->kind == 1 && p->n > 0 && !centres) return intc_internal_fail(-1, "nuclear attraction needs the centres");
    IntcOneArgs a;
    memset(&a, 0, sizeof(a));
    const double *builtin = nullptr;
    int rc = intc_internal_tables(&a.boys_cheb, &a.boys_small, &a.boys_medium, &builtin);
    if (rc) return rc;
    a.boys_mode = intc_internal_boys_mode();
    a.shell_l = p->shell_l;
    a.shell_np = p->shell_np;
    a.shell_off = p->shell_off;
    a.shell_cen = shell_cen;
    a.exps = exps;
    a.coefs = coefs;
    a.centres = centres;
    a.sph = sph ? sph : builtin;
    a.out = out;
    hipStream_t st = (hipStream_t)stream;
    for (const intc_onee_plan::Bin &b : p->bins) {
        IntcOneArgs ab = a;
        ab.it_a = p->it_a + b.first;
        ab.it_b = p->it_b + b.first;
        ab.it_c = p->it_c + b.first;
        ab.it_off = p->it_off + b.first;
        ab.it_str = p->it_str + b.first;
        ab.n = (int)b.count;
        const int grid = (int)std::min<long long>((b.count + 255) / 256, 256 * 16);
        hipError_t e = b.e->fn(&ab, grid, st);
        if (e != hipSuccess) return hip_fail("one-electron kernel launch", e);
    }
    return 0;
}

// ---------------------------------------------------------------------------
// per-call twins of the generated <full>_main for the one-electron kinds
// ---------------------------------------------------------------------------
namespace {

struct OneShell {
    const double *x, *c;
    int l, n, iskip, n_prim, n_cont, iskip_cont;
    const double *contract, *sph;
};

// One call (every contraction pair) through a plan: shells per (index,
// contraction), items, upload, execute, download, scatter to
// out[ia*iskip_a + ka*iskip_a_cont + ib*iskip_b + kb*iskip_b_cont].
int onee_per_call(int kind, int spherical, const OneShell s[2], const double *C, double *output)
{
    for (int k = 0; k < 2; ++k) {
        if (s[k].l < 0 || s[k].l > intc_one_lmax || s[k].n_prim < 1 || s[k].n_cont < 1)
            return intc_internal_fail(-1, "invalid shell arguments");
        if (s[k].n != (spherical ? nsph(s[k].l) : ncc(s[k].l)))
            return intc_internal_fail(-1, "component count n does not match l");
    }
    std::vector<int> sl, snp, soff;
    std::vector<double> cen, ex, co;
    int base[2];
    for (int k = 0; k < 2; ++k) {
        base[k] = (int)sl.size();
        for (int j = 0; j < s[k].n_cont; ++j) {
            sl.push_back(s[k].l);
            snp.push_back(s[k].n_prim);
            soff.push_back((int)ex.size());
            for (int i = 0; i < 3; ++i) cen.push_back(s[k].c[i]);
            for (int i = 0; i < s[k].n_prim; ++i) {
                ex.push_back(s[k].x[i]);
                co.push_back(s[k].contract ? s[k].contract[i + j * s[k].n_prim] : 1.0);
            }
        }
    }
    const int nblk = s[0].n * s[1].n;
    std::vector<int> items;
    std::vector<long long> off;
    for (int ka = 0; ka < s[0].n_cont; ++ka)
        for (int kb = 0; kb < s[1].n_cont; ++kb) {
            items.push_back(base[0] + ka);
            items.push_back(base[1] + kb);
            if (kind == 1) items.push_back(0);
            off.push_back((long long)off.size() * nblk);
        }
    const long long nitem = (long long)off.size();
    intc_onee_plan *p = nullptr;
    int rc = intc_onee_plan_create(&p, kind, spherical, (int)sl.size(), sl.data(), snp.data(), soff.data(),
                                   cen.data(), ex.data(), co.data(), kind == 1 ? 1 : 0, nitem, items.data(),
                                   off.data());
    if (rc) return rc;
    std::vector<double> sph_all;
    if (spherical) {
        sph_all.assign(intc_sph_off(8), 0.0);
        for (int k = 0; k < 2; ++k)
            for (int i = 0; i < ncc(s[k].l) * nsph(s[k].l); ++i) sph_all[intc_sph_off(s[k].l) + i] = s[k].sph[i];
    }
    std::vector<double> res((size_t)nitem * nblk);
    double *d_ex = nullptr, *d_co = nullptr, *d_cen = nullptr, *d_c = nullptr, *d_sph = nullptr, *d_out = nullptr;
    auto cleanup = [&]() {
        double *ptrs[] = {d_ex, d_co, d_cen, d_c, d_sph, d_out};
        for (double *q : ptrs)
            if (q) (void)hipFree(q);
        intc_onee_plan_destroy(p);
    };
#define PCHK(x)                                                        \
    do {                                                               \
        hipError_t e_ = (x);                                           \
        if (e_ != hipSuccess) { cleanup(); return hip_fail(#x, e_); }  \
    } while (0)
    PCHK(hipMalloc(&d_ex, sizeof(double) * ex.size()));
    PCHK(hipMalloc(&d_co, sizeof(double) * co.size()));
    PCHK(hipMalloc(&d_cen, sizeof(double) * cen.size()));
    PCHK(hipMalloc(&d_out, sizeof(double) * res.size()));
    PCHK(hipMemcpy(d_ex, ex.data(), sizeof(double) * ex.size(), hipMemcpyHostToDevice));
    PCHK(hipMemcpy(d_co, co.data(), sizeof(double) * co.size(), hipMemcpyHostToDevice));
    PCHK(hipMemcpy(d_cen, cen.data(), sizeof(double) * cen.size(), hipMemcpyHostToDevice));
    if (kind == 1) {
        PCHK(hipMalloc(&d_c, sizeof(double) * 3));
        PCHK(hipMemcpy(d_c, C, sizeof(double) * 3, hipMemcpyHostToDevice));
    }
    if (spherical) {
        PCHK(hipMalloc(&d_sph, sizeof(double) * sph_all.size()));
        PCHK(hipMemcpy(d_sph, sph_all.data(), sizeof(double) * sph_all.size(), hipMemcpyHostToDevice));
    }
    rc = intc_onee_plan_execute(p, d_ex, d_co, d_cen, d_c, d_sph, d_out, nullptr);
    if (rc) { cleanup(); return rc; }
    PCHK(hipDeviceSynchronize());
    PCHK(hipMemcpy(res.data(), d_out, sizeof(double) * res.size(), hipMemcpyDeviceToHost));
#undef PCHK
    cleanup();
    long long k = 0;
    for (int ka = 0; ka < s[0].n_cont; ++ka)
        for (int kb = 0; kb < s[1].n_cont; ++kb, ++k) {
            const double *blk = &res[(size_t)k * nblk];
            for (int ia = 0; ia < s[0].n; ++ia)
                for (int ib = 0; ib < s[1].n; ++ib)
                    output[(long)ia * s[0].iskip + (long)ka * s[0].iskip_cont + (long)ib * s[1].iskip +
                           (long)kb * s[1].iskip_cont] = blk[ia * s[1].n + ib];
        }
    return 0;
}

}  // namespace

#define ONE_SHELL_PARAMS(X)                                                                         \
    const double *x##X, const double c##X[3], const int l##X, const int n##X, const int iskip##X,   \
        const int n##X##_prim, const int n##X##_cont, const int iskip##X##_cont,                    \
        const double *contract_array_##X
#define ONE_MK(X, S) {x##X, c##X, l##X, n##X, iskip##X, n##X##_prim, n##X##_cont, iskip##X##_cont, contract_array_##X, S}

// primitive Cartesian (the reference example's contracted=False functions):
// per index exponent, centre, l, n, iskip
INTC_API int intc_overlap_prim_cart_main(const double xa, const double ca[3], const int la, const int na,
                                         const int iskipa, const double xb, const double cb[3], const int lb,
                                         const int nb, const int iskipb, double *work_array, double *output_array)
{
    (void)work_array;
    OneShell s[2] = {{&xa, ca, la, na, iskipa, 1, 1, 0, nullptr, nullptr},
                     {&xb, cb, lb, nb, iskipb, 1, 1, 0, nullptr, nullptr}};
    return onee_per_call(0, 0, s, nullptr, output_array);
}

INTC_API int intc_nuclear_prim_cart_main(const double xa, const double ca[3], const int la, const int na,
                                         const int iskipa, const double xb, const double cb[3], const int lb,
                                         const int nb, const int iskipb, const double cc[3], double *work_array,
                                         double *output_array)
{
    (void)work_array;
    OneShell s[2] = {{&xa, ca, la, na, iskipa, 1, 1, 0, nullptr, nullptr},
                     {&xb, cb, lb, nb, iskipb, 1, 1, 0, nullptr, nullptr}};
    return onee_per_call(1, 0, s, cc, output_array);
}

// contracted: per index exponents, centre, l, n, iskip, n_prim, n_cont,
// iskip_cont, contract_array [, sph_trans_array]
INTC_API int intc_overlap_cont_cart_main(ONE_SHELL_PARAMS(a), ONE_SHELL_PARAMS(b), double *work_array,
                                         double *output_array)
{
    (void)work_array;
    OneShell s[2] = {ONE_MK(a, nullptr), ONE_MK(b, nullptr)};
    return onee_per_call(0, 0, s, nullptr, output_array);
}

INTC_API int intc_overlap_cont_sph_main(ONE_SHELL_PARAMS(a), const double *sph_trans_array_a, ONE_SHELL_PARAMS(b),
                                        const double *sph_trans_array_b, double *work_array, double *output_array)
{
    (void)work_array;
    OneShell s[2] = {ONE_MK(a, sph_trans_array_a), ONE_MK(b, sph_trans_array_b)};
    return onee_per_call(0, 1, s, nullptr, output_array);
}

INTC_API int intc_nuclear_cont_cart_main(ONE_SHELL_PARAMS(a), ONE_SHELL_PARAMS(b), const double cc[3],
                                         double *work_array, double *output_array)
{
    (void)work_array;
    OneShell s[2] = {ONE_MK(a, nullptr), ONE_MK(b, nullptr)};
    return onee_per_call(1, 0, s, cc, output_array);
}

INTC_API int intc_nuclear_cont_sph_main(ONE_SHELL_PARAMS(a), const double *sph_trans_array_a, ONE_SHELL_PARAMS(b),
                                        const double *sph_trans_array_b, const double cc[3], double *work_array,
                                        double *output_array)
{
    (void)work_array;
    OneShell s[2] = {ONE_MK(a, sph_trans_array_a), ONE_MK(b, sph_trans_array_b)};
    return onee_per_call(1, 1, s, cc, output_array);
}
