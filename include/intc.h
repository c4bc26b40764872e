/*
 * intc.h -- C-ABI of libintc.so, the MI355X (gfx950) contracted-ERI library.
 *
 * All functions return 0 on success and a negative status on error
 * (-1 invalid argument, -2 class not built into the library, -3 HIP error);
 * intc_last_error() gives the message.  The reference's generated functions
 * return void and do no checking (src/intception/generator/callables/
 * main_source_algorithms/contracted_general.py:31-73); these return a status
 * instead and otherwise keep the reference's argument layout and meaning.
 *
 * Reference names: "<full>" below stands for the prefix the reference
 * generator gives the ERI class of the input script, e.g.
 * intception_integral_eri_abcd_m_cont_sph (integral name from
 * src/intception/dsl_extensions.py:612-627, suffixes from
 * src/intception/generator/generator.py:432-457).
 */
#ifndef INTC_H
#define INTC_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Library state                                                          */
/* ---------------------------------------------------------------------- */
const char *intc_last_error(void);
int intc_version(void);
/* Uploads the Boys tables to the current device (also done lazily). */
int intc_init(void);
/* 0: Chebyshev interpolation + downward recursion (default).
 * 1: the reference's three-branch scheme
 *    (src/intception/generator/callables/generator_boys_function.py:98-205),
 *    for bit-level parity studies. */
int intc_set_boys_mode(int mode);
/* 1 if the canonical form of (la lb|lc ld) has a kernel in this build. */
int intc_class_supported(int la, int lb, int lc, int ld, int spherical);

/* ---------------------------------------------------------------------- */
/* Per-quartet twins of the generated functions (host pointers)           */
/* ---------------------------------------------------------------------- */
/* Replaces <full>_work_array_size
 * (src/intception/generator/callables/main_source_algorithms/contracted_vrr_hrr_with_auxiliary.py:248-351).
 * Returns the same bound as the reference's formula; the device path itself
 * does not use the caller's work array. */
int intc_eri_cont_sph_work_array_size(int la, int lb, int lc, int ld, int na_prim, int nb_prim, int nc_prim,
                                      int nd_prim, int na_cont, int nb_cont, int nc_cont, int nd_cont);
int intc_eri_cont_cart_work_array_size(int la, int lb, int lc, int ld, int na_prim, int nb_prim, int nc_prim,
                                       int nd_prim, int na_cont, int nb_cont, int nc_cont, int nd_cont);

#define INTC_SHELL_ARGS(X)                                                                            \
    const double *x##X, const double c##X[3], const int l##X, const int n##X, const int iskip##X,     \
        const int n##X##_prim, const int n##X##_cont, const int iskip##X##_cont,                      \
        const double *contract_array_##X

/* Replaces <full>_main (spherical variant): argument order per index is
 * exponents, centre, l, n, iskip, n_prim, n_cont, iskip_cont,
 * contract_array (n_prim x n_cont, column-major), sph_trans_array
 * (ncc(l) x (2l+1), column-major) -- contracted_general.py:31-73 and
 * src/intception/generator/integral_wrapper.py:1892-1926.  Output element
 * out[sum_X(iX*iskipX + icontX*iskipX_cont)] as in contracted_general.py:1808-1875. */
int intc_eri_cont_sph_main(INTC_SHELL_ARGS(a), const double *sph_trans_array_a, INTC_SHELL_ARGS(b),
                           const double *sph_trans_array_b, INTC_SHELL_ARGS(c), const double *sph_trans_array_c,
                           INTC_SHELL_ARGS(d), const double *sph_trans_array_d, double *work_array,
                           double *output_array);
/* No counterpart in the reference.  intc_eri_cont_sph_main for callers that
 * cannot allocate device memory (code under HIP graph capture excluded: it
 * synchronises; tight serving loops, harnesses that forbid hipMalloc): same
 * arguments, same kernels and results, and no hipMalloc/hipFree.  Every
 * device buffer lives in one static 8 MiB device array per GPU whose address
 * intc_init() looks up, so intc_init() must have run on the current device
 * (-1 otherwise).  Limits: n_prim <= 8 and n_cont <= 2 per index, indices
 * with the same l pass the same sph_trans_array, all contraction blocks
 * together <= 1 033 376 doubles (-1 beyond).  Calls that overlap on one GPU
 * share the array and race. */
int intc_eri_cont_sph_main_noalloc(INTC_SHELL_ARGS(a), const double *sph_trans_array_a, INTC_SHELL_ARGS(b),
                                   const double *sph_trans_array_b, INTC_SHELL_ARGS(c),
                                   const double *sph_trans_array_c, INTC_SHELL_ARGS(d),
                                   const double *sph_trans_array_d, double *work_array, double *output_array);
/* Replaces <full>_main (Cartesian variant, no sph_trans_array arguments). */
int intc_eri_cont_cart_main(INTC_SHELL_ARGS(a), INTC_SHELL_ARGS(b), INTC_SHELL_ARGS(c), INTC_SHELL_ARGS(d),
                            double *work_array, double *output_array);

/* ---------------------------------------------------------------------- */
/* Batched twin (no counterpart in the reference, which evaluates one     */
/* quartet per call): the same fields for many quartets, as an SoA shell  */
/* table.  Shell s has angular momentum shell_l[s], shell_nprim[s]        */
/* primitives at exps/coefs[shell_off[s] ...] (n_cont = 1) and centre     */
/* shell_cen[3s..3s+2]; quartet q is shells quartets[4q..4q+3] (a,b,c,d)  */
/* and writes a row-major [a][b][c][d] block at out[out_off[q]].          */
/* ---------------------------------------------------------------------- */
typedef struct intc_plan intc_plan;

/* Host arrays; bins quartets by canonical class and primitive count and
 * uploads the index data.  Device memory is owned by the plan.  shell_cen,
 * exps and coefs (host copies of what execute receives) may be NULL; when
 * given, a pair of two shells with the same l is oriented so that the
 * recurrences build on the shell the primitive-pair centres lie closest to
 * (better conditioned HRR for diffuse/tight pairs far apart). */
int intc_plan_create(intc_plan **plan, int spherical, int n_shell, const int *shell_l, const int *shell_nprim,
                     const int *shell_off, const double *shell_cen, const double *exps, const double *coefs,
                     long long n_quartet, const int *quartets, const long long *out_off);
/* Same, with shell-pair data shared between quartets: pair P is shells
 * (pair_shells[2P], pair_shells[2P+1]); quartet q is (pair quartet_pairs[2q] |
 * pair quartet_pairs[2q+1]) with a row-major block over those four shells in
 * that order at out[out_off[q]]. */
int intc_plan_create_pairs(intc_plan **plan, int spherical, int n_shell, const int *shell_l, const int *shell_nprim,
                           const int *shell_off, const double *shell_cen, const double *exps, const double *coefs,
                           int n_pair, const int *pair_shells, long long n_quartet, const int *quartet_pairs,
                           const long long *out_off);
/* General contraction and caller strides, the per-quartet fields of the
 * reference's <full>_main (n_cont, iskip, iskip_cont per index;
 * contracted_general.py:44-65, 1808-1875) for a batch: shell s has
 * shell_ncont[s] contractions, column k of its coefficients (normalised,
 * n_prim x n_cont column-major, as contract_array) at coefs[coef_off[s] + k *
 * shell_nprim[s]], exponents at exps[shell_off[s]]; quartet q writes element
 * (ia, ib, ic, id) of contraction (ka, kb, kc, kd) at out[out_off[q] +
 * sum_X(iX * iskip[4q+X] + kX * iskip_cont[4q+X])] (iskip in [0, 65535]).
 * execute receives exps, coefs and shell_cen in this layout.  Routed
 * execution (which zeroes dense skipped blocks) is refused on such plans;
 * intc_plan_quartet_order lists the expanded (quartet, contractions) items. */
int intc_plan_create_general(intc_plan **plan, int spherical, int n_shell, const int *shell_l,
                             const int *shell_nprim, const int *shell_ncont, const int *shell_off,
                             const int *coef_off, const double *shell_cen, const double *exps, const double *coefs,
                             long long n_quartet, const int *quartets, const long long *out_off, const int *iskip,
                             const int *iskip_cont);
int intc_plan_num_bins(const intc_plan *plan);
/* prec: 0 = FP64 kernel, 1 = kernel with FP32 recurrences (FP64 Boys,
 * prefactors, contraction sums and contracted HRR/transform). */
int intc_plan_bin_info(const intc_plan *plan, int bin, int prec, int *cls4, int *nq, int *team_groups, int *grid,
                       int *G);
int intc_plan_set_bin_tuning(intc_plan *plan, int bin, int prec, int team_groups, int grid);
/* Device pointers.  sph: S_l for l = 0..7 concatenated, each ncc(l) x (2l+1)
 * column-major; NULL selects the built-in real solid harmonics (compiled into
 * the kernels; ignored by Cartesian plans).  stream: hipStream_t or NULL. */
int intc_plan_build_pairs(intc_plan *plan, const double *exps, const double *coefs, const double *shell_cen,
                          void *stream);
/* One class bin; the pair data must have been built (build_pairs or execute). */
int intc_plan_run_bin(intc_plan *plan, int bin, int prec, const double *sph, double *out, void *stream);
/* All quartets in FP64. */
int intc_plan_execute(intc_plan *plan, const double *exps, const double *coefs, const double *shell_cen,
                      const double *sph, double *out, void *stream);
/* Per-quartet route (device int8 per caller quartet: 0 FP64, 1 FP32
 * recurrences, 2 skip = block set to zero), compacted on the device into
 * per-class FP64/FP32 queues, then both kernels per class. */
int intc_plan_execute_routed(intc_plan *plan, const double *exps, const double *coefs, const double *shell_cen,
                             const double *sph, const signed char *route, double *out, void *stream);
/* One hipGraph holding the pair builder, the route compaction (route != NULL)
 * and every class-bin kernel; with n_streams > 1 the bins are independent
 * nodes that run concurrently.  The graph binds the given device pointers
 * (route is read at replay time); replay with intc_plan_graph_launch. */
int intc_plan_graph_create(intc_plan *plan, const double *exps, const double *coefs, const double *shell_cen,
                           const double *sph, const signed char *route, double *out, int n_streams, void *stream);
int intc_plan_graph_launch(intc_plan *plan, void *stream);
/* Routed runs: 1 (default) launches FP64 and FP32-recurrence queues; 0 treats
 * route code 1 as FP64 and launches the FP64 kernels only (no effect on
 * graphs already captured). */
int intc_plan_set_route_fp32(intc_plan *plan, int use_fp32);
/* FP32 error check of routed runs (no counterpart in the reference, whose
 * generated C is FP64 throughout): with tol > 0 every FP32-recurrence register
 * kernel estimates its quartet's error as 8 * 2^(la+lb+lc+ld) * 2^-24 times
 * the largest contracted (e 0|f 0) intermediate and queues a quartet whose
 * estimate exceeds tol; the bin's FP64 kernel then recomputes the queued
 * quartets on the same stream, overwriting their blocks.  tol = 0 turns it off
 * (the default).  Graphs keep the setting they were captured with.  -1 for a
 * negative or non-finite tol. */
int intc_plan_set_fp32_check(intc_plan *plan, double tol);
/* Per bin: quartets the last routed run recomputed in FP64 after the check. */
int intc_plan_redo_counts(intc_plan *plan, int *counts);
/* Persistent multi-class launch (no counterpart in the reference): with
 * enable = 1, execute, execute_routed and graphs captured afterwards run every
 * bin whose class the multi-class kernel holds (all spherical register
 * classes, FP64 and FP32 recurrences) in ONE launch: waves take tiles of
 * (teams per wave) x (at most tile_iters) quartet positions of one bin from a
 * device counter, bins interleaved, and run that class's code with the bin's
 * team size; a tile closes early at 1/8 of an average wave's estimated share.  Other bins still get their own kernels.  Spherical plans
 * with the built-in transforms (sph = NULL) and Boys mode 0 only (-2 / -1
 * otherwise).  grid: blocks of 256 threads, 0 = two per CU. */
int intc_plan_set_persistent(intc_plan *plan, int enable, int tile_iters, int grid);
/* Bins the persistent launch may hold (no counterpart in the reference):
 * cover[i] != 0 selects bin i (intc_plan_bin_info order); n = 0 clears the
 * selection (every bin whose class the multi-class kernel holds).  With the
 * persistent launch enabled, the selected bins run as tiles of the one launch
 * and the others keep their own class kernels beside it.  -1 unless n is 0 or
 * the number of bins. */
int intc_plan_set_persistent_bins(intc_plan *plan, int n, const int *cover);
/* Tiles and LDS doubles per wave of the last persistent eager execution. */
int intc_plan_persistent_info(const intc_plan *plan, int *n_tiles, int *lds_doubles_per_wave);
/* Primitive-pair screen (no counterpart in the reference, which contracts
 * every primitive; contracted_general.py:1083-1416): builds after this call
 * (execute, execute_routed, graph capture, build_pairs) keep only the
 * primitive pairs whose s-type Schwarz factor w = |c_a c_b| K_AB
 * sqrt(2 pi^(5/2)) / zeta * (2 zeta)^(-1/4) is >= cut; cut = 0 (the default)
 * keeps every pair and leaves results unchanged.  An integral then moves by
 * at most about cut * (sum of w over the other pair) times an angular factor
 * (measured <= 7 for l <= 2).  Graphs captured before the call are unchanged;
 * intc_plan_run_bin needs a build after it.  -1 for a negative or non-finite
 * cut. */
int intc_plan_set_prim_screen(intc_plan *plan, double cut);
/* Primitive pairs per shell pair in plan pair order: all (screened = 0) or
 * those kept by the last screened build (screened = 1). */
int intc_plan_pair_prims(const intc_plan *plan, int screened, int *out);
/* FP64/FP32/skip queue lengths per bin of the last routed execution (3 ints per bin). */
int intc_plan_route_counts(intc_plan *plan, int *counts);
/* Plan order: order[i] = caller's index of the quartet at plan position i
 * (class, then primitive count descending, then caller index).  A caller that
 * lists its quartets in this order gets the identity, and routed executions
 * then read the route array and write skipped blocks sequentially. */
int intc_plan_quartet_order(intc_plan *plan, int *order);
int intc_plan_destroy(intc_plan *plan);

/* ---------------------------------------------------------------------- */
/* One-electron kinds of the input-script API: the 2-index overlap (a|b)  */
/* and nuclear attraction (a|C|b) of the reference's example script       */
/* (examples/example_input.py:79-110), shells up to l = 3.               */
/* ---------------------------------------------------------------------- */
/* Replace the <full>_main the reference generates for that script with
 * generator_options(contracted=False) (primitive, Cartesian): per index the
 * exponent, centre, l, n, iskip (primitive_general.py:31-98, attachment order
 * dsl.py:971-972); the nuclear attraction adds the centre C as the integral's
 * extra argument after the indices.  Output out[ia*iskipa + ib*iskipb]. */
int intc_overlap_prim_cart_main(const double xa, const double ca[3], const int la, const int na, const int iskipa,
                                const double xb, const double cb[3], const int lb, const int nb, const int iskipb,
                                double *work_array, double *output_array);
int intc_nuclear_prim_cart_main(const double xa, const double ca[3], const int la, const int na, const int iskipa,
                                const double xb, const double cb[3], const int lb, const int nb, const int iskipb,
                                const double cc[3], double *work_array, double *output_array);
/* Contracted variants (contracted=True): the per-index argument layout of the
 * ERI twins above (contracted_general.py:31-73) for two indices. */
int intc_overlap_cont_cart_main(INTC_SHELL_ARGS(a), INTC_SHELL_ARGS(b), double *work_array, double *output_array);
int intc_overlap_cont_sph_main(INTC_SHELL_ARGS(a), const double *sph_trans_array_a, INTC_SHELL_ARGS(b),
                               const double *sph_trans_array_b, double *work_array, double *output_array);
int intc_nuclear_cont_cart_main(INTC_SHELL_ARGS(a), INTC_SHELL_ARGS(b), const double cc[3], double *work_array,
                                double *output_array);
int intc_nuclear_cont_sph_main(INTC_SHELL_ARGS(a), const double *sph_trans_array_a, INTC_SHELL_ARGS(b),
                               const double *sph_trans_array_b, const double cc[3], double *work_array,
                               double *output_array);

/* Batched twin (no counterpart in the reference): kind 0 = overlap, items
 * are n_item shell pairs (a, b); kind 1 = nuclear attraction, items are
 * n_item triples (a, b, c), c indexing n_centre point-charge centres.  Item i
 * writes a row-major [a][b] block at out[out_off[i]] (Cartesian or 2l+1
 * spherical components).  Host index arrays; device memory owned by the plan. */
typedef struct intc_onee_plan intc_onee_plan;
/* shell_cen, exps, coefs: host copies (may be NULL); when given, each pair is
 * built on the shell its primitive-pair centres lie closest to and the HRR
 * moves angular momentum to the other (as for the ERI plans). */
int intc_onee_plan_create(intc_onee_plan **plan, int kind, int spherical, int n_shell, const int *shell_l,
                          const int *shell_nprim, const int *shell_off, const double *shell_cen, const double *exps,
                          const double *coefs, int n_centre, long long n_item, const int *items,
                          const long long *out_off);
/* Device pointers: exps, coefs (normalised, one contraction per shell),
 * shell_cen [3 n_shell], centres [3 n_centre] (NULL for overlap), sph (S_l
 * table as above; NULL = built-in real solid harmonics), out. */
int intc_onee_plan_execute(intc_onee_plan *plan, const double *exps, const double *coefs, const double *shell_cen,
                           const double *centres, const double *sph, double *out, void *stream);
int intc_onee_plan_destroy(intc_onee_plan *plan);

#ifdef __cplusplus
}
#endif
#endif
