/*
 * ima_gpu.h -- C ABI of the MI355X engine for the IMa2p hot path.
 *
 * libima_gpu.so (built from ima2p_amd/csrc/) exports exactly the functions
 * declared here.  Every signature uses plain C types: pointers, sizes and
 * enums.  All host buffers are owned by the caller and only read or written
 * during the call; device buffers are owned by the context.
 *
 * Conventions (reference boundary: imamp.hpp:1166-1174, 1269-1280,
 * 1295-1316 declare the functions these replace):
 *   - every entry point returns IMA_OK (0) or an IMA_ERR_* code; there is
 *     no exit().  ima_ctx_last_error() returns a text for the last error;
 *   - per-genealogy outcomes the reference signals with sentinel values
 *     (REJECTINFINITESITESCONSTANT for IS, -DBL_MAX for SW) are reported
 *     both as those sentinel values in loglik[] and as status codes;
 *   - one context = one HIP stream = one host thread at a time.  Contexts
 *     are independent (no globals), so several may share one device.
 *
 * Genealogies use the reference's edge layout (imamp.hpp:651-679): for a
 * locus with n gene copies, edges 0..n-1 are tips and n..2n-2 internal;
 * up0/up1 are the daughter edges (-1 at tips), down the parent edge (-1 at
 * the root edge), time the time at the BOTTOM of the edge.
 */
#ifndef IMA_GPU_H
#define IMA_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IMA_ABI_VERSION 1

enum ima_rc {
  IMA_OK = 0,
  IMA_ERR_ARG = 1,    /* bad argument (index out of range, null pointer) */
  IMA_ERR_HIP = 2,    /* HIP runtime error */
  IMA_ERR_STATE = 3,  /* call made in the wrong state */
  IMA_ERR_SIZE = 4,   /* a size limit of the engine was exceeded */
  IMA_ERR_NODEV = 5   /* no usable device */
};

enum ima_model { IMA_MODEL_HKY = 0, IMA_MODEL_IS = 1, IMA_MODEL_SW = 2 };
enum ima_precision { IMA_PREC_FP64 = 0, IMA_PREC_FP32 = 1 };
enum ima_status {
  IMA_STATUS_OK = 0,
  IMA_STATUS_IS_REJECT = 1, /* two 0/1 nodes at a site: calc_prob_data.cpp:784-787 */
  IMA_STATUS_SW_ZERO = 2,   /* a Bessel term <= 0: calc_prob_data.cpp:876-897 */
  IMA_STATUS_IS_NOSITE = 3, /* a site with no 0/1 node (reference: IM_err exit) */
  IMA_STATUS_PRIOR_BAD = 4  /* population labels / migrations do not fit the model */
};
enum ima_lmode { IMA_LMODE_REF_COMPAT = 0, IMA_LMODE_EXACT = 1 };
/* what ima_lmode_popmig evaluates */
enum ima_popmig_kind { IMA_POPMIG_CALC = 0, IMA_POPMIG_MARGIN = 1 };

#define IMA_REJECT_IS (-1000000000.0) /* REJECTINFINITESITESCONSTANT imamp.hpp:157 */
#define IMA_MAX_GENES 256             /* gene copies per locus in this engine */

typedef struct ima_ctx ima_ctx;

/* One scoring request: the genealogy in `slot` under its locus' model.
 * HKY uses u (G->uvals[0]) and kappa; IS uses u; SW uses u and tr (the
 * time scalar applied to branch lengths). */
typedef struct {
  int32_t slot;
  int32_t flags; /* reserved, 0 */
  double u;
  double kappa;
  double tr;
} ima_item;

/* ---- context ---------------------------------------------------------- */
int ima_device_count(int *count);
/* Free and total device memory in bytes (hipMemGetInfo). */
int ima_device_memory(int device, double *free_bytes, double *total_bytes);
int ima_ctx_create(int device, ima_ctx **out);
void ima_ctx_destroy(ima_ctx *ctx);
const char *ima_ctx_last_error(const ima_ctx *ctx);
int ima_abi_version(void);

/* ---- locus data (readata.cpp:246-588 produce these encodings) ---------- */
/* HKY: seq[n][S] bases 0..3 after gap removal and pattern compression,
 * mult[S] pattern counts, totsites = sites after gap removal, pi[4]
 * empirical base frequencies (L[li].seq/mult/totsites, pi[li]). */
int ima_locus_add_hky(ima_ctx *ctx, int n, int S, const int8_t *seq,
                      const int32_t *mult, int totsites, const double pi[4],
                      int *locus);
/* IS: seq[n][S] 0/1 relative to gene copy 0 (readseqIS). */
int ima_locus_add_is(ima_ctx *ctx, int n, int S, const int8_t *seq,
                     int *locus);
/* SW: allele states travel with each genealogy (ima_slots_write A[]). */
int ima_locus_add_sw(ima_ctx *ctx, int n, int *locus);
/* IS constant sum_v log(k_v!) (calc_prob_data.cpp:8,752-765).  The
 * reference computes it once per locus from the first genealogy scored;
 * ima_locus_sumlogk_from_slot does the same from a stored genealogy. */
int ima_locus_set_sumlogk(ima_ctx *ctx, int locus, double sumlogk);
int ima_locus_sumlogk_from_slot(ima_ctx *ctx, int locus, int slot,
                                double *sumlogk);

/* ---- genealogy slots --------------------------------------------------- */
/* Reserve `count` slots, slot i holding a genealogy of locus locus_of[i].
 * Slot ids are returned consecutively from *first. */
int ima_slots_alloc(ima_ctx *ctx, int count, const int32_t *locus_of,
                    int *first);
/* Write genealogies into slots first..first+count-1.  Arrays are packed
 * slot after slot, 2n-1 entries per slot (n of that slot's locus).
 * A may be NULL; for SW loci it holds the allele state of every edge
 * (gtree[i].A[0]). */
int ima_slots_write(ima_ctx *ctx, int first, int count, const int32_t *up0,
                    const int32_t *up1, const int32_t *down,
                    const double *time, const int32_t *root,
                    const int32_t *A);
int ima_slots_read(ima_ctx *ctx, int first, int count, int32_t *up0,
                   int32_t *up1, int32_t *down, double *time, int32_t *root,
                   int32_t *A);
/* Population labels (gtree[i].pop) of slots first..first+count-1 as the
 * device holds them, packed like ima_slots_read's arrays (the device's own
 * genealogy move, IMA_MC3_TOPOLOGY, relabels moved nodes). */
int ima_slots_read_pops(ima_ctx *ctx, int first, int count, int32_t *pop);
/* Number of slots in the context, and the locus / gene-copy count of one
 * slot (slots may also be created by ima_likelihood_hky). */
int ima_slots_count(ima_ctx *ctx, int *count);
int ima_slot_info(ima_ctx *ctx, int slot, int *locus, int *n);

/* ---- batched data likelihood ------------------------------------------ */
/* Scores every item in one launch per mutation model.  Replaces the
 * per-(chain, locus) calls likelihoodHKY (calc_prob_data.cpp:584),
 * likelihoodIS (:732) and likelihoodSW (:842).  loglik[i] and status[i]
 * (may be NULL) receive each item's result. */
int ima_score(ima_ctx *ctx, const ima_item *items, int nitems, int precision,
              double *loglik, int32_t *status);
/* ima_score run `reps` times back to back on the context's stream, every
 * pass writing the same results.  With timing enabled one event pair
 * brackets all passes of a kernel class, so ima_timing_get's average is
 * the kernel's own duration rather than one launch's round trip.  A
 * measurement aid with no counterpart in the reference. */
int ima_score_repeat(ima_ctx *ctx, const ima_item *items, int nitems,
                     int precision, int reps, double *loglik, int32_t *status);
/* IS diagnostics: the 0/1 node of every site (calc_prob_data.cpp:777-821)
 * for the genealogy in `slot`; site_node has S entries. */
int ima_is_site_nodes(ima_ctx *ctx, int slot, int32_t *site_node);
/* SW per-edge terms dlikeA (calc_prob_data.cpp:879-889) for one item. */
int ima_sw_dlike(ima_ctx *ctx, const ima_item *item, double *dlikeA);

/* The allele proposal of a genealogy move at a stepwise locus,
 * finishSWupdateA (update_gtree_common.cpp:2219-2362).  The host moved
 * `edge` onto `newsisedge` through the freed node `downedge` (old sister
 * `sisedge`) and wrote the moved genealogy to `slot`; the copy fields are
 * what storeAinfo (:820-846) saved before the move, uni and flip the draws
 * of geometric() (utilities.cpp:609-617, by inversion) and bitran(). */
typedef struct {
  int32_t slot, edge, downedge, sisedge, newsisedge;
  int32_t copyA0, copyA1;  /* A of edge and sisedge before the move */
  int32_t down_was_root;   /* downedge was the root edge before the move */
  int32_t hold_down_A;     /* A of downedge's old parent (holddownA) */
  int32_t minA, maxA;      /* the locus' allele range (L[li].minA/maxA) */
  int32_t flip;            /* bitran(): the step goes down */
  double copy_dlikeA[3];   /* dlikeA of edge, sisedge, downedge before */
  double newsis_dlikeA;    /* dlikeA of newsisedge before the move */
  double u;                /* mutation scalar of the locus */
  double uni;              /* uniform() draw of geometric(), in (0, 1] */
} ima_sw_move;
/* All moves in one launch, one lane each (no two on the same slot).  The
 * proposed allele goes into the slot (A[downedge]); per move: the
 * log-likelihood change dlike[i] (0 when calc_likelihood is 0, as with
 * DONTCALCLIKELIHOODMUTATION), the Hastings term Aterm[i] (0 likewise),
 * the allele newA[i], and edge_dlikeA[4i..4i+3], the new terms of sisedge,
 * edge, newsisedge and downedge (0 for the root edge).  A move whose slot
 * does not hold edge and newsisedge under downedge fails the call with
 * IMA_ERR_ARG and changes nothing. */
int ima_sw_update_A(ima_ctx *ctx, const ima_sw_move *moves, int nmoves, int calc_likelihood,
                    double *dlike, double *Aterm, int32_t *newA, double *edge_dlikeA);

/* ---- reference-shaped single call (plain per-locus FP64 path) --------- */
/* likelihoodHKY(ci, li, mutrate, kappa, e1, e2, e3, e4) for one genealogy
 * given as host arrays: uploads it, launches one FP64 kernel for this
 * locus and returns the log-likelihood.  e1..e4 name the edges the
 * reference would recompute; the engine recomputes the whole genealogy on
 * chip, which gives the same value (the reference's committed partials are
 * those of the same genealogy). */
int ima_likelihood_hky(ima_ctx *ctx, int locus, const int32_t *up0,
                       const int32_t *up1, const int32_t *down,
                       const double *time, int root, double u, double kappa,
                       int e1, int e2, int e3, int e4, double *loglik);

/* ---- MC3 swap (swapchains.cpp:527-662, swapweight :13-34) -------------- */
/* Performs `swaptries` swap attempts among nchains chains whose total
 * log posterior sums are S[c] = sum_li pdg + probg.  beta[] is updated in
 * place (betas are swapped, never chain state).  Randomness comes from a
 * counter-based generator keyed by seed, starting at block `counter`; a
 * call reads one block of four 32-bit words per two uniforms (three or
 * more uniforms per try), so independent calls need counters that many
 * blocks apart, e.g. call index << 16.  swaps_out (may be
 * NULL) receives the number of accepted swaps; pairs_out (may be NULL,
 * 3*swaptries ints) receives (ci, cj, accepted) per try. */
int ima_swap(ima_ctx *ctx, int nchains, const double *S, double *beta,
             int swaptries, uint64_t seed, uint64_t counter, int *swaps_out,
             int32_t *pairs_out);

/* ---- L mode (jointfind.cpp:885-1047, surface_call_functions.cpp:120-173) */
/* Upload a genealogy table in the savegsampinf layout (ginfo.cpp:288-316):
 * G rows of K floats; npsize population-size and nmig migration
 * parameters, npops sampled populations. */
int ima_lmode_table(ima_ctx *ctx, const float *table, int G, int K,
                    int npsize, int nmig, int npops, int *table_id);
int ima_lmode_table_free(ima_ctx *ctx, int table_id);
/* An empty resident table for `capacity` genealogies of the savegsampinf
 * layout with npsize / nmig parameters and npops - 1 split times, filled
 * on the device by ima_mc3_record (the reference's gsampinf array sized
 * by MAXGENEALOGIESTOSAVE, imamp.hpp:149).  L-mode calls use the rows
 * recorded so far. */
int ima_lmode_table_reserve(ima_ctx *ctx, int capacity, int npsize, int nmig, int npops,
                            int *table_id);
/* Genealogies a table holds, values per genealogy, and its population-size
 * and migration parameter counts and sampled populations (any may be
 * NULL). */
int ima_lmode_table_info(ima_ctx *ctx, int table_id, int *G, int *K, int *npsize, int *nmig,
                         int *npops);
/* Rows [first, first + count) of a table, K floats each. */
int ima_lmode_table_read(ima_ctx *ctx, int table_id, int first, int count, float *rows);
/* jointp at npoints points x[npoints][npsize+nmig]; params
 * [prange0, prange1) are used; modeltype 0 all, 1 popsize-only, 2
 * migration-only (npops > 2).  mode REF_COMPAT reproduces the reference's
 * 10-nat window, list cap and eexp summation; EXACT is a plain FP64
 * log-sum-exp over every genealogy.  out[i] = -log posterior (reference
 * sign convention); ess (may be NULL) receives the effective sample
 * number (calc_ess). */
int ima_lmode_jointp(ima_ctx *ctx, int table_id, const double *x, int npoints,
                     int modeltype, int prange0, int prange1, int mode,
                     double *out, double *ess);
/* ima_lmode_jointp over DEVICE pointers that allocates nothing.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid it).  x (npoints x (npsize + nmig), values > 0), out and ess
 * (may be NULL) are device buffers the caller owns; the per (point, chunk)
 * work area is a static 16 MiB array of the library, one per device, held
 * until the process ends.  Calls that need more work area than that fail
 * with IMA_ERR_SIZE (split the points).  The launches are enqueued on the
 * context's stream and the call returns without waiting.  Every context
 * on a device shares that device's work area: calls that overlap on one
 * device race. */
int ima_lmode_jointp_noalloc(ima_ctx *ctx, int table_id, const double *x, int npoints,
                             int modeltype, int prange0, int prange1, int mode, double *out,
                             double *ess);
/* The joint posterior peak of a table by differential evolution, as
 * findjointpeaks / modelloop search it (jointfind.cpp:596-812, 1089-1183),
 * every generation's trials scored in one batched jointp launch.
 * modeltype 0: all parameters (two populations); 1: population sizes only;
 * 2: migration rates only (more populations).  xmap (may be NULL for the
 * full model) is a nested model's map over the npsize + nmig parameters:
 * xmap[i] < 0 fixes parameter i at -xmap[i], otherwise it takes varying
 * value xmap[i] (mapvals, :563-573; positions index the full vector, so a
 * migration-only model's values start at npsize).  lower/upper: every
 * parameter's bounds (setbounds :576-594 uses MINPARAMVAL and the prior
 * maxima).  popsize_multiplier: population per varying value (the
 * reference's 100).  On return peak[npsize + nmig] holds the peak's full
 * parameter vector, *logp the log posterior density there (the reference
 * prints it as log(P); 2LLR of a nested model is 2 (logp_full -
 * logp_nested)), *ess the effective sample number, *evaluations the jointp
 * points evaluated (any output but peak may be NULL).  The random numbers
 * are MT19937's from `seed`. */
int ima_lmode_joint_peak(ima_ctx *ctx, int table_id, int modeltype, const double *xmap,
                         const double *lower, const double *upper, int popsize_multiplier,
                         uint64_t seed, int mode, double *peak, double *logp, double *ess,
                         int64_t *evaluations);
/* The marginal posterior peak of parameter `param` over [lower, upper]
 * (marginalopt / findmarginpeaks, surface_call_functions.cpp:175-470): the
 * mean density of margincalc over the table, maximised (one batched launch
 * over a 1024-point grid brackets it, golden section refines it to 1e-7
 * relative, as goldenmod), with the 95% limits where the density falls
 * e^1.92 below the peak's (margin95).  out[5] = peak location, log density
 * there, lower and upper 95% limits (-1 where the density does not fall
 * that low inside the range), and for a migration rate the no-migration
 * statistic 2 log(p(peak) / p(MINPARAMVAL)) (0 otherwise).  expo /
 * prior_mean as for ima_lmode_margincalc. */
int ima_lmode_marginal_peak(ima_ctx *ctx, int table_id, int param, double lower, double upper,
                            int expo, double prior_mean, double *out);
/* The posterior density of the population migration rate 2NM = theta_i *
 * m_i / 2 at nx points x (popmig.cpp:9-351), replacing calc_popmig /
 * calc_pop_expomig and marginpopmig / marginpop_expomig: per genealogy the
 * closed-form integral of the coalescent and migration likelihoods along
 * theta m / 2 = x over theta_i's uniform prior (0, qmax) and m_i's prior
 * (uniform on (0, mprior), or exponential with mean mprior when expo != 0),
 * averaged over the table's genealogies.
 *   IMA_POPMIG_CALC: calc_popmig / calc_pop_expomig over the whole table
 *     (first, last ignored), divided by the prior density of 2NM when
 *     prob_or_like != 0 (x > 0);
 *   IMA_POPMIG_MARGIN: marginpopmig / marginpop_expomig: minus the sum over
 *     genealogies [first, last) divided by last - first + (first == 0), as
 *     the reference divides, and 1.0 for x outside [0, qmax mprior / 2]
 *     (uniform) or [0, 20 mprior] (exponential).
 * The genealogy columns read are theta_i's cc/fc/hcc/qintegrate and m_i's
 * mc/fm/mintegrate.  Integer counts are truncated from the float table as
 * the reference casts them. */
int ima_lmode_popmig(ima_ctx *ctx, int table_id, int theta_i, int mig_i, double qmax,
                     double mprior, int expo, int kind, int prob_or_like, int first, int last,
                     const double *x, int nx, double *out);
/* The peak of the 2NM density for one (theta_i, m_i) pair over genealogies
 * [first, last) (marginalopt_popmig, popmig.cpp:361-429): the
 * IMA_POPMIG_MARGIN value at 100 bins MINPARAMVAL + j upper / 100 (one
 * launch) brackets the minimum, golden section refines it (goldenmod,
 * surface_search_functions.cpp:134-184, tolerance 1e-7).  *peak is the
 * location, *mlval the density there (minus the minimum); evaluations
 * (may be NULL) counts the points evaluated. */
int ima_lmode_popmig_peak(ima_ctx *ctx, int table_id, int theta_i, int mig_i, double qmax,
                          double mprior, int expo, int first, int last, double *peak,
                          double *mlval, int64_t *evaluations);
/* The parameter comparisons of print_greater_than_tests (gtint.cpp:
 * 333-445): out[i * n + j] = P(parameter i > parameter j) for the n
 * population sizes (kind 0, gtpops) or migration rates (kind 1, gtmig) of
 * the table, averaged over its genealogies (every (G / 20000)-th of the
 * first (G / 20000) * 20000 when G > 20000, USETREESMAX).  prior_max[k] is
 * parameter k's prior upper bound; pairs whose bounds differ, the diagonal
 * and (kind 1) rates with a bound <= MINPARAMVAL get -1, which the
 * reference prints as "na".  Per genealogy the reference's closed forms or
 * its trapezoid refinement (qtrap, relative change below 1e-4) are used,
 * capped at 1. */
int ima_lmode_greater_than(ima_ctx *ctx, int table_id, int kind, int n, const double *prior_max,
                           double *out);
/* margincalc for parameter `param` at nx values x[] (logi, yadjust,
 * exponential migration prior with mean prior_mean when expo != 0). */
int ima_lmode_margincalc(ima_ctx *ctx, int table_id, int param,
                         const double *x, int nx, int logi, double yadjust,
                         int expo, double prior_mean, double *out);

/* ---- genealogy-table files (.ti), host only --------------------------- */
/* Writes G rows of K floats in the reference's .ti layout: with append = 0
 * a new file with the header block of ima_main_mpi.cpp:2121-2142 around
 * `header` (the reference puts its run's parameter text there) and the
 * line VALUESSTART; then each row as K values "%.6f\t" and a newline
 * (savegenealogyfile, output.cpp:662-685).  append = 1 adds rows only. */
int ima_ti_write(const char *path, const char *header, const float *rows, int G, int K,
                 int append);
/* Genealogies (lines after VALUESSTART) in a .ti file. */
int ima_ti_count(const char *path, int *G);
/* loadgenealogyvalues (ima_main_mpi.cpp:3216-3440) over npaths files in
 * the given order: every line after VALUESSTART is one genealogy of
 * exactly K values read as "%f" (fewer or more is IMA_ERR_ARG); with more
 * than max_rows genealogies in all, a genealogy is kept while
 * loaded / (loaded + skipped) <= max_rows / total, the first always.
 * rows holds max_rows * K floats; *loaded and *total (may be NULL) receive
 * the counts. */
int ima_ti_load(const char *const *paths, int npaths, int K, int max_rows, float *rows,
                int *loaded, int *total);

/* ---- MCMC state files (.mcf), host only -------------------------------- */
/* One locus of a .mcf file: its model, gene copies (2n - 1 edges) and, for
 * HKY, its site patterns (the conditional-likelihood lines of every
 * internal edge). */
typedef struct {
  int32_t model;
  int32_t n;
  int32_t nsites;
} ima_mcf_locus;
/* The chains' state, packed chain after chain, and within a chain locus
 * after locus: tvals [chain][nsplit]; u, kappa [chain][locus] (kappa used
 * for HKY loci, may be NULL); pi [chain][locus][4]; per edge
 * [chain][locus][2n - 1]: up0, up1, down, pop, time, A and dlikeA
 * (stepwise loci; may be NULL), mut (may be NULL); mig_count per edge and,
 * in the same order, each edge's migration events in time order in
 * mig_time / mig_to (mig_capacity entries when reading). */
typedef struct {
  double *tvals, *u, *kappa, *pi, *time, *dlikeA, *mig_time;
  int32_t *up0, *up1, *down, *mut, *pop, *A, *mig_count, *mig_to;
  int64_t mig_capacity;
} ima_mcf_state;
/* writemcf (mcmcfile.cpp:204-292): the reference's text layout, item by
 * item as awrite prints it ("<name> <type> <count> " and the values, "%d "
 * or "%.10lg ", one item a line).  The conditional likelihoods of HKY
 * internal edges are written as zeros (readmcf's init_p recomputes them). */
int ima_mcf_write(const char *path, int nchains, int nsplit, int nloci,
                  const ima_mcf_locus *loci, const ima_mcf_state *st);
/* readmcf (mcmcfile.cpp:310-442): every item's name and type checked
 * (IMA_ERR_ARG on a mismatch); a file that ends before nchains chains is
 * read again from its start; with tmax (nsplit bounds, may be NULL) a
 * chain with a split time above its bound is read over by the next one.
 * More migration events than mig_capacity is IMA_ERR_SIZE.  *nmig (may
 * be NULL) receives the events stored. */
int ima_mcf_read(const char *path, int nchains, int nsplit, int nloci,
                 const ima_mcf_locus *loci, const double *tmax, ima_mcf_state *st,
                 int64_t *nmig);

/* ---- genealogy prior (update_gtree_common.cpp:1680-2053) ----------------- */
/* Population model, flattened (the reference's poptree and parameter maps,
 * build_poptree.cpp, setup_iparams initialize.cpp:202-760).  P = npops
 * sampled populations = periods k = 0..P-1 between split times; tree
 * populations 0..2P-2, ancestral population P+k formed by the split that
 * closes period k.  plist[k][i] (P x P, -1 padded) lists the npresent[k]
 * populations of period k; popdown[q] / pope[q] give the population q
 * merges into and the period where it ends (-1 for the root population);
 * join[k] = (a, b, P+k) for k < P-1.  Population-size parameter t sums the
 * (period, position) pairs theta_wp[theta_off[t] .. theta_off[t+1]) and
 * migration parameter m the (period, from, to) triples
 * mig_wp[mig_off[m] .. mig_off[m+1]) (the itheta/imig wp lists); uniform
 * priors on [0, qmax[t]] and [0, mmax[m]]. */
int ima_model_add(ima_ctx *ctx, int npops, const int32_t *plist,
                  const int32_t *npresent, const int32_t *popdown,
                  const int32_t *pope, const int32_t *join, int npsize,
                  const int32_t *theta_off, const int32_t *theta_wp, int nmig,
                  const int32_t *mig_off, const int32_t *mig_wp,
                  const double *qmax, const double *mmax, int *model);
/* The locus' model, gene copies per sampled population (L[li].samppop) and
 * inheritance scalar h (L[li].hval). */
int ima_locus_set_model(ima_ctx *ctx, int locus, int model,
                        const int32_t *samppop, double hval);
/* Population of every edge at its lower end (gtree[i].pop) and the
 * migration events of each genealogy (gtree[i].mig, sorted by edge then
 * time; mig_to is the population the lineage enters going back in time),
 * for slots first..first+count-1, packed slot after slot. */
int ima_slots_write_pops(ima_ctx *ctx, int first, int count,
                         const int32_t *pop, const int32_t *nmig,
                         const int32_t *mig_edge, const double *mig_time,
                         const int32_t *mig_to);
/* treeweight (update_gtree_common.cpp:1680-1931) for each slot, split
 * times tvals[i][0..P-2]; outputs per slot cc[P][P], fc, hcc (P x P),
 * mc[P][P][P], fm (P x P x P), indexed [period][position][position], and
 * the total branch length.  Any output may be NULL. */
int ima_treeweight(ima_ctx *ctx, const int32_t *slots, int nslots,
                   const double *tvals, int32_t *cc, double *fc, double *hcc,
                   int32_t *mc, double *fm, double *length);
/* integrate_tree_prob (update_gtree_common.cpp:1945-2053) for groups of
 * slots (a chain's loci: group g is slots[group_off[g] .. group_off[g+1]),
 * split times tvals[g][0..P-2]): the groups' summed weights integrated over
 * the uniform priors.  probg[g], qint[g][npsize] and mint[g][nmig] (may be
 * NULL) are the reference's probg, qintegrate and mintegrate. */
int ima_prior(ima_ctx *ctx, const int32_t *slots, const int32_t *group_off,
              int ngroups, const double *tvals, double *probg, double *qint,
              double *mint);
/* Room for at least cap migration events in each of slots first ..
 * first + count - 1 (the reference grows each edge's mig array on demand,
 * checkmig, update_gtree_common.cpp; a slot here holds a fixed region).
 * Events already written keep their values. */
int ima_slots_reserve_migs(ima_ctx *ctx, int first, int count, int cap);
/* changet_NW's migration update (update_mig_tNW, update_t_NW.cpp:392-792,
 * with addmigration_NW :238-333, simmpath_tNW :144-236 and getmprob_NW
 * :335-372) of the genealogies in slots[0..nslots) -- a chain's loci --
 * for the move of split `period` from tvals[period] to newt (strictly
 * between its neighbours, 0 and TIMEMAX at the ends): every edge that
 * overlaps the interval between the two times gets its migrations there
 * drawn again so that its populations fit newt.  Node times do not move.
 * The slots' populations and migrations are rewritten in place; updated[i]
 * is 1 for a changed genealogy, 0 for one that lies below the interval
 * (the reference leaves it as it is), 2 for one whose new migrations would
 * exceed its slot's room or the genealogy prior's event limit (left as it
 * is; the caller rejects the move).  migweight[i] receives genealogy i's
 * Hastings term (the reference's migweight sums them).  Genealogy i draws
 * from Philox streams keyed by seed, i and the edge, at counter `counter`.
 * The caller then scores the new genealogies with ima_treeweight at the
 * new split times and decides with exp(beta dprobg + migweight)
 * (changet_NW :983-1000), restoring its own copy on rejection. */
int ima_t_nw_update(ima_ctx *ctx, const int32_t *slots, int nslots, const double *tvals,
                    int period, double newt, uint64_t seed, uint64_t counter,
                    double *migweight, int32_t *updated);
/* integrate_tree_prob (update_gtree_common.cpp:1945-2053) with the
 * reference's own argument: a chain's all-loci weights allgweight, which
 * IMa2p keeps up to date one locus at a time (sum_subtract_treeinfo,
 * ginfo.cpp:249-285).  wsum[g] is group g's record of R = 3P^2 + 2P^3 + 1
 * doubles: cc[P][P] | fc[P][P] | hcc[P][P] | mc[P][P][P] | fm[P][P][P] |
 * length (ima_treeweight's outputs, summed), counts as whole numbers.
 * probg[g], qint[g][npsize], mint[g][nmig] as ima_prior's. */
int ima_integrate_tree_prob(ima_ctx *ctx, int model, const double *wsum, int ngroups,
                            double *probg, double *qint, double *mint);

/* ---- device-resident MC3 sweep ------------------------------------------ */
/* All heated chains on this device: slots first_slot + c*nloci + l hold the
 * genealogy of chain c at locus l.  A sweep proposes one genealogy change
 * per (chain, locus), scores all of them in one batched launch per model
 * and accepts each with the reference's Metropolis rule
 * (update_gtree.cpp:919-927); with IMA_MC3_SCALE, alternate sweeps also
 * rescale each chain's genealogies and re-score every locus before one
 * accept decision per chain (the full re-scoring of changet_RY1,
 * update_t_RY.cpp:343-388); every sweep ends with `swaptries` swap
 * attempts (swapchains.cpp:527-662).  u/kappa are per (chain, locus);
 * scale_width is the width of the log rescaling step. */
#define IMA_MC3_GRAPH 1 /* replay each sweep as a captured hipGraph */
#define IMA_MC3_SCALE 2 /* whole-chain rescaling on alternate sweeps */
#define IMA_MC3_TIMED 4 /* eager launches with HIP events per scoring pass */
/* With IMA_MC3_SCALE and the genealogy prior (two or more populations), the
 * alternate-sweep time update is the reference's rubber-band update of one
 * split time (changet_RY1, update_t_RY.cpp:223-517): split k drawn
 * uniformly, its new time in getnewt's reflected window
 * (update_gtree_common.cpp:2501-2518), node and migration times between
 * the neighbouring split times stretched by beforesplit / aftersplit
 * (update_t_RY.cpp:59-82), Hastings term (ecd + emd) log t_d_hterm +
 * (ecu + emu) log t_u_hterm, every locus re-scored.  Without this flag it
 * is a global rescale of every time by exp(w (U - 1/2)). */
#define IMA_MC3_RY1 8
/* Mutation rates and kappa (changeu, update_mc_params.cpp:24-382, as qupdate
 * calls it every fifth step, ima_main_mpi.cpp:1917-1945): the sweep whose
 * counter q has q % 5 == 4 is followed by a round in which every chain, for
 * j = 0 .. nloci-1 (only j = 0 with two loci), picks k != j uniformly,
 * moves log(u_j/u_k) by a uniform window of log(10000)/nloci reflected at
 * +-3 log(10000), sets u_j *= d, u_k /= d (d = exp(half the move)), moves
 * each HKY locus' kappa (window 2, reflected at 0 and 100), re-scores both
 * loci and accepts with U < min(1, exp(beta sum(new - old))).  A lone HKY
 * locus gets changekappa (:387-429) instead.  Steps touching disjoint loci
 * are scored together; the result is the sequential loop's. */
#define IMA_MC3_UPDATE_U 16
/* With IMA_MC3_UPDATE_U: one step per scoring launch, in the loop's order
 * (the same result, for checking the grouped schedule). */
#define IMA_MC3_U_SERIAL 32
/* The genealogy update is the reference's own move (updategenealogy,
 * update_gtree.cpp:723-966): an edge's top slides along the genealogy
 * (slider_nomigration :79-253) through joinsisdown / splitsisdown
 * (:375-528), so topologies change; the root move's Hastings term
 * (slideweight) and, at stepwise loci, the moved node's allele proposal
 * with its Hastings term (finishSWupdateA, update_gtree_common.cpp:
 * 2219-2362) enter the decision.  Proposals are made on the device in
 * proposal slots the engine allocates on first use, scored in one batched
 * launch and decided in each chain's locus order; accepted ones replace
 * the chains' genealogies.  For a model with migration parameters the
 * move is slider (:256-372) with addmigration (:589-665): the edge's and,
 * when it ends at the root, the new sister's migration paths are drawn
 * again period by period (mwork_single_edge, mwork_two_edges, simmpath,
 * update_gtree_common.cpp:330-1583) and getmprob's forward and reverse
 * path probabilities enter the Hastings term; the engine gives each
 * chain's slot and its proposal slot room for twice the chain's
 * migrations plus 32 on first use, and a proposal beyond that room or
 * the prior's event bound is rejected (IMA_ERR_SIZE when the room does
 * not fit LDS). */
#define IMA_MC3_TOPOLOGY 64
/* With IMA_MC3_SCALE and the genealogy prior of a model with migration:
 * each chain's time update is changet_NW (update_t_NW.cpp:920-1032) with
 * probability 1/2, as qupdate picks between RY1 and NW (ima_main_mpi.cpp:
 * 1867-1911), else the RY1 or rescale update above.  NW moves one split
 * time as RY1 does (getnewt), leaves every node where it is, draws the
 * migrations of each genealogy's edges that overlap the interval between
 * the old and new times again (ima_t_nw_update, on the device), scores
 * the genealogies' weights at the new time and accepts with U < min(1,
 * exp(beta dprobg + migweight)).  On first use every chain slot gets room
 * for twice its migrations plus 32; a genealogy whose new migrations
 * would not fit rejects its chain's move.  Without migration parameters
 * (or without the prior) the flag changes nothing: the reference draws NW
 * only for models with migration (ima_main_mpi.cpp:1808-1814). */
#define IMA_MC3_NW 128
int ima_mc3_setup(ima_ctx *ctx, int nchains, int nloci, int first_slot,
                  const double *u, const double *kappa, const double *beta,
                  uint64_t seed, int precision, int swaptries,
                  double scale_width);
int ima_mc3_run(ima_ctx *ctx, int nsweeps, int flags);
/* pdg[nchains*nloci], beta[nchains]; stats[8] = genealogy proposals,
 * accepts, rescale proposals, accepts, swap tries, swaps, mutation-rate
 * steps, accepts; sweeps done. */
int ima_mc3_read(ima_ctx *ctx, double *pdg, double *beta, uint64_t *stats,
                 uint64_t *sweeps);
/* The chains' current mutation-rate scalars and kappa, u[nchains*nloci] and
 * kappa[nchains*nloci] (either may be NULL); the device copy of
 * C[ci]->G[li].uvals[0] and kappaval (imamp.hpp:967-968). */
int ima_mc3_read_rates(ima_ctx *ctx, double *u, double *kappa);

/* Adds the genealogy prior to the sweep: every MC3 slot's locus has a
 * population model (ima_locus_set_model) and its population labels and
 * migrations (ima_slots_write_pops); tvals[c][0..P-2] are chain c's split
 * times, uniform prior on [0, tmax].  The genealogy proposal then keeps
 * each node inside its period and between its edges' migrations, the
 * rescaling also scales migration and split times, and acceptance is
 * beta (dprobg + dpdg) with probg from integrate_tree_prob over the
 * chain's all-loci weights (update_gtree.cpp:827-961).  Call after
 * ima_mc3_setup. */
int ima_mc3_set_prior(ima_ctx *ctx, const double *tvals, double tmax);
/* Loci decided per group by the sweep's accept step (1..6; set_prior picks
 * the largest that fits LDS).  A group of D loci evaluates each
 * parameter's integral under all 2^D - 1 (locus, earlier decisions)
 * scenarios at once; decisions are the same for every D.  No counterpart
 * in the reference, which decides one locus at a time. */
int ima_mc3_set_prior_group(ima_ctx *ctx, int loci);
/* Several uncoupled analyses in one batched sweep: chains [k*n, (k+1)*n)
 * form analysis k (n = chains_per_analysis, which must divide the chain
 * count; ima_mc3_setup starts with one analysis of all chains).  Each
 * analysis swaps only among its own chains, drawing from a swap stream of
 * its own (analysis 0 keeps the stream of a single analysis, so n = C
 * changes nothing); everything else of a sweep is per chain already.  The
 * caller gives each analysis its own heating ladder through beta.  No
 * counterpart in the reference, where one run is one analysis: it is for
 * callers that run many small analyses (replicates, served requests) and
 * want them in the same launches. */
int ima_mc3_set_analyses(ima_ctx *ctx, int chains_per_analysis);
/* Host-only planning, no device needed: the accept group size set_prior
 * would choose for a population model with these sizes (ima_model_add's
 * npops, npsize, nmig and the lengths of theta_wp / mig_wp in entries) and
 * the dynamic LDS the sweep's per-chain kernels then take.  Every launch
 * stays within 64 KiB of dynamic LDS; IMA_ERR_SIZE (with *loci = 0) when
 * even one locus per group does not fit. */
int ima_mc3_prior_plan(int npops, int npsize, int ntheta_wp, int nmig, int nmig_wp,
                       int *loci, int64_t *lds_bytes);
int ima_mc3_prior_group(ima_ctx *ctx, int *loci);
/* Autotuner of the sweep: times `sweeps` captured sweeps for every group
 * size that fits, each from the same saved state, keeps the fastest
 * (*best; 0 without the prior) and restores the chains -- genealogies,
 * likelihoods, weights, split times, rates and counters -- so the sweeps
 * that follow are those of an untuned run.  `flags` are the ima_mc3_run
 * flags the caller will use (IMA_MC3_GRAPH is implied, IMA_MC3_TIMED and
 * IMA_MC3_U_SERIAL are ignored): the group size is timed on those sweeps.
 * ms_per_sweep (may be NULL, 6 entries) receives each size's time, -1
 * where it was not tried.  At the chosen size it then times the HKY site
 * blocks of 64 and 128 patterns the same way and keeps the faster, unless
 * ima_set_site_block fixed the size (ima_mc3_site_block reads it). */
int ima_mc3_tune(ima_ctx *ctx, int sweeps, int flags, int *best, double *ms_per_sweep);
/* HKY site-block size: every HKY work unit (one wavefront) covers 64
 * patterns, one per lane, or 128, two per lane, which share the program,
 * coefficient and tip reads.  0 (the default) uses 64 for ima_score and
 * ima_mc3_setup, and lets ima_mc3_tune time both on the chains' sweeps and
 * keep the faster; 64 or 128 fixes the size for the calls that follow.
 * ima_mc3_site_block reads the size the chains' sweeps use. */
int ima_set_site_block(ima_ctx *ctx, int patterns);
int ima_mc3_site_block(ima_ctx *ctx, int *patterns);
/* Decisions for genealogy proposals made by IMa2p's own updategenealogy
 * (update_gtree.cpp:723-966: slider / joinsisdown / splitsisdown /
 * addmigration on the host).  The host wrote chain c's proposal at locus l
 * into slot prop_first + c*nloci + l (ima_slots_write, ima_slots_write_pops,
 * ima_sw_update_A for the allele at a stepwise locus); those slots must not
 * overlap the chains' own.  One call scores every proposal in one batched
 * launch per model with its genealogy weights, then decides each chain's
 * loci in order, as the sequence of updategenealogy calls in qupdate does
 * (ima_main_mpi.cpp:1821-1841): accept when
 *   U < min(1, exp(beta (dprobg + gbeta (newpdg - pdg)) + h))
 * (update_gtree.cpp:919-927) with probg from integrate_tree_prob over the
 * chain's all-loci weights after the earlier loci' decisions, h =
 * hastings[i] (migweight + slideweight + Atermsum; NULL: 0) and U =
 * uniforms[i] (the uniform() of the decision; NULL: the engine's own
 * counter-based stream).  An incompatible infinite-sites genealogy or a
 * zero stepwise term is rejected (rejectIS).  gbeta is the reference's
 * thermodynamic-integration multiplier (1 otherwise).  Accepted proposals
 * replace the chains' genealogies (slots, pdg, weights, probg); rejected
 * ones leave them as they were.  accepted (may be NULL) receives 1 per
 * accepted item.  The chains' pending device proposals are never touched:
 * call between ima_mc3_run calls. */
int ima_mc3_decide(ima_ctx *ctx, int prop_first, const double *hastings,
                   const double *uniforms, double gbeta, int32_t *accepted);
/* savegsampinf (ginfo.cpp:319-377), as savegenealogyinfo calls it every
 * recording interval: appends the cold chain's genealogy summary (the last
 * chain whose beta is exactly 1.0, whichiscoldchain ima_main_mpi.cpp:
 * 3023-3033) to a table made by ima_lmode_table_reserve with the sweep's
 * parameters, on the device, without waiting: per population-size
 * parameter the count of cc and the float sums of fc and hcc over its
 * positions, per migration parameter the count of mc and the float sum of
 * fm, qintegrate, mintegrate, the summed pdg, probg and the split times,
 * all as float.  IMA_ERR_SIZE once the table is full (the reference stops
 * at MAXGENEALOGIESTOSAVE). */
int ima_mc3_record(ima_ctx *ctx, int table_id);
/* The chains' all-loci weights wall[nchains][R] (allgweight, R as for
 * ima_integrate_tree_prob) and their integrals qint[nchains][npsize +
 * nmig] (allpcalc.qintegrate then mintegrate); either may be NULL. */
int ima_mc3_read_weights(ima_ctx *ctx, double *wall, double *qint);
/* probg[nchains] and split times tvals[nchains][P-1] of the sweep. */
int ima_mc3_read_prior(ima_ctx *ctx, double *probg, double *tvals);
/* Migration times of one slot as the device holds them (the rescaling
 * moves them); mig_time has that slot's migration count entries. */
int ima_slots_read_mig(ima_ctx *ctx, int slot, double *mig_time);
/* The migration events of one slot as the device holds them: *count, and
 * (any may be NULL) each event's edge, time and entered population, sorted
 * by edge then time -- what a .mcf file records per edge. */
int ima_slots_read_mig_events(ima_ctx *ctx, int slot, int32_t *count, int32_t *mig_edge,
                              double *mig_time, int32_t *mig_to);

/* ---- serving ----------------------------------------------------------- */
/* A request loop for independent IM analyses and L-mode queries arriving
 * over time (no counterpart in the reference, which serves one analysis
 * per process).  Requests wait in arrival order and are admitted, first
 * fit, while their HBM footprint fits the budget; admitted L-mode queries
 * on one table are batched into shared launches, admitted analyses run in
 * quanta of sweeps on streams of their own, all overlapping on the device.
 * ima_server_step makes progress without blocking (except while building
 * an admitted analysis' context); latency runs from submission to the step
 * that sees the request finish. */
typedef struct ima_server ima_server;
enum { IMA_REQ_WAITING = 0, IMA_REQ_RUNNING = 1, IMA_REQ_DONE = 2, IMA_REQ_FAILED = 3 };
enum { IMA_REQ_LMODE = 0, IMA_REQ_ANALYSIS = 1 };

/* One IM analysis, flat (every array is copied at submission).  Loci:
 * model[l] (IMA_MODEL_*), n[l] gene copies, S[l] patterns (HKY) or sites
 * (IS), totsites[l], pi[l][4]; seq packs the HKY and IS loci' n x S data in
 * locus order, mult the HKY pattern counts; samppop[l][npops].  The
 * population model takes ima_model_add's arrays.  Genealogies are packed
 * chain-major (slot c*nloci + l, 2n-1 entries each): up0, up1, down, time,
 * pop, A (may be NULL), root[c*nloci + l], nmigev[c*nloci + l] and the
 * migrations (edge, time, destination) packed in slot order.  u and kappa
 * per (chain, locus), beta per chain, tvals[c][npops-1] split times with
 * uniform prior on [0, tmax]; nsweeps sweeps of ima_mc3_run with
 * IMA_MC3_GRAPH | IMA_MC3_SCALE. */
typedef struct {
  int nloci, nchains, npops;
  const int32_t *model, *n, *S, *totsites;
  const int8_t *seq;
  const int32_t *mult;
  const double *pi;
  const int32_t *samppop;
  const int32_t *plist, *npresent, *popdown, *pope, *join;
  int npsize;
  const int32_t *theta_off, *theta_wp;
  int nmig;
  const int32_t *mig_off, *mig_wp;
  const double *qmax, *mmax;
  const int32_t *up0, *up1, *down, *root, *A, *pop, *nmigev, *mig_edge, *mig_to;
  const double *time, *mig_time;
  const double *u, *kappa, *beta, *tvals;
  double tmax;
  uint64_t seed;
  int precision, nsweeps;
} ima_analysis;

/* budget_bytes <= 0: 80% of the device's free memory.  An analysis' sweeps
 * are ima_mc3_run's with IMA_MC3_GRAPH | IMA_MC3_TOPOLOGY | IMA_MC3_SCALE |
 * IMA_MC3_RY1 | IMA_MC3_NW | IMA_MC3_UPDATE_U (updategenealogy's move, the
 * rubber-band or NW split-time update when it has two or more
 * populations, the mutation-rate rounds every fifth sweep).  It
 * runs in quanta of sweeps: the first is one sweep, later ones about 4 ms of GPU
 * time at its measured sweep rate, at most `quantum` sweeps (<= 0: 25).
 * Short quanta keep L-mode launches from queueing behind long runs of
 * sweeps on the hardware queues that streams share. */
int ima_server_create(int device, double budget_bytes, int quantum, ima_server **out);
void ima_server_destroy(ima_server *srv);
const char *ima_server_last_error(const ima_server *srv);
/* A resident genealogy table (ima_lmode_table's arguments). */
int ima_server_table(ima_server *srv, const float *table, int G, int K, int npsize,
                     int nmig, int npops, int *table_id);
/* jointp at npoints points (ima_lmode_jointp's arguments). */
int ima_server_submit_lmode(ima_server *srv, int table_id, const double *x, int npoints,
                            int modeltype, int prange0, int prange1, int mode,
                            int64_t *req);
int ima_server_submit_analysis(ima_server *srv, const ima_analysis *a, int64_t *req);
/* Progress: completions, admissions, launches.  *pending (may be NULL)
 * receives the requests not yet finished. */
int ima_server_step(ima_server *srv, int *pending);
/* state (IMA_REQ_*), latency (-1 until finished) and the HBM footprint the
 * request is charged: estimated from its sizes while it waits, measured
 * from its context once admitted. */
int ima_server_status(ima_server *srv, int64_t req, int *state, double *latency_s,
                      double *bytes);
int ima_server_lmode_result(ima_server *srv, int64_t req, double *out, double *ess);
/* pdg[nchains*nloci], probg[nchains], beta[nchains], stats[8] as
 * ima_mc3_read's. */
int ima_server_analysis_result(ima_server *srv, int64_t req, double *pdg, double *probg,
                               double *beta, uint64_t *stats);
/* Forget a finished request (frees its host copy and results). */
int ima_server_release(ima_server *srv, int64_t req);
/* p50 / p99 / mean latency in seconds of finished requests of a kind
 * (IMA_REQ_LMODE, IMA_REQ_ANALYSIS, -1 for all). */
int ima_server_latency(ima_server *srv, int kind, int *count, double *p50, double *p99,
                       double *mean);
int ima_server_usage(ima_server *srv, double *used_bytes, double *budget_bytes,
                     int *waiting, int *running);

/* ---- timing ------------------------------------------------------------ */
/* Per-kernel HIP-event timing on the context's stream (off by default).
 * ima_timing_get returns, for kernel class k (0 HKY FP64, 1 HKY FP32,
 * 2 IS, 3 SW, 4 L-mode, 5 swap, 6 MC3 sweep, 7 the sweep's accept step,
 * 8 ima_prior's integrals), the number of timed launches and their total
 * milliseconds.  Classes 6 and 7 are timed by ima_mc3_run with
 * IMA_MC3_TIMED. */
int ima_timing_enable(ima_ctx *ctx, int on);
int ima_timing_get(ima_ctx *ctx, int kclass, int64_t *launches,
                   double *total_ms);
int ima_timing_reset(ima_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
