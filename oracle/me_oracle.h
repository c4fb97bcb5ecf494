/*
 * me_oracle.h — CPU FP64 oracle of the Milne–Eddington inversion path.
 *
 * TEST INFRASTRUCTURE ONLY.  This restates the reference algorithm
 * (HaoLi-70/Milne-Eddington, Src/FADDEEVA.c, Src/ME_SOLVER.c, Src/LM_FIT.c,
 * Src/SVD_TEMPLATE.h, Src/INV_INIT.c) on the CPU so that tests, the smoke
 * check and the benchmark's CPU-baseline leg can check and time the device
 * path.  Nothing in the product links, loads or calls it.
 *
 * How each function is pinned (tests/ hold the checks):
 *   orc_faddeeva      bit-for-bit against the reference's own FADDEEVA.c built
 *                     into oracle/_ref (fixtures in tests/golden), and against
 *                     scipy.special.wofz to the nfigures accuracy
 *   orc_svd/svbksb    bit-for-bit against the reference's SVD.c in oracle/_ref
 *                     (fixtures), reconstruction A = U W V^T, numpy singular values
 *   orc_synth         closed forms (B=0, theta=0, eta0->0, continuum limit,
 *                     phi+pi invariance, line-centre symmetry) and an
 *                     independent 4x4 propagation-matrix solve of the
 *                     Unno-Rachkovsky system
 *   orc_synth jac     central finite differences of orc_synth
 *   hessian/gradient  numpy brute force from J, ivnoise, residual
 *   first trial solve numpy SVD pseudo-inverse with the same truncation loop
 *   trial k>=2 (H1)   clip(U^T g) with U from the _ref SVD of the previous matrix
 *   bounds/folds      closed forms
 *   philox            Random123 known-answer vectors (Salmon et al. 2011)
 *   init guess        hand-derived values on constructed profiles; the
 *                     branch thresholds, H4 and H7 on constructed grids
 *   run policy H3/H9  run counts and save order on constructed chi2 sequences
 *                     through the orc_run_policy hook
 *   LM loop control   trial counts under the damping bound, the damping reset
 *                     per iteration, the accept and convergence tests, the
 *                     num_iter cap and the regularisation ratio updates, on
 *                     constructed chi2/penalty sequences through orc_lm_replay,
 *                     which runs the same control function as the oracle's LM
 *   Get_Regul         the regularised normal equations against the
 *                     Gauss-Newton equations of the augmented residual vector
 *                     (numpy), with fixed parameters shifting the packed index
 *   Random_Jump       runs 1-3 by hand; runs >= 4 by recovering the eight
 *                     draws from the jumped parameters (which draw moves which
 *                     parameter, guess or best as the base), Box-Muller draws
 *                     tested for N(0,1) (Kolmogorov-Smirnov) and independence
 *   learned policy    constructed cases: a prediction at the truth converges in
 *                     one run; a bad prediction falls back to the heuristic
 *                     runs and keeps the better record
 * Parity unpinned: the end-to-end INVERSION records themselves — the reference's
 * ME_SOLVER.c/LM_FIT.c include fitsio.h, which the image lacks, so they cannot
 * be built here; each stage is pinned as listed instead.
 */
#ifndef ME_ORACLE_H
#define ME_ORACLE_H

#include <stdint.h>
#include "../include/meinv.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Faddeeva branch ids (same order as me_stats.n_fadd) */
enum { ORC_F916 = 0, ORC_FHUI = 1, ORC_FAS1 = 2, ORC_FAS2 = 3, ORC_FAS3 = 4,
       ORC_FAS4 = 5, ORC_FY0 = 6, ORC_FNU0 = 7, ORC_FHUM = 8 };

typedef struct orc_fadd {
    int nfig, ifig, nmax;
    double rmin, log_rmin, sqrt_log_rmin, a, aa;
    double e[50];
} orc_fadd;

void orc_fadd_init(orc_fadd *F, int nfigures);
int  orc_faddeeva(double nu, double y, const orc_fadd *F, double *H, double *L);
void orc_faddeeva_vec(int64_t n, const double *nu, const double *y, int nfigures,
                      double *H, double *L, int *branch);

/* Golub–Reinsch SVD of an m x n row-major matrix (m >= n); a is replaced by U.
 * Returns 0, or 1 when a singular value needed SVD_MAX_ITER QR iterations. */
int  orc_svd(double *a, int m, int n, double *w, double *v);
void orc_svbksb(const double *u, const double *v, const double *w, int m, int n,
                const double *b, double *x);

/* Philox4x32-10 */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* context: configuration plus INV_INIT derived constants */
typedef struct orc_ctx {
    me_config cfg;
    int nl, ncut, ic, i0, i1, npar;
    double *lambda;
    double bshift;
    double limits[9][2];
    double wsq[4];
    orc_fadd F;
    int regul_any;
} orc_ctx;

orc_ctx *orc_ctx_create(const me_config *cfg, const double *lambda, int nl);
void     orc_ctx_destroy(orc_ctx *c);
void     orc_ctx_info(const orc_ctx *c, int *ic, int *i0, int *i1, int *ncut, double *vlos);

/* forward synthesis, syn [4][nl]; jac [4][9][nl] or NULL; step[7] from cfg */
void orc_synth(const orc_ctx *c, const double *par, const double *step, double *syn, double *jac,
               int64_t *fcount);
void orc_synth_batch(const orc_ctx *c, const double *par, int64_t n, double *syn, double *jac);

void orc_bounds_check(double *p, const double lim[9][2]);

/* Init_Guess on one pixel profile (prof points at the pixel inside a
 * contiguous [n][4][nl] buffer; pix and n locate it for the H7 reads).
 * Returns 1 when mean(I) < Icriteria.  Outputs guess[9], limit7[2], step7,
 * norm0, normp. */
int orc_init_guess(const orc_ctx *c, const double *prof, int64_t pix, int64_t n,
                   double *guess, double *limit7, double *step7, double *norm0, double *normp);

/* H2: how LM->chisq (the previous LM_FIT's last trial chi2) seeds the
 * regularisation ratio test of the next LM_FIT (Src/LM_FIT.c:557). */
#define ORC_H2_SEQUENTIAL 0  /* one state carried across pixels in order, 0 at start */
#define ORC_H2_PER_PIXEL  1  /* no state before a pixel's first run: test is false  */

typedef struct orc_pixstats {
    int32_t iters, runs, trials, status;
} orc_pixstats;

/* orc_invert with the learned policy and the restart key:
 *   learned   != 0: init holds the learned initialiser's predictions; one LM
 *             run from each, heuristic runs (Init_Guess, H3/H9) only when its
 *             chi2/normp exceeds chisq_criteria, the better record kept
 *             (status ME_PIX_FALLBACK)
 *   pix_base  global index of pixel 0: the Philox counter of restart jumps */
void orc_invert_ex(const orc_ctx *c, const double *prof, int64_t n, double *res, double *fit,
                   const double *init, int learned, int64_t pix_base, int h2_mode, int nthreads,
                   orc_pixstats *pst, int64_t *ev);

/* Invert n pixels (double profiles).  init may be NULL.  pst may be NULL.
 * ev (may be NULL) receives event counts: [0]=jac [1]=trial [2]=hess [3]=svd
 * [4]=bk [5..12]=Faddeeva branches.  nthreads > 1 splits pixels into chunks of
 * 20 handed out dynamically (each thread has its own H2 state). */
void orc_invert(const orc_ctx *c, const double *prof, int64_t n, double *res, double *fit,
                const double *init, int h2_mode, int nthreads, orc_pixstats *pst, int64_t *ev);

/* single LM stages: normal equations at par (no regularisation), and the
 * truncated SVD solve that overwrites H with U */
void orc_normal_eq(const orc_ctx *c, const double *par, const double *prof, double step7, double norm0,
                   double *H, double *g, double *chi2);
void orc_solve(const orc_ctx *c, double *H, const double *g, double *sol);

/* Synthetic map: same sampler and Philox streams as the device generator.
 * ranges [9][2]; writes prof [n][4][nl] (float) and truth [n][9]. */
void orc_gen_map(const orc_ctx *c, uint64_t seed, const double *ranges, double ff_frac,
                 double ff_bmax, double sigma, int64_t pix0, int64_t n, float *prof, double *truth);

/* Run-policy hook for tests: replays INVERSION's save/break policy
 * (Src/LM_FIT.c:701-724) on a given per-run chi2 sequence.  Returns the
 * number of runs executed; *best_run is the index whose record is kept. */
int orc_run_policy(int nruns, double criteria, const double *chisq_seq, int nseq, int *best_run);

/* LM-control hook for tests: runs the oracle's LM loop control (the function
 * orc_invert's LM uses) with the start chi2 chi0 and penalty pen0, and the
 * k-th trial returning trial_chi[k] / trial_pen[k] (the last entry repeats).
 * *lm_chisq is the H2 state in and out.  Returns the iteration count; writes
 * the trial count, the final chi2 and the regularisation ratio handed to each
 * normal-matrix build (up to nratio_max, *nratios = number of builds). */
int orc_lm_replay(const me_config *cfg, int regul_any, double chi0, double pen0, double *lm_chisq,
                  const double *trial_chi, const double *trial_pen, int ntrial, int *trials, double *chisq,
                  double *ratios, int nratio_max, int *nratios);

/* Get_Regul with the matrix flag on a packed npar x npar H and g[npar]
 * (in place); returns the penalty */
double orc_regul_probe(const orc_ctx *c, const double *par, double ratio, double *H, double *g);

/* Random_Jump for run irun of a pixel with restart key `key`: best, guess,
 * lim [9][2] in, par out */
void orc_jump_probe(uint64_t seed, int64_t key, int irun, const double *best, const double *guess,
                    const double *lim, double *par);

/* four N(0,1) draws (two Box-Muller pairs) of Philox block (key, c2, stream) */
void orc_gauss4_probe(uint64_t seed, int64_t key, uint32_t c2, uint32_t stream, double *g);

#ifdef __cplusplus
}
#endif

#endif
