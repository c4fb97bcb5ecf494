/*
 * meinv.h — C ABI of the MI355X Milne–Eddington inversion engine (libmeinv.so).
 *
 * The library is a drop-in for the per-pixel inversion loop of MEINV
 * (HaoLi-70/Milne-Eddington).  Every entry point takes plain pointers and
 * sizes; nothing in these signatures is a C++ or torch type.
 *
 * Units follow the reference after its control-file conversion
 * (Src/RINPUT.c:242-288): B [G], theta/phi [rad], v_los [km/s],
 * Doppler width [mÅ], damping a, eta0, S = S0+S1, beta = S0/S.
 *
 * Profile layout: [n][4][nl], Stokes I,Q,U,V blocks, the C order of the FITS
 * cube [nl,4,nx,ny] (Src/IO.c:288-377).  Result layout: [n][10] =
 * 9 parameters + chi2/normp, zeros for a pixel that is not inverted
 * (Src/LM_FIT.c:791-803).
 */
#ifndef MEINV_H
#define MEINV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int me_status;
#define ME_OK        0
#define ME_EINVAL    1   /* bad argument or configuration */
#define ME_EDEVICE   2   /* HIP runtime error or no device */
#define ME_ENOMEM    3   /* device allocation failed */
#define ME_ENOTREADY 4   /* learned initialiser requested but not trained/loaded */

/* precision of synthesis, Faddeeva, Jacobian and normal matrix.  The 9x9 SVD
 * solve is FP64 (unless me_config.svd_fp32 opts into FP32 on the FP32 path)
 * and the chi2 accept/reject test FP64 in both. */
#define ME_FP64 0
#define ME_FP32 1

/* where the LM starting point comes from */
#define ME_INIT_HEURISTIC 0  /* Init_Guess of the reference (Src/ME_SOLVER.c:445-676) */
#define ME_INIT_LEARNED   1  /* MLP prediction, one LM run, fallback to heuristic runs */

/* per-pixel status bits (me_stats.status) */
#define ME_PIX_INVALID   1   /* NaN or mean(I) < Icriteria: zero record */
#define ME_PIX_SVD_FAIL  2   /* a 9x9 SVD hit 50 QR iterations (the reference aborts, Src/SVD_TEMPLATE.h:283-286) */
#define ME_PIX_FALLBACK  4   /* learned start missed chisq_criteria, heuristic runs done */

/*
 * Control-file keywords consumed by the inversion loop, in internal units.
 * Defaults (me_config_defaults) are the reference's keyword table,
 * Src/RINPUT.c:400-458.
 */
typedef struct me_config {
    double lambda0;            /* `lines` [Å]                               */
    double geff;               /* `lines` effective Landé factor             */
    int    num_iter;           /* 60                                        */
    int    num_run;            /* 4 (up to 4*num_run LM runs, LM_FIT.c:701) */
    double convg_criteria;     /* 1e-3                                      */
    double chisq_criteria;     /* 1.0                                       */
    double weights[4];         /* 1,1,1,1                                   */
    double damp;               /* 1e-3                                      */
    double lambda_accept;      /* 5                                         */
    double lambda_reject;      /* 5                                         */
    double lambda_limits[2];   /* 1e-5, 1e3                                 */
    double bounds[9][2];       /* Bounds_*; [1],[2] in rad (never applied)  */
    int    fixed[9];           /* *_fixed: parameter held at its starting value (npar < 9) */
    int    regul[9];           /* *_Regularization flag                     */
    double regul_value[9];     /* constant (theta, phi already x pi)        */
    double regul_weight[9];    /* weight                                    */
    double vcoeffi;            /* 100                                       */
    double lcoeffi;            /* 250                                       */
    double icriteria;          /* 100                                       */
    double svd_threshold;      /* 1e-6                                      */
    int    hmi_ref;            /* 0                                         */
    int    nfigures;           /* 6 (1..8)                                  */
    double step[9];            /* 1000,2,2,2,20,0.3,30,1e5,0.3              */
    int    output_fit;         /* 0                                         */
    /* engine settings (no counterpart keyword) */
    int      precision;        /* ME_FP64 | ME_FP32                         */
    int      init_mode;        /* ME_INIT_HEURISTIC | ME_INIT_LEARNED       */
    uint64_t seed;             /* Philox4x32-10 key of the restart jumps    */
    int      block_size;       /* 0 = 256; else 64/128/256 (me_ctx_autotune) */
    int      waves_per_cu;     /* 0 = all that fit; resident waves per CU    */
    /* upper damping bound of the learned run only (ME_INIT_LEARNED); 4e-3
     * by default, 0 = lambda_limits[1].  The damping never enters the solve
     * (trial k >= 2 of an iteration solves with the previous trial's U), it
     * only bounds the trials of one iteration: damp * lambda_reject^r > bound
     * ends the run after r rejected trials, so with the default damping
     * (1e-3, x5 per rejection) 4e-3 allows one trial per iteration, 6e-3 two
     * (the second solves with the normal matrix's U; later ones would solve
     * with the U of an orthogonal matrix, which rounding decides).  The
     * heuristic runs of the fallback always use lambda_limits[1]. */
    double   learned_lambda_limit;
    /* upper bound of S in the learned run, in units of the profile's I
     * maximum; 2 by default, 0 keeps Init_Guess's 1.4.  Init_Guess starts S at
     * Imax and bounds it to [0.6, 1.4] Imax (Src/ME_SOLVER.c:662-666), so a
     * line that never reaches the continuum inside the wavelength window
     * (S above 1.4 Imax) cannot be fitted; the learned run starts from the
     * network's S and may go up to this bound.  Every heuristic run, the
     * fallback's included, keeps [0.6, 1.4] Imax. */
    double   learned_s_limit;
    /* FP32 path only: 0 (default) keeps the 9x9 SVD solve in FP64, as the
     * normal matrix's conditioning needs; 1 runs it in FP32, an opt-in
     * throughput mode.  The FP32 SVD fits in half the register file, and its
     * kernel build runs two waves per SIMD (8 per CU), which hides the SVD's
     * dependency latency: hmi4k plain 21 % faster with the converged
     * fraction kept (82.25 % against 82.40 %), but fewer pixels at the FP64
     * solution: against the FP64 oracle, 77 % of 2048 hmi4k plain pixels
     * within 0.1 G / 0.01 deg / 1 m/s against 87 % with the FP64 SVD, and
     * 90 % of learned pixels against 96 %.  Learned maps gain no speed from
     * it, and the nl = 60 and 112 maps lose 15 %.  Results depend on this
     * setting, never on the launch shape. */
    int      svd_fp32;
} me_config;

/* Aggregate counters and optional per-pixel arrays filled by me_invert_*.
 * Event counts feed the algorithmic op model (FP32/FP64 FLOP per event). */
typedef struct me_stats {
    /* caller-owned per-pixel arrays, host memory, may be NULL */
    int32_t *iters;            /* LM iterations summed over runs            */
    int32_t *runs;             /* LM runs executed                          */
    int32_t *trials;           /* trial syntheses                           */
    int32_t *status;           /* ME_PIX_* bits                             */
    /* aggregates over the batch */
    int64_t n_pixels, n_valid, n_converged;
    int64_t n_runs, n_iters;
    int64_t n_jac;             /* syntheses with Jacobian                   */
    int64_t n_trial;           /* syntheses without Jacobian                */
    int64_t n_hess;            /* normal-matrix builds                      */
    int64_t n_svd;             /* 9x9 SVDs                                  */
    int64_t n_bk;              /* back-substitutions                        */
    int64_t n_fadd[8];         /* Faddeeva calls by branch: 916, Hui p6, asym1, asym2, asym3, asym4, y=0, nu=0
                                  (ME_FP32 evaluates one rational approximation everywhere and counts each call
                                  by the region the FP64 evaluation would take) */
    int64_t n_svd_fail;
    int64_t n_fallback;
    double  kernel_ms;         /* inversion kernel time, HIP events on the context stream */
    double  total_ms;          /* whole call, including staging             */
} me_stats;

typedef struct me_ctx me_ctx;

/* Fill cfg with the reference keyword defaults (Src/RINPUT.c:400-458). */
void me_config_defaults(me_config *cfg);

/*
 * Create a context for one wavelength grid.  Replaces INIT_INV's derived
 * constants (Src/INV_INIT.c:121-205): Zeeman shift, ncut, line-centre index,
 * Init_Guess window, v_los bounds, Faddeeva table.  Allocates device state on
 * the current HIP device and creates the context's stream.
 */
me_status me_ctx_create(const me_config *cfg, const double *lambda, int nl, me_ctx **out);
void      me_ctx_destroy(me_ctx *ctx);

/* Global index of the first pixel of the next batches (default 0).  Restart
 * jumps at irun >= 4 draw from Philox4x32-10 with counter (base + index of the
 * pixel in the batch, run, stream), so a pixel gets the same draws whichever
 * batch it is inverted in: a host that splits a map into batches, or resumes
 * an interrupted run, passes each batch's offset in the map.  Replaces the
 * reference's time-seeded RNG state (Src/RANDOM_NUMBER.c:25-61, used by
 * Random_Jump, Src/LM_FIT.c:185-197), which is not reproducible at all. */
me_status me_ctx_set_pixel_base(me_ctx *ctx, int64_t base);

/* Derived constants of the context: ic, i0, i1, ncut and v_los bounds. */
me_status me_ctx_info(const me_ctx *ctx, int *ic, int *i0, int *i1, int *ncut, double vlos_bounds[2]);

/* Whether ME_FP32 inversions classify every Faddeeva call by the region the
 * FP64 evaluation would use and count it in me_stats.n_fadd (on after
 * me_ctx_create).  The classification costs a few per cent of kernel time;
 * results do not depend on it.  No counterpart in the reference: the counts
 * feed the benchmark's operation model. */
me_status me_ctx_set_region_counts(me_ctx *ctx, int on);

/* Launch-shape autotuner of the persistent inversion kernel: inverts the
 * first n_sample device profiles under every candidate block size (64, 128,
 * 256 lanes) and resident-wave count per CU (one wave per SIMD, and two with
 * the FP32 SVD) that the occupancy allows (best of two runs each, the context's
 * init mode and precision), keeps the fastest in the context's block_size /
 * waves_per_cu and reports it.  The sample should hold some 64 pixels per
 * resident lane (4M pixels on 256 CUs): a smaller one is timed by its slowest
 * pixels.  Results never depend on the launch shape.  No counterpart in the
 * reference (its MPI chunk size, nprofiles, is not a kernel shape). */
me_status me_ctx_autotune(me_ctx *ctx, const void *d_prof, int prof_is_f64, int64_t n_sample,
                          int *block_size, int *waves_per_cu, double *kernel_ms);

/*
 * Invert n pixels.  Replaces INVERSION_MULTI (Src/LM_FIT.c:746-811,
 * declared Src/LM_FIT.h:43-44) together with the MPI pixel scatter of
 * Src/ME_INV.c:139-283: one launch of a persistent work-queue kernel.
 *   prof  [n][4][nl] host float32 (FITS order)
 *   res   [n][10]    host double (out)
 *   fit   [n][4][nl] host double (out, may be NULL; written when non-NULL)
 *   init  [n][9]     host double explicit starting points, or NULL (learned
 *                    mode: NULL seeds each pixel with the network's prediction)
 *   stats            may be NULL
 * A bad pixel never aborts the batch: it gets a zero record and a status bit.
 */
me_status me_invert_batch(me_ctx *ctx, const float *prof, int64_t n, double *res,
                          double *fit, const double *init, me_stats *stats);

/* Same, with FP64 profiles (the reference reads every BITPIX into double). */
me_status me_invert_batch_f64(me_ctx *ctx, const double *prof, int64_t n, double *res,
                              double *fit, const double *init, me_stats *stats);

/* Same, with every buffer already resident in device memory (HBM).
 * prof_is_f64 selects float32 (0) or float64 (1) profiles. */
me_status me_invert_device(me_ctx *ctx, const void *d_prof, int prof_is_f64, int64_t n,
                           double *d_res, double *d_fit, const double *d_init, me_stats *stats);

/*
 * me_invert_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The kernel's per-lane scratch is a static device array of the library
 * (190 doubles for each of at most 65536 resident lanes, 99.6 MB per GPU,
 * held until the process ends) instead of the context's buffer; the work and
 * event counters are the context's and are cleared by a fill kernel.  Same
 * results as me_invert_device.  Limits: no best-fit profiles; learned mode
 * needs d_init (predictions from me_mlp_predict_device, which may allocate
 * on its first call); stats may be NULL or hold NULL per-pixel arrays
 * (ME_EINVAL otherwise); an explicit launch shape (block_size /
 * waves_per_cu) needing more than 65536 lanes is ME_EINVAL, the default one
 * is capped there.  Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (from two contexts or two host threads) race.
 */
me_status me_invert_device_noalloc(me_ctx *ctx, const void *d_prof, int prof_is_f64, int64_t n,
                                   double *d_res, const double *d_init, me_stats *stats);

/*
 * Forward synthesis with optional analytic Jacobian.  Replaces
 * Milne_Eddington_Single (Src/ME_SOLVER.c:101-366).
 *   par    [n][9]        host double
 *   stokes [n][4][nl]    host double (out)
 *   jac    [n][4][9][nl] host double (out, may be NULL); row p scaled by step[p]
 *                        (step[7] as given in the config)
 */
me_status me_synth_batch(me_ctx *ctx, const double *par, int64_t n, double *stokes, double *jac);

/*
 * Synthetic Stokes map on the device: truths from Philox4x32-10
 * (key = seed, counter = pixel, block, stream), forward synthesis in FP64,
 * Gaussian noise N(0, (sigma*S)^2), written as float32 [n][4][nl].
 *   ranges     [9][2] uniform ranges of the 9 truths
 *   ff_frac    fraction of field-free pixels (B ~ U(0, ff_bmax)); 0 disables
 *   d_prof     device float32 [n][4][nl] (out)
 *   d_truth    device double [n][9] (out, may be NULL)
 */
me_status me_gen_map_device(me_ctx *ctx, uint64_t seed, const double *ranges, double ff_frac,
                            double ff_bmax, double sigma, int64_t n, float *d_prof, double *d_truth);

/* ---- learned initialiser: MLP 4*nl -> 256 -> 256 -> 256 -> 10, GELU, BF16 MFMA ---- */

/* Train the initialiser on profiles synthesised on the device from the truth
 * ranges (same sampler as me_gen_map_device).  Returns the final training loss. */
me_status me_mlp_train(me_ctx *ctx, uint64_t seed, const double *ranges, double ff_frac,
                       double ff_bmax, double sigma, int64_t n_train, int batch, int epochs,
                       double lr, double *final_loss);

/* Predict starting points for n device-resident profiles: d_init [n][9] double. */
me_status me_mlp_predict_device(me_ctx *ctx, const float *d_prof, int64_t n, double *d_init);

/* Export / import the FP32 master weights (host buffer of me_mlp_num_params() floats). */
int64_t   me_mlp_num_params(const me_ctx *ctx);
me_status me_mlp_get_params(const me_ctx *ctx, float *params);
me_status me_mlp_set_params(me_ctx *ctx, const float *params);

/* Use a network trained elsewhere (e.g. exported from another context with
 * me_mlp_get_params): its weights and the truth ranges [9][2] it was trained
 * on, which decode its outputs.  The context's learned mode is ready after
 * this call.  No counterpart in the reference. */
me_status me_mlp_import(me_ctx *ctx, const float *params, const double *ranges);

/* One forward+backward pass of the MLP on a host batch, for gradient checks:
 *   x [batch][4*nl] normalised inputs, y [batch][10] targets,
 *   grad [me_mlp_num_params] (out), returns the mean squared error. */
me_status me_mlp_grad(me_ctx *ctx, const float *x, const float *y, int batch,
                      float *grad, double *loss);

/* Library build identification (gfx target, HIP version). */
const char *me_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MEINV_H */
