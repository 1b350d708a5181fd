/*
 * lcf_hip.h -- C ABI of the MI355X (gfx950) batched light-curve fitting
 * library.  This is the boundary a `hip` backend of the reference crate binds
 * (a src/gpu_hip.rs next to src/gpu_types.rs, selected like the `cuda`
 * backend in src/gpu.rs:12-18 and src/lib.rs:5-10).
 *
 * Conventions
 *   - Every call returns 0 on success or a negative LCF_E* code; the message
 *     of the last failure on a context is lcf_last_error(ctx) (of the calling
 *     thread's last context-free call: lcf_last_error(NULL)).  This replaces
 *     the reference's `Err(format!("CUDA error {code}: {msg}"))`
 *     (src/gpu_cuda.rs:170-186).
 *   - Plain host pointers and sizes only.  Inputs are read, outputs written,
 *     before the call returns (each call synchronises its context stream).
 *   - Model ids are GpuModelName::model_id (src/gpu_types.rs:51-62):
 *     0 Bazin, 1 Villar, 2 Tde, 3 Arnett, 4 Magnetar, 5 ShockCooling,
 *     6 Afterglow, 7 MetzgerKN.  Per-source parameter rows are padded to
 *     LCF_MAX_PARAMS doubles.
 *   - One context per host thread (the reference GpuContext is tied to the
 *     calling thread's current device, src/gpu_cuda.rs:297-312).
 */
#ifndef LCF_HIP_H
#define LCF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LCF_MAX_PARAMS 7
#define LCF_N_MODELS 8

/* status codes */
#define LCF_OK 0
#define LCF_E_INVALID (-1)   /* bad argument (size, model id, null pointer) */
#define LCF_E_HIP (-2)       /* HIP runtime error (message has the HIP text) */
#define LCF_E_NOMEM (-3)     /* device allocation failed */
#define LCF_E_INTERNAL (-4)  /* inconsistent internal state (a bug: please report) */

/* precision of model evaluation / swarm state in the fitting kernels */
#define LCF_FP32 0           /* FP32 + Kahan-compensated chi-square (default) */
#define LCF_FP64 1           /* FP64 throughout */

/* work distribution for batched fits */
#define LCF_SCHED_PERSISTENT 0  /* one persistent launch, size-sorted device queue (default); a PSO fit
                                   of two or more FP32 models 0-6 (<= 64 particles, no gate) runs all of
                                   them in that one launch, mixed model families on one queue */
#define LCF_SCHED_PER_CLASS 1   /* one launch per power-of-two size class, synchronised */
#define LCF_SCHED_TUNED 2       /* PSO fits: one persistent launch per size bucket with the context's
                                   autotuned launch shape (lcf_ctx_autotune_pso); other fits as
                                   LCF_SCHED_PERSISTENT */
#define LCF_SCHED_PER_MODEL 3   /* PSO fits: one persistent launch per model (per-model launch
                                   times); other fits as LCF_SCHED_PERSISTENT */

/* PSO flags */
#define LCF_PSO_STALL_FROM_FIRST_EVAL 1u /* opt-in: stall baseline = first global best
                                            (the reference's GPU path starts it at +inf) */

typedef struct lcf_ctx lcf_ctx;
typedef struct lcf_batch lcf_batch;

/* ---- context: GpuContext::new (src/gpu_cuda.rs:307-312) ---------------- */
int lcf_ctx_create(int device, lcf_ctx** out);
void lcf_ctx_destroy(lcf_ctx* ctx);
const char* lcf_last_error(const lcf_ctx* ctx);
/* Precision (LCF_FP32 / LCF_FP64) and schedule (LCF_SCHED_*) used by the
 * batched fits of this context.  No counterpart in the reference (FP64 only,
 * one launch per PSO iteration). */
int lcf_ctx_set_precision(lcf_ctx* ctx, int precision);
int lcf_ctx_set_schedule(lcf_ctx* ctx, int schedule);
/* Number of compute units of the context's device. */
int lcf_ctx_num_cus(const lcf_ctx* ctx);
/* Waits for all work of the context's stream and device.  Every batch call
 * already returns synchronised (as the reference's end with
 * cudaDeviceSynchronize, src/gpu_cuda.rs:353-354); this is for callers that
 * time or order work around them. */
int lcf_ctx_synchronize(lcf_ctx* ctx);

/* ---- autotuned launch shapes of the PSO swarms ------------------------- */
/* The reference hard-codes its launch shapes (src/gpu_cuda.rs:1029-1031,
 * 1365-1368).  Here the shape of the persistent swarm launches -- the LDS
 * observation tile (records staged per wave; a tile below the band's size is
 * re-staged every iteration) and the waves per CU -- is measured per
 * band-size bucket: bucket k holds bands of at most 64 << k observations
 * (k < LCF_TUNE_BUCKETS - 1; the last holds the rest).
 * lcf_ctx_autotune_pso times, for every bucket present in `data` and every
 * model, the candidate shapes on a sample of that bucket's bands with the
 * given swarm settings (two launches each, the second timed) and keeps the
 * fastest in the context, for the context's precision and the particles'
 * lane layout.  Fits use the table under LCF_SCHED_TUNED; the shape changes
 * no result (every band is fitted by one wavefront whatever the grid, and
 * the tile only splits the staging of its observations).
 * Per model it then times the tuned per-bucket launches against the single
 * persistent launch of the default shape over all of `data` and records a
 * whole-batch entry (bucket LCF_TUNE_BUCKETS, tile 0, waves 0) when the
 * single launch is not slower; LCF_SCHED_TUNED then keeps that launch.
 * Adjacent buckets of equal shape share one launch.
 * The table is read and written as entries of LCF_TUNE_FIELDS int32:
 * {precision, model, particles per lane (1, 2, 4), bucket, tile,
 * waves per CU}; lcf_ctx_get_tuning returns the number of entries (writing
 * at most max_entries), lcf_ctx_set_tuning replaces the table (n = 0
 * clears it). */
#define LCF_TUNE_BUCKETS 7
#define LCF_TUNE_FIELDS 6
int lcf_ctx_autotune_pso(lcf_ctx* ctx, const lcf_batch* data, const int32_t* models, int n_models,
                         int n_particles, int max_iters, int stall_iters, uint64_t seed);
int lcf_ctx_get_tuning(const lcf_ctx* ctx, int32_t* out, int max_entries);
int lcf_ctx_set_tuning(lcf_ctx* ctx, const int32_t* entries, int n_entries);

/* ---- resident source data: GpuBatchData (src/gpu_cuda.rs:236-292) ------ */
/* offsets[n_sources+1] index the concatenated per-observation arrays
 * (BatchSource fields, src/gpu_types.rs:122-130); is_upper is 0/1.  A batch
 * belongs to a device, not to a context: every context on the device that
 * holds it may use it. */
int lcf_batch_create(lcf_ctx* ctx, int64_t n_sources, const int64_t* offsets,
                     const double* times, const double* flux, const double* obs_var,
                     const int32_t* is_upper, const double* upper_flux,
                     lcf_batch** out);
/* GpuBatchData::new(sources) (src/gpu_cuda.rs:251-285) with the reference's
 * signature: no context; the arrays go to the calling thread's current device
 * (hipSetDevice, or lcf_ctx_create, which makes its device current), as the
 * reference's cudaMalloc uploads do.  Any context on that device may use the
 * batch.  On failure the message is lcf_last_error(NULL), the calling
 * thread's last context-free error. */
int lcf_batch_create_current(int64_t n_sources, const int64_t* offsets, const double* times,
                             const double* flux, const double* obs_var, const int32_t* is_upper,
                             const double* upper_flux, lcf_batch** out);
void lcf_batch_destroy(lcf_batch* batch);
int64_t lcf_batch_n_sources(const lcf_batch* batch);
/* GpuBatchData::n_obs_per_source: observation count, 1 when out of range. */
int64_t lcf_batch_n_obs(const lcf_batch* batch, int64_t source_idx);

/* ---- forward evaluation: GpuContext::eval_batch (src/gpu_cuda.rs:317-360) */
/* out[d*n_times + i] = model(params[d*n_params(model) ..], times[i]), FP64. */
int lcf_eval_batch(lcf_ctx* ctx, int model_id, const double* params, int64_t n_draws,
                   const double* times, int64_t n_times, double* out);

/* ---- batched PSO --------------------------------------------------------- */
/* GpuContext::batch_pso (src/gpu_cuda.rs:368-567).
 * out_params[n_sources*LCF_MAX_PARAMS], out_cost[n_sources]. */
int lcf_batch_pso(lcf_ctx* ctx, int model_id, const lcf_batch* data, int n_particles,
                  int max_iters, int stall_iters, uint64_t seed,
                  double* out_params, double* out_cost);

/* GpuContext::batch_model_select (src/gpu_cuda.rs:573-620): Bazin first,
 * the other models of ALL_GPU_MODELS only where Bazin's cost >= gate.
 * out_model[n_sources] holds model ids. */
int lcf_batch_model_select(lcf_ctx* ctx, const lcf_batch* data, int n_particles,
                           int max_iters, int stall_iters, double bazin_good_enough,
                           int32_t* out_model, double* out_params, double* out_cost);

/* GpuContext::batch_all_models (src/gpu_cuda.rs:624-637): every model of
 * ALL_GPU_MODELS on every source, in that order.
 * out_params[8][n_sources][LCF_MAX_PARAMS], out_cost[8][n_sources]; row m is
 * ALL_GPU_MODELS[m] (ids 0,3,2,6,1,4,5,7). */
int lcf_batch_all_models(lcf_ctx* ctx, const lcf_batch* data, int n_particles,
                         int max_iters, int stall_iters,
                         double* out_params, double* out_cost);

/* General form behind the three calls above: fits models[0..n_models) in
 * order, seed `seed`; when use_gate != 0, models[1..] run only for sources
 * whose models[0] cost is >= gate.  Any output pointer may be NULL.
 *   model_params[n_models][n_sources][7], model_cost[n_models][n_sources],
 *   model_iters[n_models][n_sources] (iterations executed),
 *   best_model/best_cost[n_sources], best_params[n_sources][7].
 * Random numbers: the initial swarm from Philox4x32-10 keyed by seed with the
 * source index in the counter, the velocity updates from one xoshiro128**
 * stream per (source, particle) seeded by a Philox block, so a source's fit
 * does not depend on the batch it is in (the reference's sequential
 * SmallRng stream is not reproduced).  The FP32 cost loop's
 * censored-likelihood table is built once per device by lcf_ctx_create.
 * With FP32 precision the swarms run in FP32 and the cost returned for each
 * model (and compared by the model choice) is the reference's FP64 cost at
 * the returned parameters, from one FP64 evaluation per band and model. */
int lcf_batch_fit_models(lcf_ctx* ctx, const lcf_batch* data, const int32_t* models,
                         int n_models, int use_gate, double gate, int n_particles,
                         int max_iters, int stall_iters, uint64_t seed, uint32_t flags,
                         double* model_params, double* model_cost, int32_t* model_iters,
                         int32_t* best_model, double* best_params, double* best_cost);

/* ---- batched fit without device allocation ------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory inside the call (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): lcf_batch_fit_models with the same
 * arguments, outputs and results, but its device intermediates (per-model
 * results, best-model outputs, model ids, the work-queue counter) live in one
 * fixed scratch array per device (LCF_NOALLOC_SCRATCH_BYTES, held until the
 * process ends) instead of being allocated per call; the counter is cleared
 * by a kernel, not a memset.  Batches of more than LCF_NOALLOC_MAX_BANDS
 * sources are rejected with LCF_E_INVALID.  Every call on one device shares
 * that device's array: calls that overlap on a device (several contexts or
 * threads) race and must be serialised by the caller. */
#define LCF_NOALLOC_MAX_BANDS 16384
#define LCF_NOALLOC_SCRATCH_BYTES 10027520
int lcf_batch_fit_models_noalloc(lcf_ctx* ctx, const lcf_batch* data, const int32_t* models,
                                 int n_models, int use_gate, double gate, int n_particles,
                                 int max_iters, int stall_iters, uint64_t seed, uint32_t flags,
                                 double* model_params, double* model_cost, int32_t* model_iters,
                                 int32_t* best_model, double* best_params, double* best_cost);

/* ---- greedy MultiBazin ---------------------------------------------------- */
/* GpuContext::batch_pso_multi_bazin (src/gpu_cuda.rs:650-963): for K = 1..4
 * Bazin components a swarm over [ln A, t0, ln tau_rise, ln tau_fall] x K, B,
 * ln sigma_x with t0 bounded by the batch time range -/+ 30 d; particle 0 of
 * K >= 2 starts from the K-1 solution plus a component at the largest
 * detection residual; best K by BIC = 2 cost n + (4K+2) ln n, stopping once
 * a K's BIC exceeds the previous one by more than 2.  Outputs per source:
 * best_k, params[18] (4 best_k + 2 used, zero padded; all zero and bic = inf
 * when no K gave a finite BIC), cost, bic, per_k_cost[4] / per_k_bic[4]
 * (NaN after the stop).  The reference's `sources` argument (time ranges and
 * residuals) is read from `data` on the device. */
int lcf_batch_multi_bazin(lcf_ctx* ctx, const lcf_batch* data, int n_particles, int max_iters,
                          int stall_iters, uint64_t seed, int32_t* out_best_k, double* out_params,
                          double* out_cost, double* out_bic, double* out_per_k_cost,
                          double* out_per_k_bic);

/* ---- nonparametric band pipeline (FP64) --------------------------------- */
#define LCF_NP_NFEAT 50  /* NonparametricBandResult fields (src/nonparametric.rs:74-160), band name excluded */
/* process_band (src/nonparametric.rs:1228-1525) for every band of a batch, in
 * one pass on the device: replaces the per-band CPU work of fit_nonparametric
 * (src/nonparametric.rs:1085-1149) and the GPU fit + host feature loop of
 * fit_nonparametric_batch_gpu_with_opts (src/nonparametric.rs:799-1081).
 * Bands are CSR (offsets[n_bands+1]) over raw magnitudes: times, mags,
 * mag_errs; the prediction grid of band b is n_grid points on
 * [t_lo[b], t_hi[b]] (the source's time range; reference: 50 points) and its
 * duration sets the length-scale candidates max(duration / f, 0.1) for f in
 * ls_factors (reference: 3, 6, 12, 24, 48) crossed with amp_candidates
 * (reference: 0.3, 1, 3, 10).  Detections (SNR >= 3) of at most dense_max
 * (reference: 100; at most 200) are fitted by a dense GP selected by train
 * RMS, larger bands by FITC with n_inducing points (reference: 30, at most
 * 32) selected by approx_nlml, plus a dense GP on max_subsample detections
 * (reference: 25) kept for thermal reuse.
 * Outputs: out_status[b] 0 no result (fewer than 5 observations or
 * detections, no usable fit, empty time range), 1 dense, 2 FITC;
 * out_features [n_bands][LCF_NP_NFEAT] in NonparametricBandResult field
 * order (n_obs and n_upper_limits as doubles, NaN for None);
 * out_pred / out_std (rescaled) / out_dpred (analytic d mean/dt)
 * [n_bands][n_grid]; out_state [n_bands][2 S + 4] with S = max(dense_max,
 * max_subsample): alpha[S], x[S], amp, 1/(2 l^2), mean, m of the dense GP
 * kept for thermal reuse (the reference's trained_gps map). */
int lcf_batch_nonparametric(lcf_ctx* ctx, int64_t n_bands, const int64_t* offsets, const double* times,
                            const double* mags, const double* mag_errs, const double* t_lo,
                            const double* t_hi, int n_grid, const double* amp_candidates, int n_amp,
                            const double* ls_factors, int n_ls, int dense_max, int max_subsample,
                            int n_inducing, int32_t* out_status, double* out_features, double* out_pred,
                            double* out_std, double* out_dpred, double* out_state);

/* ---- blackbody colour-temperature fit (FP64) ------------------------------ */
/* fit_thermal (src/thermal.rs:222-418) for n_src sources at once, after the
 * host's band choice (reference band = bluest of u > g > r with >= 5 points;
 * other bands with a known effective wavelength): reference-band
 * observations (CSR ref_off: times, mags, errors) for fit_ref_band_gp when no
 * prefit GP is given; the other bands' observations (CSR col_off) with their
 * wavelengths; lambda_ref and the source's duration.  gp_state rows
 * (2 state_cap + 4 doubles: alpha[state_cap], x[state_cap], amp, 1/(2 l^2),
 * mean, m; state_cap >= 32) hold the prefit reference GP (e.g. a row of
 * lcf_batch_nonparametric) or m = 0 to fit it on the device (the row is then
 * filled).  out_status: 0 None, 1 result without fit (fewer than 3 colours),
 * 2 fitted; out [n_src][8]: log_temp_peak, log_temp_latest, cooling_rate,
 * log_temp_peak_err, cooling_rate_err, chi2, best cost, n_color_obs;
 * out_col_used[i] = 1 for the colours used.  The two particle-swarm runs use
 * Philox4x32-10 keyed by seed. */
int lcf_batch_thermal(lcf_ctx* ctx, int64_t n_src, const int64_t* ref_offsets, const double* ref_times,
                      const double* ref_mags, const double* ref_errs, const int64_t* col_offsets,
                      const double* col_times, const double* col_mags, const double* col_errs,
                      const double* col_lambda, const double* lambda_ref, const double* duration,
                      double* gp_state, int state_cap, uint64_t seed, int32_t* out_status, double* out,
                      int32_t* out_col_used);

/* fit_bb_temperature (src/gp2d.rs:275-324) for n_sets SEDs (CSR off:
 * wavelengths in Angstrom, magnitudes, errors): the 200-point log10 T grid
 * on [3, 5.5] with the weighted least-squares offset, chi2 / max(n - 2, 1);
 * out_ok[s] = 0 (outputs NaN) for a set of fewer than 2 points.  Used on the
 * 2D GP predictions by extract_thermal_from_gp2d (src/gp2d.rs:445-476) and
 * fit_gp_2d_batch_gpu (src/gp2d.rs:516-724). */
int lcf_batch_bb_temperature(lcf_ctx* ctx, int64_t n_sets, const int64_t* offsets, const double* wavelengths,
                             const double* mags, const double* errs, double* out_log_t, double* out_chi2,
                             double* out_offset, int32_t* out_ok);

/* ---- batched SVI ----------------------------------------------------------- */
/* SVI flags */
#define LCF_SVI_CPU_STOPPING 1u  /* early stop after 200 steps / stall 50 (svi_fit constants,
                                    src/parametric.rs:2262-2263) instead of 50 / 20 */
#define LCF_SVI_INFLATE_SIGMA 2u /* add ln 4 to the returned log sigma (src/parametric.rs:2441-2446) */
#define LCF_SVI_FD_GRADIENT 4u   /* numerical (forward-difference) model gradients instead of the
                                    analytic adjoint: the plain comparison path */
#define LCF_SVI_FIXED_STEPS 8u   /* run all n_steps Adam steps (no early stop) */
#define LCF_SVI_SHARED_STREAM 32u /* every job draws job 0's noise, so a job's fit does not depend on
                                     its position in the batch (the CPU fit's fresh SmallRng(42) per
                                     band, src/parametric.rs:2239) */
#define LCF_SVI_EXACT_TAIL 16u   /* upper-limit adjoint = the exact derivative of the log_normal_cdf
                                    the ELBO sums, instead of the reference's phi / max(Phi, 1e-300)
                                    with Phi from erf_approx (cuda/svi.cu:200-204), which reaches
                                    1e280 in the lower tail and freezes Adam */

/* GpuContext::batch_svi_fit (src/gpu_cuda.rs:1315-1422): one job per source
 * of `data` (the SviBatchInput list is parallel to the batch).  Rows of
 * pso_params / prior_centers / prior_widths / out_mu / out_log_sigma are
 * padded to LCF_MAX_PARAMS; se_idx[j] must be n_params(model)-1 as produced by
 * svi_model_meta (src/parametric.rs:3277-3290).  Seed 42. */
int lcf_batch_svi_fit(lcf_ctx* ctx, const lcf_batch* data, const int32_t* model_ids,
                      const double* pso_params, const int32_t* se_idx,
                      const double* prior_centers, const double* prior_widths, int n_steps,
                      int n_samples, double lr, double* out_mu, double* out_log_sigma,
                      double* out_elbo);
/* General form: explicit seed, LCF_SVI_* flags, steps executed per job. */
int lcf_batch_svi_fit_ex(lcf_ctx* ctx, const lcf_batch* data, const int32_t* model_ids,
                         const double* pso_params, const double* prior_centers,
                         const double* prior_widths, int n_steps, int n_samples, double lr,
                         uint64_t seed, uint32_t flags, double* out_mu, double* out_log_sigma,
                         double* out_elbo, int32_t* out_steps);
/* One Monte-Carlo ELBO estimate per job at (mu, log_sigma) with the noise of
 * step `step`, and its exact gradient (the adjoint kernel; LCF_SVI_FD_GRADIENT
 * selects forward differences).  No counterpart entry point in the reference:
 * it exposes one step of batch_svi_fit (cuda/svi.cu:140-244) for gradient
 * checks and autograd bindings. */
int lcf_svi_elbo_grad(lcf_ctx* ctx, const lcf_batch* data, const int32_t* model_ids,
                      const double* mu, const double* log_sigma, const double* prior_centers,
                      const double* prior_widths, int n_samples, uint64_t seed, int step,
                      uint32_t flags, double* out_elbo, double* out_grad_mu,
                      double* out_grad_log_sigma);

/* ---- batched Laplace approximation (FP64) ---------------------------------- */
#define LCF_LAPLACE_FD4 0      /* reference: 4-point central differences of the objective */
#define LCF_LAPLACE_GRAD_FD 1  /* forward over adjoint: central differences of the analytic gradient */

/* laplace_fit (src/parametric.rs:2623-2693) for every job of `data` (one
 * band per job, model_ids[j], MAP point map_params[j][7]): log_sigma[j][7]
 * (shape parameters from the Hessian of the negative log posterior with
 * sigma_x fixed; sigma_x's slot is ln of its prior width), elbo[j] = -f(MAP),
 * hessian[j][36] (row-major (np-1)^2 block, may be NULL).  The reference runs
 * this per band on the CPU inside fit_parametric; batched here. */
int lcf_batch_laplace(lcf_ctx* ctx, const lcf_batch* data, const int32_t* model_ids,
                      const double* map_params, int mode, double* out_log_sigma, double* out_elbo,
                      double* out_hessian);

/* ---- CPU-path swarm fits (FP64) ------------------------------------------- */
/* The fits of the reference's CPU path, which fit_parametric runs per band
 * (src/parametric.rs:1160-1849, 3813-4131): one job is one
 *   LCF_CPSO_SWARM_BAND   pso_minimize / pso_minimize_seeded on
 *                         PsoCost::cost_from_slice of band `band` and model
 *                         `model` (population priors when use_prior; t0 held
 *                         at t0_value and left out of the search vector when
 *                         fixed_t0, as profile_t0_search does);
 *   LCF_CPSO_SWARM_JOINT  pso_minimize[_seeded] on fit_parametric_multiband's
 *                         joint Villar cost over bands band .. band+n_bands-1
 *                         of `data` (reference band first; dim = 7 n_bands,
 *                         n_bands <= 9);
 *   LCF_CPSO_SWARM_MULTIBAZIN  pso_minimize[_seeded] on multi_bazin_cost
 *                         (src/parametric.rs:983-1006) of band `band` with
 *                         K = `model` components (dim = 4 K + 2), the swarms
 *                         of fit_multi_bazin (:1008-1157);
 *   LCF_CPSO_POLISH       lbfgs_refine from x0 (start_cost = its cost; 7 pairs,
 *                         max_iters iterations, gradient norm 1e-6,
 *                         backtracking Armijo line search), kept only if it
 *                         lowers the cost.
 * The swarm is the reference's: inertia 0.9 -> 0.4, c1 = c2 = 1.5, velocity
 * clamp and wall absorption, global best updated after every particle,
 * relative-gain stall rule; particle 0 at x0 when seeded.  Random numbers:
 * Philox4x32-10 keyed by `seed`, counter word `stream` (the reference's
 * SmallRng(seed) stream is not reproduced).  lo / hi / x0 hold `dim`
 * entries (dim <= 64, n_particles <= 256, at most 4096 observations per
 * job).  Outputs: out_x[n_jobs][64] (best position), out_cost, out_iters
 * (iterations executed). */
#define LCF_CPSO_SWARM_BAND 0
#define LCF_CPSO_SWARM_JOINT 1
#define LCF_CPSO_POLISH 2
#define LCF_CPSO_SWARM_MULTIBAZIN 3
typedef struct lcf_pso_job {
    int32_t kind, model, band, n_bands, dim, n_particles, max_iters, stall_iters, use_prior, fixed_t0, seeded;
    uint32_t stream;
    uint64_t seed;
    double t0_value, start_cost;
    double lo[64], hi[64], x0[64];
} lcf_pso_job;
int lcf_cpu_pso(lcf_ctx* ctx, const lcf_batch* data, const lcf_pso_job* jobs, int64_t n_jobs, double* out_x,
                double* out_cost, int32_t* out_iters);

/* ---- batched finalize: profile refinement + magnitude chi2 (FP64) ------------ */
#define LCF_FINALIZE_T0 1        /* profile_t0_refine */
#define LCF_FINALIZE_ALL 2       /* profile_refine_all */
#define LCF_FINALIZE_MAG_CHI2 4  /* magnitude-space chi2 */
#define LCF_FINALIZE_DEFAULT 7
#define LCF_FINALIZE_SVI_CHI2 8  /* magnitude chi2 from mu only, no PSO fallback (fit_parametric_multiband, :4069-4085) */

/* The per-band post-processing that finalize_parametric_with_gpu_svi
 * (src/parametric.rs:3298-3421), finalize_all_models_with_gpu_svi
 * (:3429-3553) and finalize_parametric_from_gpu (:3126-3253) run on the host
 * after the GPU fits, batched on the device: for band j of `data` (the
 * BandFitData arrays: normalised flux, obs_var, upper flags / fluxes) with
 * model_ids[j], PSO parameters pso_params[j][7] and posterior mu[j][7],
 * log_sigma[j][7]:
 *   LCF_FINALIZE_T0   profile_t0_refine (:2788-3008): 25-point t0 grid on
 *                     [t_first - 30, t_peak + 5] with the amplitude (Bazin,
 *                     TDE: and baseline) refitted by weighted least squares;
 *                     best t0 -> out_mu, Delta LL = 0.5 width -> log sigma;
 *   LCF_FINALIZE_ALL  profile_refine_all (:2705-2786): 15-point scans over
 *                     the CPU PSO bounds, sigma = max(SVI sigma, profile
 *                     half-width, population-prior width);
 *   LCF_FINALIZE_MAG_CHI2  out_mag_chi2[j] (:3366-3393) from flux[],
 *                     flux_err[] (normalised, per observation, the batch's
 *                     order), mag_obs[] = -2.5 log10(raw flux) + 23.9 and
 *                     peak_flux[j]; NaN when no term qualifies.
 * Empty bands pass mu / log_sigma through with a NaN chi2.  Bands are limited
 * to 4096 observations. */
int lcf_batch_finalize(lcf_ctx* ctx, const lcf_batch* data, const int32_t* model_ids,
                       const double* pso_params, const double* mu, const double* log_sigma,
                       const double* flux, const double* flux_err, const double* mag_obs,
                       const double* peak_flux, int flags, double* out_mu, double* out_log_sigma,
                       double* out_mag_chi2);

/* ---- batched dense GP (FP64) -------------------------------------------------- */
/* GpuContext::batch_gp_fit (src/gpu_cuda.rs:976-1143): bands given as CSR
 * (times, mags, noise_var); training set = at most min(max_subsample, 25)
 * points, every (amp, ls) candidate with ls >= 0.1 scored by the exact
 * negative log marginal likelihood; outputs the reference's 54-double state
 * per band (alpha[25], x[25], amp, 1/(2 l^2), y mean, m), mean and std at the
 * n_query shared query times, and the mean at every observation. */
int lcf_batch_gp_fit(lcf_ctx* ctx, int64_t n_bands, const int64_t* offsets, const double* times,
                     const double* mags, const double* noise_var, const double* query_times,
                     int n_query, const double* amp_candidates, int n_amp,
                     const double* ls_candidates, int n_ls, int max_subsample, double* out_state,
                     double* out_pred, double* out_std, double* out_pred_obs);
/* General form: training-set cap up to 200 (LDS-resident factorisation),
 * subsample rule (0: cuda/gp.cu:189-191, 1: src/gp.rs:15-19), criterion
 * (0: NLML, 1: train RMS, src/nonparametric.rs:1318-1335), query grid shared
 * (query_per_band = 0, query[n_query]) or per band (query[n_bands][n_query]),
 * state rows of 2*cap+4 doubles (alpha[cap], x[cap], amp, 1/(2 l^2), mean, m)
 * and the time derivative of the predictive mean (out_dpred may be NULL). */
int lcf_batch_gp_fit_ex(lcf_ctx* ctx, int64_t n_bands, const int64_t* offsets, const double* times,
                        const double* mags, const double* noise_var, const double* query,
                        int n_query, int query_per_band, const double* amp_candidates, int n_amp,
                        const double* ls_candidates, int n_ls, int cap, int subsample_rule,
                        int criterion, double* out_state, double* out_pred, double* out_std,
                        double* out_dpred, double* out_pred_obs);

/* GpuContext::batch_gp_2d (src/gpu_cuda.rs:1152-1306): per source, an
 * anisotropic squared-exponential GP over (time, log10 wavelength); training
 * set = at most min(max_subsample, 40) points; every (amp, lst, lsw)
 * candidate with lst >= 0.1 and lsw >= 0.001 scored by the exact NLML;
 * outputs the reference's 125-double state (alpha[40], t[40], w[40], amp,
 * 1/(2 lt^2), 1/(2 lw^2), y mean, m), mean / std at the (qt[k], qw[k]) query
 * points, and the train RMS over at most 100 observations (NaN on failure). */
int lcf_batch_gp2d_fit(lcf_ctx* ctx, int64_t n_sources, const int64_t* offsets, const double* times,
                       const double* log_waves, const double* mags, const double* noise_var,
                       const double* query_times, const double* query_waves, int n_query,
                       const double* amp_candidates, int n_amp, const double* lst_candidates,
                       int n_lst, const double* lsw_candidates, int n_lsw, int max_subsample,
                       double* out_state, double* out_pred, double* out_std, double* out_train_rms);
/* General 2D form, as lcf_batch_gp_fit_ex: cap up to 200, state rows of
 * 3*cap+5 doubles (alpha[cap], t[cap], w[cap], amp, 1/(2 lt^2), 1/(2 lw^2),
 * mean, m), per-source query grids, d(mean)/dt, mean at every observation. */
int lcf_batch_gp2d_fit_ex(lcf_ctx* ctx, int64_t n_sources, const int64_t* offsets, const double* times,
                          const double* log_waves, const double* mags, const double* noise_var,
                          const double* query_times, const double* query_waves, int n_query,
                          int query_per_source, const double* amp_candidates, int n_amp,
                          const double* lst_candidates, int n_lst, const double* lsw_candidates,
                          int n_lsw, int cap, int subsample_rule, int criterion, double* out_state,
                          double* out_pred, double* out_std, double* out_dpred, double* out_pred_obs);

/* fit_gp_2d (src/gp2d.rs:336-439), the CPU path's 2D GP, batched: per
 * source (concatenated bands: times, log10 wavelengths, mags, noise
 * variances) a training subsample of at most 200 points at floor(i n / 200),
 * the per-source grid amp {0.3, 1, 3, 10} x lt {dur/4, dur/8, dur/16} x
 * lw {2 wr, wr, wr/2} (dur, wr: time and log-wavelength ranges of the
 * subsample, floored at 1 d and 0.01 dex), selection by the first strict
 * minimum of the train RMS, then predict_with_std (:181-209) at the source's
 * queries (query_t / query_w [n][n_query]).  out_hyp[n][3] = amp, lt, lw,
 * out_train_rms[n], out_n_train[n] (0 and NaN when no candidate factorises or
 * the source has fewer than 3 points). */
int lcf_batch_fit_gp_2d(lcf_ctx* ctx, int64_t n_sources, const int64_t* offsets, const double* times,
                        const double* log_waves, const double* mags, const double* noise_var,
                        const double* query_t, const double* query_w, int n_query, double* out_hyp,
                        double* out_train_rms, int32_t* out_n_train, double* out_pred, double* out_std);

/* ---- 2D GP hyper-parameter fit with the Cholesky adjoint (FP64) -------------
 * The optimised counterpart of the hyper-parameter grid of batch_gp_2d /
 * fit_gp_2d (src/gpu_cuda.rs:1152-1306, src/gp2d.rs:336-439, 516-724): per
 * source (CSR offsets, times, log10 wavelengths, magnitudes, noise variances,
 * at most 2047 points) the separable squared-exponential GP
 *   K = a exp(-dt^2 / (2 lt^2) - dw^2 / (2 lw^2)) + diag(noise_var + sn^2)
 * on y - mean(y), theta = ln (a, lt, lw, sn) ([n][4]).  The negative log
 * marginal likelihood and its exact gradient 1/2 tr((K^-1 - alpha alpha^T)
 * dK/dtheta) come from a blocked FP64 Cholesky and its adjoint; L-BFGS
 * (6 pairs over the free components -- a component whose gradient presses
 * against its bound is held for the iteration, and a step the projection
 * leaves without decrease falls back to free steepest descent -- projected
 * backtracking Armijo search that gives up once a step moves theta by
 * <= 1e-10, box a in [1e-6, 1e6],
 * lt in [0.1, 1e4], lw in [1e-3, 10], sn in [1e-6, 10]) runs at most
 * max_iters iterations from theta0, stopping when the projected gradient's
 * largest component is <= grad_tol.  out_status: 0 converged, 1 iteration
 * limit, 2 no further decrease along the search direction, 3 K not positive
 * definite at theta0 (or no points).  Optional queries (n_query > 0; shared,
 * or [n][n_query] when query_per_source) give the posterior mean (+ mean(y))
 * and standard deviation at the returned theta.  out_evals ([n][2], may be
 * NULL) counts the NLML and gradient evaluations of each source. */
int lcf_batch_gp2d_train(lcf_ctx* ctx, int64_t n_sources, const int64_t* offsets, const double* times,
                         const double* log_waves, const double* mags, const double* noise_var,
                         const double* theta0, int max_iters, double grad_tol, const double* query_times,
                         const double* query_waves, int n_query, int query_per_source, double* out_theta,
                         double* out_nlml, double* out_grad, int32_t* out_iters, int32_t* out_status,
                         double* out_pred, double* out_std, int32_t* out_evals);

/* General form: flags LCF_GP2D_FD_GRADIENT runs the same L-BFGS on central
 * finite differences of the NLML (h = 1e-5 max(|theta_k|, 1), eight
 * factorisations per gradient; the plain form the adjoint replaces; out_grad
 * is still the exact gradient at the returned point), and the context's
 * schedule applies (LCF_SCHED_PER_CLASS: one synchronised launch per
 * power-of-two size class).  out_evals counts every factorisation. */
#define LCF_GP2D_FD_GRADIENT 1u
int lcf_batch_gp2d_train_ex(lcf_ctx* ctx, int64_t n_sources, const int64_t* offsets, const double* times,
                            const double* log_waves, const double* mags, const double* noise_var,
                            const double* theta0, int max_iters, double grad_tol, const double* query_times,
                            const double* query_waves, int n_query, int query_per_source, uint32_t flags,
                            double* out_theta, double* out_nlml, double* out_grad, int32_t* out_iters,
                            int32_t* out_status, double* out_pred, double* out_std, int32_t* out_evals);

/* The same objective and gradient at theta ([n][4], not clamped): one
 * evaluation per source; out_status 0, or 3 when K is not positive definite
 * (NaN outputs).  Backs the autograd binding and the finite-difference check. */
int lcf_gp2d_nlml_grad(lcf_ctx* ctx, int64_t n_sources, const int64_t* offsets, const double* times,
                       const double* log_waves, const double* mags, const double* noise_var,
                       const double* theta, double* out_nlml, double* out_grad, int32_t* out_status);

/* Timing of the last lcf_batch_fit_models-family call on this context:
 * milliseconds between HIP events around the kernel launches (inputs already
 * resident, outputs still on the device). */
double lcf_last_kernel_ms(const lcf_ctx* ctx);
/* Milliseconds of the launch(es) of model slot `slot` in that call (when the
 * models ran as one mixed-family launch: slot 0 holds that launch, the other
 * slots 0). */
double lcf_last_model_ms(const lcf_ctx* ctx, int slot);

#ifdef __cplusplus
}
#endif
#endif
