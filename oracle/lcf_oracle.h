/*
 * lcf_oracle.h -- CPU FP64 restatement of the lightcurve-fitting batched
 * fitting path.  TEST INFRASTRUCTURE ONLY: this library is the parity checker
 * for the HIP product (tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg load it).  Nothing in the product links or calls it.
 *
 * Every function cites the reference file:line whose semantics it restates
 * (paths relative to the reference repository root).
 *
 * Conventions
 *   - model ids follow GpuModelName::model_id (src/gpu_types.rs:51-62):
 *     0 Bazin, 1 Villar, 2 Tde, 3 Arnett, 4 Magnetar, 5 ShockCooling,
 *     6 Afterglow, 7 MetzgerKN.
 *   - a "batch" is the CSR layout of GpuBatchData::new (src/gpu_cuda.rs:251-285):
 *     offsets[n_src+1] (int64 here), times/flux/obs_var/upper_flux (f64),
 *     is_upper (i32).
 *   - return value 0 = ok, negative = invalid argument.
 */
#ifndef LCF_ORACLE_H
#define LCF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- model metadata ---------------------------------------------------- */
int orc_n_params(int model_id);
/* GPU PSO box bounds, GpuModelName::pso_bounds (src/gpu_types.rs:64-103). */
int orc_pso_bounds(int model_id, double* lower, double* upper);
/* CPU PSO box bounds, pso_bounds (src/parametric.rs:1541-1579). */
int orc_pso_bounds_cpu(int model_id, double* lower, double* upper);
/* population_priors (src/parametric.rs:1588-1621): centers/widths[np]. */
int orc_population_priors(int model_id, double* centers, double* widths);
/* prior_params (src/parametric.rs:2072-2115). */
int orc_prior_params(int model_id, double* centers, double* widths);

/* ---- model evaluation ---------------------------------------------------- */
/* eval_model / MetzgerKN batch (src/parametric.rs:512-538, 544-648); this is
 * also GpuContext::eval_batch semantics (cuda/models.cu:146-282, 387-407).
 * out[d*n_t+i] = f(params[d*np..], times[i]). */
int orc_eval_batch(int model_id, const double* params, int64_t n_draws,
                   const double* times, int64_t n_times, double* out);
/* Per-observation MetzgerKN used by the GPU cost kernel
 * (cuda/models.cu:429-514): 100-point grid re-solved per time. */
double orc_metzger_kn_single(const double* p, double t);
/* eval_model_grad (src/parametric.rs:527-538): out[np] = d f / d p.
 * Returns -1 for MetzgerKN (no analytic gradient). */
int orc_eval_grad(int model_id, const double* p, double t, double* out);

/* ---- likelihood helpers ------------------------------------------------- */
double orc_log_normal_cdf(double x);   /* src/parametric.rs:882-901 */
double orc_dlog_normal_cdf(double x);  /* its exact derivative (the consistent SVI tail adjoint) */
double orc_erf_approx(double x);       /* src/parametric.rs:867-880 */

/* ---- PSO cost -------------------------------------------------------------- */
/* GPU-path cost: kernel batch_pso_cost (cuda/models.cu:516-584) plus the
 * host-side population prior / n^2 (src/gpu_cuda.rs:461-480). */
double orc_pso_cost_gpu(int model_id, const double* p, const double* t,
                        const double* f, const double* v, const int32_t* up,
                        const double* uf, int64_t n);
/* CPU-path cost: PsoCost::cost_from_slice (src/parametric.rs:1303-1370),
 * with the Villar constraint and MetzgerKN rescale. */
double orc_pso_cost_cpu(int model_id, const double* p, const double* t,
                        const double* f, const double* v, const int32_t* up,
                        const double* uf, int64_t n);
/* PsoCost::gradient (src/parametric.rs:1383-1446). */
int orc_pso_cost_grad_cpu(int model_id, const double* p, const double* t,
                          const double* f, const double* v, const int32_t* up,
                          const double* uf, int64_t n, double* grad);

/* ---- counter-based RNG ---------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- batched PSO (GPU semantics) ------------------------------------------ */
/* flags for orc_batch_pso / orc_batch_model_select */
#define ORC_PSO_STALL_FROM_FIRST_EVAL 1u  /* opt-in: stall baseline = first gbest */

/* GpuContext::batch_pso (src/gpu_cuda.rs:368-567): synchronous swarm update,
 * Philox stream keyed by (seed, source index) in place of SmallRng.
 * out_params[n_src*np], out_cost[n_src], out_iters[n_src] (may be NULL). */
int orc_batch_pso(int model_id, int64_t n_src, const int64_t* offsets,
                  const double* t, const double* f, const double* v,
                  const int32_t* up, const double* uf,
                  int n_particles, int max_iters, int stall_iters,
                  uint64_t seed, uint32_t flags,
                  double* out_params, double* out_cost, int32_t* out_iters,
                  int n_threads);

/* GpuContext::batch_model_select (src/gpu_cuda.rs:573-620) over an explicit
 * model list (models[0] is the gate model); seed 42 as the reference.
 * out_model[n_src], out_params[n_src*7] (padded with 0), out_cost[n_src]. */
int orc_batch_model_select(int64_t n_src, const int64_t* offsets,
                           const double* t, const double* f, const double* v,
                           const int32_t* up, const double* uf,
                           const int32_t* models, int n_models,
                           int n_particles, int max_iters, int stall_iters,
                           double gate, uint32_t flags,
                           int32_t* out_model, double* out_params,
                           double* out_cost, int n_threads);

/* GpuContext::batch_pso_multi_bazin (src/gpu_cuda.rs:650-963), Philox
 * stream; outputs per source best_k, params[18] (zero padded), cost, bic,
 * per_k_cost[4], per_k_bic[4] (NaN after the early stop). */
/* batch_pso_cost_multi_bazin (cuda/models.cu:319-367) for one parameter
 * vector with k components. */
double orc_multi_bazin_cost(const double* p, int k, const double* t, const double* f, const double* v,
                            const int32_t* up, const double* uf, int64_t n);
int orc_batch_multi_bazin(int64_t n_src, const int64_t* offsets, const double* t, const double* f,
                          const double* v, const int32_t* up, const double* uf, int n_particles,
                          int max_iters, int stall_iters, uint64_t seed, int32_t* best_k,
                          double* params, double* cost, double* bic, double* per_k_cost,
                          double* per_k_bic, int n_threads);

/* ---- SVI (GPU semantics, cuda/svi.cu:48-330) ------------------------------ */
#define ORC_SVI_RNG_PHILOX 0     /* stream of the HIP kernels */
#define ORC_SVI_RNG_REFERENCE 1  /* reference xorshift64, seeded 42 + job*123456789 */
#define ORC_SVI_CPU_STOPPING 1u  /* flag: svi_fit constants, min 200 / stall 50 */
#define ORC_SVI_INFLATE_SIGMA 2u /* flag: add ln 4 to log sigma (src/parametric.rs:2441-2446) */
#define ORC_SVI_FIXED_STEPS 8u   /* flag: no early stop, all n_steps steps */
#define ORC_SVI_EXACT_TAIL 16u   /* flag: upper-limit adjoint = exact derivative of log_normal_cdf */
#define ORC_SVI_SHARED_STREAM 32u /* flag: every job draws job 0's Philox noise */

/* Philox normals of (job, step, sample): eps[np]. */
int orc_svi_normals(uint64_t seed, uint32_t job, uint32_t step, uint32_t sample, int np, double* eps);
/* One step's Monte-Carlo ELBO for noise eps[S][np] and its full gradient,
 * entropy term included (grad_mode 0 analytic model gradient, 1 forward
 * differences). */
double orc_svi_elbo_grad(int model_id, const double* mu, const double* log_sigma,
                         const double* prior_c, const double* prior_w, const double* eps, int S,
                         const double* t, const double* f, const double* v, const int32_t* up,
                         const double* uf, int64_t n, int grad_mode, double* grad_mu,
                         double* grad_log_sigma);
/* GpuContext::batch_svi_fit (src/gpu_cuda.rs:1315-1422).  Per job: model id,
 * pso[7], prior centers/widths [7] (padded rows).  Outputs mu/log_sigma
 * [n_jobs*7], elbo[n_jobs], steps executed, and (diagnostic, may be NULL) the
 * largest |gradient| handed to Adam over the fit, +inf if one was non-finite. */
int orc_batch_svi(int64_t n_jobs, const int64_t* offsets, const double* t, const double* f,
                  const double* v, const int32_t* up, const double* uf, const int32_t* model_ids,
                  const double* pso, const double* prior_c, const double* prior_w, int n_steps,
                  int n_samples, double lr, int rng_mode, uint64_t seed, int grad_mode,
                  uint32_t flags, double* out_mu, double* out_log_sigma, double* out_elbo,
                  int32_t* out_steps, double* out_gmax, int n_threads);
/* svi_prior_for_model (src/parametric.rs:3259-3273). */
int orc_svi_prior_for_model(int model_id, const double* pso, double* centers, double* widths);

/* ---- dense 1D GP (src/sparse_gp.rs:72-176, cuda/gp.cu:124-367) ----------- */
#define ORC_GP_SUBSAMPLE_GPU 0   /* idx = floor(i (n-1)/(m-1) + 1/2) (cuda/gp.cu:189-191) */
#define ORC_GP_SUBSAMPLE_CPU 1   /* idx = floor((i + 1/2) n / m)     (src/gp.rs:15-19) */
#define ORC_GP_SELECT_NLML 0     /* exact NLML (GPU path) */
#define ORC_GP_SELECT_TRAIN_RMS 1 /* train RMS (CPU dense path) */

/* DenseGP::fit + predict_with_std + d(mean)/dt + NLML + train_rms.
 * Returns 1 when the Cholesky fails. Output pointers may be NULL. */
int orc_dense_gp(const double* t, const double* y, const double* nv, int n, double amp, double ls,
                 const double* q, int nq, double* mean, double* stdv, double* dmean, double* nlml,
                 double* train_rms);
/* GpuContext::batch_gp_fit (src/gpu_cuda.rs:976-1143) with the subsample cap
 * max_sub (the reference's is min(max_subsample, 25)).  state rows (stride
 * >= 2*max_sub+4): alpha[max_sub], x[max_sub], amp, 1/(2 l^2), ybar, m.
 * pred/std/dpred [n_bands*nq], pred_obs [n_obs] (NaN where no fit). */
int orc_batch_gp_fit(int64_t n_bands, const int64_t* offsets, const double* t, const double* y,
                     const double* noise_var, const double* q, int nq, const double* amps, int n_amp,
                     const double* lss, int n_ls, int max_sub, int subsample_rule, int criterion,
                     double* state, int state_stride, double* pred, double* stdv, double* dpred,
                     double* pred_obs, int n_threads);

/* GpuContext::batch_gp_2d (src/gpu_cuda.rs:1152-1306) with cap max_sub
 * (the reference's is min(max_subsample, 40)); state rows (stride >=
 * 3*max_sub+5): alpha, t, w [max_sub each], amp, 1/(2lt^2), 1/(2lw^2), ybar, m. */
int orc_batch_gp2d_fit(int64_t n_src, const int64_t* offsets, const double* t, const double* w,
                       const double* y, const double* noise_var, const double* qt, const double* qw,
                       int nq, const double* amps, int n_amp, const double* lst, int n_lst,
                       const double* lsw, int n_lsw, int max_sub, int subsample_rule, int criterion,
                       double* state, int state_stride, double* pred, double* stdv, double* dpred,
                       double* pred_obs, int n_threads);

/* ---- nonparametric band pipeline (src/nonparametric.rs:1228-1525) ------- */
#define ORC_NP_NFEAT 50  /* NonparametricBandResult fields (src/nonparametric.rs:74-160), band name excluded */
/* process_band per CSR band of raw magnitudes (times, mags, mag errors),
 * grid of n_grid points on [t_lo[b], t_hi[b]] (duration t_hi - t_lo), dense
 * GP when detections <= dense_max, else FITC with n_inducing points and a
 * dense GP on max_sub subsampled detections for thermal reuse.  status[b]:
 * 0 no result, 1 dense, 2 FITC.  feat [n_bands][ORC_NP_NFEAT] (NaN = None),
 * pred / std (rescaled) / d(mean)/dt [n_bands][n_grid], state rows (stride
 * >= 2 cap + 4, cap >= max(dense_max, max_sub)): alpha[cap], x[cap], amp,
 * 1/(2 l^2), y mean, m of the GP kept for thermal reuse. */
int orc_batch_nonparametric(int64_t n_bands, const int64_t* offsets, const double* times, const double* mags,
                            const double* mag_errs, const double* t_lo, const double* t_hi, int n_grid,
                            const double* amps, int n_amp, const double* ls_factors, int n_ls, int dense_max,
                            int max_sub, int n_inducing, int32_t* status, double* feat, double* pred,
                            double* stdv, double* dpred, double* state, int state_stride, int n_threads);
/* compute_fine_grid_metrics' eleven metrics of a curve with a given peak
 * index, in FineGridMetrics order (src/nonparametric.rs:490-554). */
int orc_np_curve_metrics(const double* t, const double* m, int n, int peak_idx, double* out11);
/* von Neumann ratio, pre-peak rms, rise amplitude / noise, Stetson K, excess
 * variance, CUSUM range (src/nonparametric.rs:586-724). */
int orc_np_raw_features(const double* t, const double* m, const double* e, int n, double t_peak,
                        double peak_mag, double* out6);
/* Grid statistics of compute_predictive_features (sigma_f, peak_to_peak,
 * dfdt_max, dfdt_min, frac_of_peak, post_var_mean, post_var_max, skewness,
 * kurtosis, n_inflections), then power-law index and chi2, post-peak
 * monotonicity, local maxima (src/nonparametric.rs:266-472, 730-774). */
int orc_np_grid_features(const double* tg, const double* p, const double* sd, int n, int peak_idx,
                         double* out14);
/* SparseGP (FITC) fit with m inducing points + predict_with_std + approx_nlml
 * (src/sparse_gp.rs:183-328).  Returns 1 when a factorisation fails. */
int orc_sparse_gp(const double* t, const double* y, const double* nv, int n, double amp, double ls, int m,
                  const double* q, int nq, double* mean, double* stdv, double* nlml);

/* ---- blackbody colour-temperature fit (src/thermal.rs:222-418) ---------- */
/* bb_color_mag (src/thermal.rs:75-92). */
double orc_bb_color_mag(double lambda_ref, double lambda, double temp);
/* fit_thermal per source after the host's band choice: reference-band
 * observations (CSR ref_off) for fit_ref_band_gp, the other bands'
 * observations with their wavelengths (CSR col_off), lambda_ref, duration,
 * gp_state rows (alpha[cap], x[cap], amp, 1/(2 l^2), mean, m; m = 0: fit the
 * reference GP here, the row is then filled).  status: 0 None, 1 result
 * without a fit (fewer than 3 colours), 2 fitted; out[8]: log_temp_peak,
 * log_temp_latest, cooling_rate, log_temp_peak_err, cooling_rate_err, chi2,
 * best cost, n_color_obs; col_used[i] = 1 for colours kept. */
int orc_batch_thermal(int64_t n_src, const int64_t* ref_off, const double* ref_t, const double* ref_m,
                      const double* ref_e, const int64_t* col_off, const double* col_t, const double* col_m,
                      const double* col_e, const double* col_lambda, const double* lambda_ref,
                      const double* duration, double* gp_state, int state_stride, uint64_t seed, int32_t* status,
                      double* out, int32_t* col_used, int n_threads);

/* fit_bb_temperature (src/gp2d.rs:275-324): out3 = log10 T, chi2, offset;
 * returns 1 (no fit) for fewer than 2 points. */
int orc_fit_bb_temperature(const double* lam, const double* mag, const double* err, int n, double* out3);

/* ---- Laplace (src/parametric.rs:2464-2693) ---------------------------------- */
#define ORC_LAPLACE_FD4 0      /* reference: 4-point central differences of f */
#define ORC_LAPLACE_GRAD_FD 1  /* central differences of the analytic gradient */
double orc_neg_log_posterior(int model_id, const double* p, const double* t, const double* f,
                             const double* v, const int32_t* up, const double* uf, int64_t n);
int orc_symmetric_eigen(const double* mat, int n, double* evals, double* evecs);
/* log_sigma[np] (shape params from the Hessian, sigma_x = ln prior width),
 * elbo = -f(MAP), hessian[(np-1)^2] (may be NULL). */
int orc_laplace_fit(int model_id, const double* map_params, const double* t, const double* f,
                    const double* v, const int32_t* up, const double* uf, int64_t n, int mode,
                    double* log_sigma, double* elbo, double* hessian);

/* ---- finalize (src/parametric.rs:2705-3008, 3298-3421) ----------------------
 * Per band (CSR offsets, BandFitData arrays: normalised flux / flux_err, obs
 * variance, upper-limit flags and fluxes, observed magnitudes, band peak flux)
 * and model: profile_t0_refine, then profile_refine_all, then the magnitude
 * chi2; mu / log_sigma / pso are [n_bands][7].  Empty bands pass through with
 * a NaN chi2. */
#define ORC_FINALIZE_T0 1
#define ORC_FINALIZE_ALL 2
#define ORC_FINALIZE_MAG_CHI2 4
#define ORC_FINALIZE_SVI_CHI2 8  /* chi2 from mu only (no PSO fallback) */
int orc_finalize(int64_t n_bands, const int64_t* off, const double* t, const double* f, const double* v,
                 const int32_t* up, const double* uf, const double* fe, const double* mag_obs,
                 const double* peak, const int32_t* model_ids, const double* pso, const double* mu,
                 const double* log_sigma, int flags, double* out_mu, double* out_log_sigma,
                 double* out_mag_chi2, int n_threads);

/* fit_gp_2d (src/gp2d.rs:336-439) per source on concatenated training arrays
 * (times, log10 wavelengths, mags, noise variances): hyp[n][3] = amp, lt, lw,
 * train_rms[n], n_train[n] (0: no fit), predict_with_std at per-source
 * queries qt/qw[n][nq]. */
int orc_batch_fit_gp_2d(int64_t n_src, const int64_t* off, const double* t, const double* w, const double* y,
                        const double* nv, const double* qt, const double* qw, int nq, double* hyp,
                        double* train_rms, int32_t* n_train, double* pred, double* stdv, int n_threads);

/* ---- 2D GP hyper-parameter fit (new; R & W 2.30 / 5.9, L-BFGS) -------------
 * theta = ln (amplitude, time length scale, log10-wavelength length scale,
 * extra noise sigma).  nlml_grad returns 1 when K is not positive definite.
 * train: status 0 gradient below gtol, 1 max_iters reached, 2 line search
 * failed, 3 K not positive definite at theta0 (or empty source); evals[n][2]
 * counts NLML and gradient evaluations (initial, trials, final). */
int orc_gp2d_nlml_grad(const double* t, const double* w, const double* y, const double* nv, int n,
                       const double* theta, double* nlml, double* grad);
int orc_batch_gp2d_train(int64_t n_src, const int64_t* off, const double* t, const double* w, const double* y,
                         const double* nv, const double* theta0, int max_iters, double gtol, double* theta,
                         double* nlml, double* grad, int32_t* iters, int32_t* status, int32_t* evals,
                         int n_threads);

/* ---- CPU-path swarm fits (src/parametric.rs:1160-1849, 3813-4131) ---------
 * One job = one pso_minimize[_seeded] run (kind 0: PsoCost of band `band`,
 * model `model`, priors when use_prior, t0 held at t0_value when fixed_t0;
 * kind 1: the joint Villar cost over bands band .. band + n_bands - 1) or one
 * lbfgs_refine from x0 (kind 2, start_cost = cost at x0).  lo / hi / x0 hold
 * `dim` entries.  Outputs: x[n][64], cost[n], iterations[n]. */
#define ORC_CPSO_SWARM_BAND 0
#define ORC_CPSO_SWARM_JOINT 1
#define ORC_CPSO_POLISH 2
#define ORC_CPSO_SWARM_MULTIBAZIN 3  /* model = K components, dim = 4 K + 2 */
typedef struct orc_pso_job {
    int32_t kind, model, band, n_bands, dim, n_particles, max_iters, stall_iters, use_prior, fixed_t0, seeded;
    uint32_t stream;
    uint64_t seed;
    double t0_value, start_cost;
    double lo[64], hi[64], x0[64];
} orc_pso_job;
int orc_cpu_pso(const int64_t* off, const double* t, const double* flux, const double* obs_var,
                const int32_t* is_upper, const double* upper_flux, int64_t n_obs, const orc_pso_job* jobs,
                int64_t n_jobs, double* out_x, double* out_cost, int32_t* out_iters, int n_threads);

/* ---- fit_parametric (src/parametric.rs:3563-4131) -------------------------
 * The CPU parametric pipeline for n_src sources: source s has bands
 * src_off[s] .. src_off[s+1]-1 of the band CSR band_off over (times, flux
 * values, flux errors), in the order that stands for the band map's
 * iteration order.  method 0 = Svi, 1 = Laplace; forced_model >= 0 runs
 * fit_parametric_model with that model (no joint route).  One record of
 * ORC_FP_RECORD doubles per input band (NaN where unset):
 *   [0] rank of the band's result in its source's list (-1: none), [1] model,
 *   [2..8] pso params, [9] pso chi2, [10..16] posterior mu, [17..23] log
 *   sigma, [24] elbo, [25] n_obs, [26] magnitude chi2, [27..34] per-model
 *   chi2 by model id (NaN = None), [35..42] key present (0/1), [43] route
 *   (0 per band, 1 joint), [44] MultiBazin present, [45] best K, [46..63]
 *   params, [64] cost, [65] BIC, [66..69] per-K cost, [70..73] per-K BIC,
 *   [74..129] per-model params [8][7] (fit_all_models). */
#define ORC_FP_RECORD 130
int orc_fit_parametric(int64_t n_src, const int64_t* src_off, const int64_t* band_off, const double* t,
                       const double* flux, const double* err, int fit_all_models, int method,
                       int forced_model, double* out, int n_threads);

#ifdef __cplusplus
}
#endif
#endif
