//! HIP (MI355X / gfx950) backend of the lightcurve-fitting crate: the
//! module a `hip` feature adds as `src/gpu_hip.rs`, next to `src/gpu_cuda.rs`,
//! re-exported by `src/gpu.rs` like the `cuda` backend (src/gpu.rs:12-18).
//!
//! Every public item keeps the signature of its `src/gpu_cuda.rs`
//! counterpart, so callers compile unchanged:
//!   GpuBatchData::new(&[BatchSource])          src/gpu_cuda.rs:251
//!   GpuBatchData::n_obs_per_source             :289
//!   GpuContext::new                            :309
//!   GpuContext::eval_batch                     :317
//!   GpuContext::batch_pso                      :368
//!   GpuContext::batch_model_select             :573
//!   GpuContext::batch_all_models               :624
//!   GpuContext::batch_pso_multi_bazin          :650
//!   GpuContext::batch_gp_fit                   :976
//!   GpuContext::batch_gp_2d                    :1152
//!   GpuContext::batch_svi_fit                  :1315
//!
//! Each wrapper packs the caller's slices into the flat arrays of one
//! `liblcf_hip.so` call (declared in include/lcf_hip.h) and unpacks the
//! outputs into the reference's result types.  The swarms, GP searches and
//! SVI loops run entirely on the device; nothing is iterated here.
//!
//! Build: link `liblcf_hip.so` (see build_hip in INTEGRATION.md).  The
//! tests/c_driver/lcf_driver.c program makes the same C calls, in the same
//! order and with the same argument packing, and runs in the GPU test suite.

use std::ffi::{c_char, c_int, CStr};

use super::gpu_types::*;
use crate::parametric::MultiBazinResult;

#[repr(C)]
struct LcfCtx {
    _p: [u8; 0],
}
#[repr(C)]
struct LcfBatch {
    _p: [u8; 0],
}

extern "C" {
    fn lcf_ctx_create(device: c_int, out: *mut *mut LcfCtx) -> c_int;
    fn lcf_ctx_destroy(ctx: *mut LcfCtx);
    fn lcf_last_error(ctx: *const LcfCtx) -> *const c_char;
    fn lcf_batch_create_current(n: i64, off: *const i64, t: *const f64, f: *const f64, v: *const f64,
                                up: *const i32, uf: *const f64, out: *mut *mut LcfBatch) -> c_int;
    fn lcf_batch_destroy(b: *mut LcfBatch);
    fn lcf_batch_n_obs(b: *const LcfBatch, idx: i64) -> i64;
    fn lcf_eval_batch(ctx: *mut LcfCtx, model: c_int, p: *const f64, nd: i64, t: *const f64, nt: i64,
                      out: *mut f64) -> c_int;
    fn lcf_batch_pso(ctx: *mut LcfCtx, model: c_int, b: *const LcfBatch, np: c_int, it: c_int,
                     stall: c_int, seed: u64, params: *mut f64, cost: *mut f64) -> c_int;
    fn lcf_batch_model_select(ctx: *mut LcfCtx, b: *const LcfBatch, np: c_int, it: c_int, stall: c_int,
                              gate: f64, model: *mut i32, params: *mut f64, cost: *mut f64) -> c_int;
    fn lcf_batch_all_models(ctx: *mut LcfCtx, b: *const LcfBatch, np: c_int, it: c_int, stall: c_int,
                            params: *mut f64, cost: *mut f64) -> c_int;
    fn lcf_batch_multi_bazin(ctx: *mut LcfCtx, b: *const LcfBatch, np: c_int, it: c_int, stall: c_int,
                             seed: u64, best_k: *mut i32, params: *mut f64, cost: *mut f64, bic: *mut f64,
                             per_k_cost: *mut f64, per_k_bic: *mut f64) -> c_int;
    fn lcf_batch_gp_fit(ctx: *mut LcfCtx, nb: i64, off: *const i64, t: *const f64, m: *const f64,
                        nv: *const f64, q: *const f64, nq: c_int, amp: *const f64, na: c_int,
                        ls: *const f64, nl: c_int, max_sub: c_int, state: *mut f64, pred: *mut f64,
                        std: *mut f64, pred_obs: *mut f64) -> c_int;
    fn lcf_batch_gp2d_fit(ctx: *mut LcfCtx, ns: i64, off: *const i64, t: *const f64, w: *const f64,
                          m: *const f64, nv: *const f64, qt: *const f64, qw: *const f64, nq: c_int,
                          amp: *const f64, na: c_int, lst: *const f64, nlt: c_int, lsw: *const f64,
                          nlw: c_int, max_sub: c_int, state: *mut f64, pred: *mut f64, std: *mut f64,
                          train_rms: *mut f64) -> c_int;
    fn lcf_batch_svi_fit(ctx: *mut LcfCtx, b: *const LcfBatch, model: *const i32, pso: *const f64,
                         se: *const i32, pc: *const f64, pw: *const f64, steps: c_int, samples: c_int,
                         lr: f64, mu: *mut f64, ls: *mut f64, elbo: *mut f64) -> c_int;
}

const MAXP: usize = 7; // LCF_MAX_PARAMS: per-source parameter rows are padded to 7 doubles
const MB_MAXK: usize = 4; // MultiBazin components tried (K = 1..4)
const MB_ROW: usize = 4 * MB_MAXK + 2; // MultiBazin parameter row: 4K + 2 used, zero padded
const GP_STATE: usize = 54; // batch_gp_fit state row: alpha[25], x[25], amp, 1/(2 l^2), mean, m
const GP2D_STATE: usize = 125; // batch_gp_2d state row: alpha[40], t[40], w[40], amp, 1/(2 lt^2), 1/(2 lw^2), mean, m

fn last_error(ctx: *const LcfCtx, rc: c_int) -> String {
    // lcf_last_error(NULL) is the calling thread's last context-free failure.
    let msg = unsafe { CStr::from_ptr(lcf_last_error(ctx)) }.to_string_lossy().into_owned();
    format!("HIP error {rc}: {msg}")
}

/// Device-resident observations of many sources (one packed record per
/// observation plus the CSR offsets), on the device that was current when
/// they were uploaded.
pub struct GpuBatchData {
    b: *mut LcfBatch,
    /// Number of sources.
    pub n_sources: usize,
}

impl Drop for GpuBatchData {
    fn drop(&mut self) {
        unsafe { lcf_batch_destroy(self.b) }
    }
}

impl GpuBatchData {
    /// Upload packed source data to the current HIP device.
    pub fn new(sources: &[BatchSource]) -> Result<Self, String> {
        let mut off: Vec<i64> = Vec::with_capacity(sources.len() + 1);
        off.push(0);
        let (mut t, mut f, mut v, mut up, mut uf) = (Vec::new(), Vec::new(), Vec::new(), Vec::new(), Vec::new());
        for s in sources {
            t.extend_from_slice(&s.times);
            f.extend_from_slice(&s.flux);
            v.extend_from_slice(&s.obs_var);
            up.extend(s.is_upper.iter().map(|&b| b as i32));
            uf.extend_from_slice(&s.upper_flux);
            off.push(t.len() as i64);
        }
        let mut b = std::ptr::null_mut();
        let rc = unsafe {
            lcf_batch_create_current(sources.len() as i64, off.as_ptr(), t.as_ptr(), f.as_ptr(), v.as_ptr(),
                                     up.as_ptr(), uf.as_ptr(), &mut b)
        };
        if rc != 0 {
            return Err(last_error(std::ptr::null(), rc));
        }
        Ok(Self { b, n_sources: sources.len() })
    }

    /// Number of observations for a given source index (1 out of range).
    pub fn n_obs_per_source(&self, source_idx: usize) -> usize {
        unsafe { lcf_batch_n_obs(self.b, source_idx as i64) as usize }
    }
}

/// A HIP context (stream, work queue, timing events) on one device.  Like
/// the reference's GpuBatchData it owns device state and is neither Send
/// nor Sync: create one per host thread.
pub struct GpuContext {
    ctx: *mut LcfCtx,
}

impl Drop for GpuContext {
    fn drop(&mut self) {
        unsafe { lcf_ctx_destroy(self.ctx) }
    }
}

impl GpuContext {
    fn check(&self, rc: c_int) -> Result<(), String> {
        if rc == 0 { Ok(()) } else { Err(last_error(self.ctx, rc)) }
    }

    /// Create a context on the given device (typically 0) and make that
    /// device current for the calling thread.
    pub fn new(device: i32) -> Result<Self, String> {
        let mut ctx = std::ptr::null_mut();
        let rc = unsafe { lcf_ctx_create(device, &mut ctx) };
        if rc != 0 || ctx.is_null() {
            return Err(format!("HIP error {rc}: device {device} unavailable"));
        }
        Ok(Self { ctx })
    }

    /// Evaluate a parametric model for many parameter draws at many time
    /// points: `n_draws x n_times`, row-major.
    pub fn eval_batch(&self, model: GpuModelName, params: &[f64], times: &[f64], n_draws: usize)
                      -> Result<Vec<f64>, String> {
        let n_params = model.n_params();
        if params.len() != n_draws * n_params {
            return Err(format!("params length {} != n_draws({}) x n_params({})", params.len(), n_draws, n_params));
        }
        let mut out = vec![0.0; n_draws * times.len()];
        self.check(unsafe {
            lcf_eval_batch(self.ctx, model.model_id(), params.as_ptr(), n_draws as i64, times.as_ptr(),
                           times.len() as i64, out.as_mut_ptr())
        })?;
        Ok(out)
    }

    /// PSO for one model across many sources; one `BatchPsoResult` per source.
    pub fn batch_pso(&self, model: GpuModelName, data: &GpuBatchData, n_particles: usize, max_iters: usize,
                     stall_iters: usize, seed: u64) -> Result<Vec<BatchPsoResult>, String> {
        let n = data.n_sources;
        let (mut p, mut c) = (vec![0.0; n * MAXP], vec![0.0; n]);
        self.check(unsafe {
            lcf_batch_pso(self.ctx, model.model_id(), data.b, n_particles as c_int, max_iters as c_int,
                          stall_iters as c_int, seed, p.as_mut_ptr(), c.as_mut_ptr())
        })?;
        let np = model.n_params();
        Ok((0..n).map(|s| BatchPsoResult { params: p[s * MAXP..s * MAXP + np].to_vec(), cost: c[s] }).collect())
    }

    /// Model selection over every GPU model (Bazin first, gated by
    /// `bazin_good_enough`): best model and its result per source.
    pub fn batch_model_select(&self, data: &GpuBatchData, n_particles: usize, max_iters: usize,
                              stall_iters: usize, bazin_good_enough: f64)
                              -> Result<Vec<(GpuModelName, BatchPsoResult)>, String> {
        let n = data.n_sources;
        let (mut m, mut p, mut c) = (vec![0i32; n], vec![0.0; n * MAXP], vec![0.0; n]);
        self.check(unsafe {
            lcf_batch_model_select(self.ctx, data.b, n_particles as c_int, max_iters as c_int,
                                   stall_iters as c_int, bazin_good_enough, m.as_mut_ptr(), p.as_mut_ptr(),
                                   c.as_mut_ptr())
        })?;
        (0..n)
            .map(|s| {
                let model = ALL_GPU_MODELS
                    .iter()
                    .copied()
                    .find(|x| x.model_id() == m[s])
                    .ok_or_else(|| format!("HIP backend returned unknown model id {}", m[s]))?;
                let np = model.n_params();
                Ok((model, BatchPsoResult { params: p[s * MAXP..s * MAXP + np].to_vec(), cost: c[s] }))
            })
            .collect()
    }

    /// PSO for every GPU model on all sources: one (model, results) per model.
    pub fn batch_all_models(&self, data: &GpuBatchData, n_particles: usize, max_iters: usize,
                            stall_iters: usize) -> Result<Vec<(GpuModelName, Vec<BatchPsoResult>)>, String> {
        let n = data.n_sources;
        let k = ALL_GPU_MODELS.len();
        let (mut p, mut c) = (vec![0.0; k * n * MAXP], vec![0.0; k * n]);
        self.check(unsafe {
            lcf_batch_all_models(self.ctx, data.b, n_particles as c_int, max_iters as c_int,
                                 stall_iters as c_int, p.as_mut_ptr(), c.as_mut_ptr())
        })?;
        Ok(ALL_GPU_MODELS
            .iter()
            .enumerate()
            .map(|(i, &m)| {
                let np = m.n_params();
                let rows = (0..n)
                    .map(|s| {
                        let o = (i * n + s) * MAXP;
                        BatchPsoResult { params: p[o..o + np].to_vec(), cost: c[i * n + s] }
                    })
                    .collect();
                (m, rows)
            })
            .collect())
    }

    /// Greedy MultiBazin (K = 1..4, BIC) for many sources.  `sources` must be
    /// the data the batch was built from; the device reads the time ranges
    /// and residuals it needs from the batch itself.
    pub fn batch_pso_multi_bazin(&self, data: &GpuBatchData, sources: &[BatchSource], n_particles: usize,
                                 max_iters: usize, stall_iters: usize, seed: u64)
                                 -> Result<Vec<MultiBazinResult>, String> {
        let n = data.n_sources;
        assert_eq!(n, sources.len());
        let mut best_k = vec![0i32; n];
        let mut params = vec![0.0; n * MB_ROW];
        let (mut cost, mut bic) = (vec![0.0; n], vec![0.0; n]);
        let (mut kc, mut kb) = (vec![0.0; n * MB_MAXK], vec![0.0; n * MB_MAXK]);
        self.check(unsafe {
            lcf_batch_multi_bazin(self.ctx, data.b, n_particles as c_int, max_iters as c_int,
                                  stall_iters as c_int, seed, best_k.as_mut_ptr(), params.as_mut_ptr(),
                                  cost.as_mut_ptr(), bic.as_mut_ptr(), kc.as_mut_ptr(), kb.as_mut_ptr())
        })?;
        Ok((0..n)
            .map(|s| {
                let k = best_k[s] as usize;
                // no K with a finite BIC: the reference keeps best_k 1 and empty params
                let p = if bic[s].is_finite() { params[s * MB_ROW..s * MB_ROW + 4 * k + 2].to_vec() } else { Vec::new() };
                MultiBazinResult {
                    best_k: k.max(1),
                    params: p,
                    cost: cost[s],
                    bic: bic[s],
                    per_k_cost: kc[s * MB_MAXK..(s + 1) * MB_MAXK].to_vec(),
                    per_k_bic: kb[s * MB_MAXK..(s + 1) * MB_MAXK].to_vec(),
                }
            })
            .collect())
    }

    /// DenseGP fits of many bands: subsample, hyper-parameter search, and
    /// predictions at the shared `query_times` and at each band's points.
    pub fn batch_gp_fit(&self, bands: &[GpBandInput], query_times: &[f64], amp_candidates: &[f64],
                        ls_candidates: &[f64], max_subsample: usize) -> Result<Vec<GpBandOutput>, String> {
        use crate::sparse_gp::DenseGP;
        let nb = bands.len();
        if nb == 0 {
            return Ok(Vec::new());
        }
        let nq = query_times.len();
        assert!(nq == 50, "batch_gp_fit expects 50 query points");
        let mut off = vec![0i64];
        let (mut t, mut m, mut nv) = (Vec::new(), Vec::new(), Vec::new());
        for b in bands {
            t.extend_from_slice(&b.times);
            m.extend_from_slice(&b.mags);
            nv.extend_from_slice(&b.noise_var);
            off.push(t.len() as i64);
        }
        let mut st = vec![0.0; nb * GP_STATE];
        let (mut pg, mut sg, mut po) = (vec![0.0; nb * nq], vec![0.0; nb * nq], vec![0.0; t.len()]);
        self.check(unsafe {
            lcf_batch_gp_fit(self.ctx, nb as i64, off.as_ptr(), t.as_ptr(), m.as_ptr(), nv.as_ptr(),
                             query_times.as_ptr(), nq as c_int, amp_candidates.as_ptr(),
                             amp_candidates.len() as c_int, ls_candidates.as_ptr(), ls_candidates.len() as c_int,
                             max_subsample as c_int, st.as_mut_ptr(), pg.as_mut_ptr(), sg.as_mut_ptr(),
                             po.as_mut_ptr())
        })?;
        Ok((0..nb)
            .map(|b| {
                let s = &st[b * GP_STATE..(b + 1) * GP_STATE];
                let k = s[53] as usize;
                // the DenseGP of src/gpu_cuda.rs:1093-1132, rebuilt from the
                // device-selected training set and hyper-parameters
                let dense_gp = if k > 0 && k <= 25 {
                    let n = bands[b].times.len();
                    let pick = |v: &[f64]| -> Vec<f64> {
                        if n <= k {
                            v.to_vec()
                        } else {
                            let step = (n - 1) as f64 / (k - 1) as f64;
                            (0..k).map(|i| v[((i as f64 * step + 0.5) as usize).min(n - 1)]).collect()
                        }
                    };
                    DenseGP::fit(&s[25..25 + k], &pick(&bands[b].mags), &pick(&bands[b].noise_var), s[50],
                                 (0.5 / s[51]).sqrt())
                } else {
                    None
                };
                let (a, e) = (off[b] as usize, off[b + 1] as usize);
                GpBandOutput {
                    dense_gp,
                    pred_grid: pg[b * nq..(b + 1) * nq].to_vec(),
                    std_grid: sg[b * nq..(b + 1) * nq].to_vec(),
                    pred_at_obs: po[a..e].to_vec(),
                }
            })
            .collect())
    }

    /// 2D (time x log10 wavelength) GP fits of many sources: predictions at
    /// the (query_times[k], query_waves[k]) points and the chosen
    /// hyper-parameters per source.
    pub fn batch_gp_2d(&self, sources: &[Gp2dInput], query_times: &[f64], query_waves: &[f64],
                       amp_candidates: &[f64], lst_candidates: &[f64], lsw_candidates: &[f64],
                       max_subsample: usize) -> Result<Vec<Gp2dOutput>, String> {
        let ns = sources.len();
        if ns == 0 {
            return Ok(Vec::new());
        }
        let nq = query_times.len();
        assert_eq!(nq, query_waves.len(), "query_times and query_waves must have same length");
        let n_hp = amp_candidates.len() * lst_candidates.len() * lsw_candidates.len();
        assert!(n_hp <= 64, "max 64 hyperparameter combos for 2D GP");
        let mut off = vec![0i64];
        let (mut t, mut w, mut m, mut nv) = (Vec::new(), Vec::new(), Vec::new(), Vec::new());
        for s in sources {
            t.extend_from_slice(&s.times);
            w.extend_from_slice(&s.waves);
            m.extend_from_slice(&s.mags);
            nv.extend_from_slice(&s.noise_var);
            off.push(t.len() as i64);
        }
        let mut st = vec![0.0; ns * GP2D_STATE];
        let (mut pg, mut sg, mut rms) = (vec![0.0; ns * nq], vec![0.0; ns * nq], vec![0.0; ns]);
        self.check(unsafe {
            lcf_batch_gp2d_fit(self.ctx, ns as i64, off.as_ptr(), t.as_ptr(), w.as_ptr(), m.as_ptr(), nv.as_ptr(),
                               query_times.as_ptr(), query_waves.as_ptr(), nq as c_int, amp_candidates.as_ptr(),
                               amp_candidates.len() as c_int, lst_candidates.as_ptr(),
                               lst_candidates.len() as c_int, lsw_candidates.as_ptr(),
                               lsw_candidates.len() as c_int, max_subsample as c_int, st.as_mut_ptr(),
                               pg.as_mut_ptr(), sg.as_mut_ptr(), rms.as_mut_ptr())
        })?;
        Ok((0..ns)
            .map(|s| {
                let state = &st[s * GP2D_STATE..(s + 1) * GP2D_STATE];
                let pred_grid = pg[s * nq..(s + 1) * nq].to_vec();
                let std_grid = sg[s * nq..(s + 1) * nq].to_vec();
                let n_train = sources[s].times.len();
                if state[124] as usize >= 3 {
                    Gp2dOutput {
                        pred_grid,
                        std_grid,
                        amp: state[120],
                        ls_time: (0.5 / state[121]).sqrt(),
                        ls_wave: (0.5 / state[122]).sqrt(),
                        train_rms: rms[s],
                        n_train,
                        success: true,
                    }
                } else {
                    Gp2dOutput {
                        pred_grid,
                        std_grid,
                        amp: 0.0,
                        ls_time: 0.0,
                        ls_wave: 0.0,
                        train_rms: f64::NAN,
                        n_train,
                        success: false,
                    }
                }
            })
            .collect())
    }

    /// SVI for many sources at once (each with its own model, PSO start and
    /// prior); `log_sigma` includes the reference's sigma inflation.
    pub fn batch_svi_fit(&self, data: &GpuBatchData, inputs: &[SviBatchInput], n_steps: usize,
                         n_samples: usize, lr: f64) -> Result<Vec<SviBatchOutput>, String> {
        let n = inputs.len();
        if n == 0 {
            return Ok(Vec::new());
        }
        assert_eq!(n, data.n_sources, "SVI inputs must match batch data size");
        let (mut ids, mut se) = (vec![0i32; n], vec![0i32; n]);
        let (mut pso, mut pc, mut pw) = (vec![0.0; n * MAXP], vec![0.0; n * MAXP], vec![0.0; n * MAXP]);
        for (i, inp) in inputs.iter().enumerate() {
            let k = inp.pso_params.len().min(MAXP);
            pso[i * MAXP..i * MAXP + k].copy_from_slice(&inp.pso_params[..k]);
            let kc = inp.prior_centers.len().min(MAXP);
            pc[i * MAXP..i * MAXP + kc].copy_from_slice(&inp.prior_centers[..kc]);
            let kw = inp.prior_widths.len().min(MAXP);
            pw[i * MAXP..i * MAXP + kw].copy_from_slice(&inp.prior_widths[..kw]);
            ids[i] = inp.model_id as i32;
            se[i] = inp.se_idx as i32;
        }
        let (mut mu, mut ls, mut elbo) = (vec![0.0; n * MAXP], vec![0.0; n * MAXP], vec![0.0; n]);
        self.check(unsafe {
            lcf_batch_svi_fit(self.ctx, data.b, ids.as_ptr(), pso.as_ptr(), se.as_ptr(), pc.as_ptr(), pw.as_ptr(),
                              n_steps as c_int, n_samples as c_int, lr, mu.as_mut_ptr(), ls.as_mut_ptr(),
                              elbo.as_mut_ptr())
        })?;
        Ok(inputs
            .iter()
            .enumerate()
            .map(|(i, inp)| {
                let k = inp.pso_params.len().min(MAXP);
                SviBatchOutput {
                    mu: mu[i * MAXP..i * MAXP + k].to_vec(),
                    log_sigma: ls[i * MAXP..i * MAXP + k].to_vec(),
                    elbo: elbo[i],
                }
            })
            .collect())
    }
}
