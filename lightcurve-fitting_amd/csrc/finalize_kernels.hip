// finalize_kernels.hip -- batched profile refinement and magnitude chi2 of the
// parametric fits (FP64).
//
// The reference finishes every GPU-fitted band on the host
// (finalize_parametric_with_gpu_svi, src/parametric.rs:3298-3421, and its two
// siblings): profile_t0_refine (:2788-3008), profile_refine_all (:2705-2786)
// and the magnitude-space chi2 (:3366-3393).  Here one wavefront takes one
// band (persistent queue, one launch per model) with the band's observations
// staged in LDS, and every log-likelihood the reference evaluates becomes one
// lane task that walks the observations in order, so each sum rounds as the
// reference's sequential loop does:
//   pass 1  lanes 0-24: the refitted t0 grid (a weighted least-squares pass
//           over the shape curve for the amplitude, and the baseline of Bazin
//           and TDE, then the likelihood); lanes 32-56: the same grid without
//           refit, which sets the t0 width;
//   pass 2  the likelihood at the updated mean and at the 15 (np - 1) points
//           of the per-parameter scans (at most 91 tasks, two rounds);
//   chi2    lanes over observations evaluate the SVI and PSO curves (MetzgerKN:
//           lane 0 and lane 1 integrate one curve each), lane 0 adds the
//           terms in order.
// The decisions between passes (best grid point, the walk out to the
// Delta LL = 0.5 threshold, the widths) are taken by every lane from the LDS
// results, so they are wave-uniform without broadcasts.  Built without
// multiply-add contraction.
#include "finalize_kernels.hpp"
#include "metzger_batch.hpp"
#include "model_tables.hpp"
#include "models_fit.hpp"

namespace lcf {

namespace {

constexpr double FIN_PI = 3.14159265358979323846;
constexpr int T0_GRID = 25;    // profile_t0_refine grid
constexpr int SCAN_GRID = 15;  // profile_refine_all grid
constexpr int FIN_VALS = 128;

// SviModel::t0_idx (src/parametric.rs:152-163) by model id
constexpr int t0_index(int m) { return (m == 0 || m == 2) ? 2 : (m == 1 || m == 7) ? 3 : 1; }

// Model curve over a band: closed forms per time, MetzgerKN integrated once
// per parameter vector over the band's largest phase (metzger_kn_eval_batch).
template <int MODEL>
struct Curve {
    Model<MODEL, double> mdl;
    __device__ __forceinline__ void setup(const double* p, const double4*, int) { mdl.setup(p); }
    __device__ __forceinline__ double at(double t) const {
        double f;
        (void)mdl.pred(t, f);
        return f;
    }
};

template <>
struct Curve<7> {
    double gt[MKN_GRID], gl[MKN_GRID];
    double t0;
    bool live;
    __device__ __forceinline__ void setup(const double* p, const double4* o, int n) {
        double phmax = 0.01;
        for (int i = 0; i < n; i++) phmax = fmax(phmax, o[i].x - p[3]);
        t0 = p[3];
        live = metzger_grid(p, phmax, gt, gl);
    }
    __device__ __forceinline__ double at(double t) const { return live ? metzger_interp(gt, gl, t - t0) : 0.0; }
};

template <int MODEL>
struct Fin {
    static constexpr int NP = ModelTraits<MODEL>::NP;
    static constexpr int SE = NP - 1;
    static constexpr int T0 = t0_index(MODEL);
    static constexpr bool HAS_B = (MODEL == 0 || MODEL == 2);

    // CPU pso_bounds (src/parametric.rs:1541-1579)
    __device__ static double lower(int j) { return (MODEL == 0 && j == 1) ? -BAZIN_B_CPU : c_lo[MODEL][j]; }
    __device__ static double upper(int j) { return (MODEL == 0 && j == 1) ? BAZIN_B_CPU : c_hi[MODEL][j]; }

    // One observation's Gaussian / censored log-likelihood term (callers skip
    // non-finite predictions).
    __device__ static double ll_term(const double4& o, double f, double sx2) {
        const double tv = o.z + sx2;
        if (o.w != 0.0) return log_normal_cdf<double>((o.y - f) / ::sqrt(tv));
        const double r = o.y - f;
        return -0.5 * (r * r / tv + ::log(2.0 * FIN_PI * tv));
    }

    // Log-likelihood of p over the band, observations in order
    // (src/parametric.rs:2717-2726).
    __device__ static double loglik(const double (&p)[NP], const double4* o, int n, Curve<MODEL>& cv) {
        const double sx = ::exp(p[SE]);
        const double sx2 = sx * sx;
        cv.setup(p, o, n);
        double ll = 0.0;
        for (int i = 0; i < n; i++) {
            const double4 ob = o[i];
            const double f = cv.at(ob.x);
            if (__builtin_isfinite(f)) ll += ll_term(ob, f, sx2);
        }
        return ll;
    }

    // Likelihood at t0 after the closed-form amplitude (and baseline) refit
    // (src/parametric.rs:2824-2947).
    __device__ static double refit_ll(const double* mu, double t0, const double4* o, int n, Curve<MODEL>& cv) {
        double p[NP];
#pragma unroll
        for (int j = 0; j < NP; j++) p[j] = mu[j];
        p[T0] = t0;
        const double sx = ::exp(mu[SE]);
        const double sx2 = sx * sx;
        const double sa = p[0];
        p[0] = 0.0;
        double sb = 0.0;
        if constexpr (HAS_B) {
            sb = p[1];
            p[1] = 0.0;
        }
        cv.setup(p, o, n);
        if constexpr (HAS_B) {
            double sw = 0.0, sy = 0.0, sf = 0.0, syf = 0.0, sff = 0.0;
            for (int i = 0; i < n; i++) {
                const double4 ob = o[i];
                const double s = cv.at(ob.x);
                if (ob.w != 0.0 || !__builtin_isfinite(s)) continue;
                const double w = 1.0 / (ob.z + sx2);
                sw += w;
                sy += w * ob.y;
                sf += w * s;
                syf += w * ob.y * s;
                sff += w * s * s;
            }
            const double det = sw * sff - sf * sf;
            p[0] = sa;
            p[1] = sb;
            if (fabs(det) > 1e-20) {
                const double aa = (sw * syf - sf * sy) / det;
                const double bo = (sff * sy - sf * syf) / det;
                if (aa > 1e-10) {
                    p[0] = ::log(aa);
                    p[1] = bo;
                }
            }
        } else {
            double num = 0.0, den = 0.0;
            for (int i = 0; i < n; i++) {
                const double4 ob = o[i];
                const double s = cv.at(ob.x);
                if (ob.w != 0.0 || !__builtin_isfinite(s)) continue;
                const double w = 1.0 / (ob.z + sx2);
                num += w * ob.y * s;
                den += w * s * s;
            }
            p[0] = (den > 1e-20 && num / den > 1e-10) ? ::log(num / den) : sa;
        }
        cv.setup(p, o, n);
        double sc = 1.0;
        if constexpr (MODEL == 7) {
            // MetzgerKN predictions rescaled to their detected peak (:2952-2965)
            double mx = -INFINITY;
            for (int i = 0; i < n; i++)
                if (o[i].w == 0.0) mx = fmax(mx, cv.at(o[i].x));
            if (mx > 1e-10 && __builtin_isfinite(mx)) sc = fmin(fmax(1.0 / mx, 0.1), 10.0);
        }
        double ll = 0.0;
        for (int i = 0; i < n; i++) {
            const double4 ob = o[i];
            const double f = cv.at(ob.x) * sc;
            if (__builtin_isfinite(f)) ll += ll_term(ob, f, sx2);
        }
        return ll;
    }

    // profile_t0_refine (src/parametric.rs:2788-3008)
    __device__ static void t0_refine(const FinalizeArgs& a, const double4* o, double* vals, int n, int64_t base,
                                     double (&mu)[NP], double (&ls)[NP], Curve<MODEL>& cv) {
        const int lane = lane_id();
        double tf = INFINITY, pv = -INFINITY;
        int pk = 0x7fffffff;
        for (int i = lane; i < n; i += WAVE) {
            tf = fmin(tf, o[i].x);
            const double f = a.flux[base + i];
            if (f > pv) {
                pv = f;
                pk = i;
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            tf = fmin(tf, __shfl_xor(tf, off, WAVE));
            const double ov = __shfl_xor(pv, off, WAVE);
            const int oi = __shfl_xor(pk, off, WAVE);
            if (ov > pv || (ov == pv && oi < pk)) {
                pv = ov;
                pk = oi;
            }
        }
        if (pv == -INFINITY) pk = 0;  // no flux above -inf: the reference keeps index 0
        const double lo = tf - 30.0, hi = o[pk].x + 5.0;
        if (lo >= hi) return;
        auto grid = [&](int g) { return lo + (hi - lo) * (double)g / (double)(T0_GRID - 1); };
        double v = 0.0;
        if (lane < T0_GRID) {
            v = refit_ll(mu, grid(lane), o, n, cv);
        } else if (lane >= 32 && lane < 32 + T0_GRID) {
            double p[NP];
#pragma unroll
            for (int j = 0; j < NP; j++) p[j] = mu[j];
            p[T0] = grid(lane - 32);
            v = loglik(p, o, n, cv);
        }
        vals[lane] = v;
        __syncthreads();
        double best_t0 = mu[T0], best_ll = -INFINITY;
        for (int g = 0; g < T0_GRID; g++) {
            const double l = vals[g];
            if (l > best_ll) {
                best_ll = l;
                best_t0 = grid(g);
            }
        }
        const double thr = best_ll - 0.5;
        double lo1 = best_t0, hi1 = best_t0;
        for (int g = T0_GRID - 1; g >= 0; g--) {
            const double t0 = grid(g);
            if (t0 >= best_t0) continue;
            if (vals[32 + g] >= thr) lo1 = t0;
            else break;
        }
        for (int g = 0; g < T0_GRID; g++) {
            const double t0 = grid(g);
            if (t0 <= best_t0) continue;
            if (vals[32 + g] >= thr) hi1 = t0;
            else break;
        }
        mu[T0] = best_t0;
        ls[T0] = ::log(fmax((hi1 - lo1) / 2.0, 0.1));
        __syncthreads();
    }

    // profile_refine_all (src/parametric.rs:2705-2786)
    __device__ static void refine_all(const double4* o, double* vals, int n, const double (&mu)[NP],
                                      double (&ls)[NP], Curve<MODEL>& cv) {
        const int lane = lane_id();
        constexpr int E = 1 + SCAN_GRID * (NP - 1);
        for (int b0 = 0; b0 < E; b0 += WAVE) {
            const int e = b0 + lane;
            if (e < E) {
                double p[NP];
#pragma unroll
                for (int j = 0; j < NP; j++) p[j] = mu[j];
                if (e > 0) {
                    const int j = (e - 1) / SCAN_GRID, g = (e - 1) % SCAN_GRID;
                    const double l = lower(j), h = upper(j);
                    const double val = l + (h - l) * (double)g / (double)(SCAN_GRID - 1);
#pragma unroll
                    for (int k = 0; k < NP - 1; k++)
                        if (k == j) p[k] = val;
                }
                vals[e] = loglik(p, o, n, cv);
            }
        }
        __syncthreads();
        const double thr = vals[0] - 0.5;
#pragma unroll
        for (int j = 0; j < NP - 1; j++) {
            const double svi_sigma = ::exp(ls[j]);
            const double l = lower(j), h = upper(j);
            if (h - l < 1e-10) continue;
            double lo1 = mu[j], hi1 = mu[j];
            for (int g = 0; g < SCAN_GRID; g++) {
                const double val = l + (h - l) * (double)g / (double)(SCAN_GRID - 1);
                if (vals[1 + j * SCAN_GRID + g] >= thr) {
                    lo1 = fmin(lo1, val);
                    hi1 = fmax(hi1, val);
                }
            }
            const double prof = fmax((hi1 - lo1) / 2.0, 1e-6);
            ls[j] = ::log(fmax(fmax(svi_sigma, prof), c_pw[MODEL][j]));
        }
        __syncthreads();
    }

    // Magnitude-space chi2 (src/parametric.rs:3366-3393); the value is valid
    // on lane 0.  Overwrites the LDS observations.
    __device__ static double mag_chi2(const FinalizeArgs& a, double4* o, int n, int64_t base, int band,
                                      const double (&mu)[NP], Curve<MODEL>& cv) {
        const int lane = lane_id();
        const double peak = a.peak[band];
        double q[NP];
        __syncthreads();
        if constexpr (MODEL == 7) {
            if (lane < 2) {
#pragma unroll
                for (int j = 0; j < NP; j++) q[j] = lane == 0 ? mu[j] : a.pso[(int64_t)band * 7 + j];
                cv.setup(q, o, n);
                for (int i = 0; i < n; i++) {
                    const double f = cv.at(o[i].x);
                    if (lane == 0) o[i].z = f;
                    else o[i].w = f;
                }
            }
        } else {
            Curve<MODEL> cp;
#pragma unroll
            for (int j = 0; j < NP; j++) q[j] = a.pso[(int64_t)band * 7 + j];
            cv.setup(mu, o, n);
            cp.setup(q, o, n);
            for (int i = lane; i < n; i += WAVE) {
                const double t = o[i].x;
                o[i].z = cv.at(t);
                o[i].w = cp.at(t);
            }
        }
        __syncthreads();
        int cnt = 0;
        for (int i = lane; i < n; i += WAVE) cnt += (o[i].z * peak > 0.0) ? 1 : 0;
        cnt = wave_sum(cnt);
        const bool use_pso = !(a.flags & FIN_SVI_CHI2) && cnt < n / 2;
        for (int i = lane; i < n; i += WAVE) {
            const double pf = (use_pso ? o[i].w : o[i].z) * peak;
            const double f = a.flux[base + i];
            double inc = 0.0, term = 0.0;
            if (pf > 0.0 && f * peak > 0.0) {
                const double mp = -2.5 * ::log10(pf) + 23.9;
                const double me = 1.0857 * a.flux_err[base + i] / f;
                if (me > 0.0) {
                    const double r = a.mag_obs[base + i] - mp;
                    term = r * r / (me * me);
                    inc = 1.0;
                }
            }
            o[i].x = inc;
            o[i].y = term;
        }
        __syncthreads();
        double s = 0.0;
        int c = 0;
        if (lane == 0)
            for (int i = 0; i < n; i++)
                if (o[i].x != 0.0) {
                    s += o[i].y;
                    c++;
                }
        return c > 0 ? s / (double)c : NAN;
    }

    __device__ static void run(const FinalizeArgs& a, double4* lds, double* vals, int band) {
        const int lane = lane_id();
        const int64_t base = a.offsets[band];
        const int n = (int)(a.offsets[band + 1] - base);
        const double4* g = reinterpret_cast<const double4*>(a.obs) + base;
        __syncthreads();
        for (int i = lane; i < n; i += WAVE) lds[i] = g[i];
        __syncthreads();
        double mu[NP], ls[NP];
#pragma unroll
        for (int j = 0; j < NP; j++) {
            mu[j] = a.mu_in[(int64_t)band * 7 + j];
            ls[j] = a.ls_in[(int64_t)band * 7 + j];
        }
        Curve<MODEL> cv;
        double chi2 = NAN;
        if (n > 0) {
            if (a.flags & FIN_T0) t0_refine(a, lds, vals, n, base, mu, ls, cv);
            if (a.flags & FIN_ALL) refine_all(lds, vals, n, mu, ls, cv);
            if (a.flags & FIN_MAG_CHI2) chi2 = mag_chi2(a, lds, n, base, band, mu, cv);
        }
        if (lane == 0) {
            for (int j = 0; j < 7; j++) {
                const int64_t k = (int64_t)band * 7 + j;
                a.mu_out[k] = a.mu_in[k];
                a.ls_out[k] = a.ls_in[k];
            }
#pragma unroll
            for (int j = 0; j < NP; j++) {
                a.mu_out[(int64_t)band * 7 + j] = mu[j];
                a.ls_out[(int64_t)band * 7 + j] = ls[j];
            }
            a.mag_chi2[band] = chi2;
        }
    }
};

}  // namespace

template <int MODEL>
__global__ __launch_bounds__(64) void finalize_kernel(FinalizeArgs a) {
    extern __shared__ double4 fin_lds[];
    __shared__ double vals[FIN_VALS];
    for (;;) {
        const int t = wave_next_task(a.queue);
        if (t >= a.n_tasks) break;
        Fin<MODEL>::run(a, fin_lds, vals, a.order[t]);
        wave_reconverge();
    }
}

static const void* fin_ptr(int model) {
    switch (model) {
        case 0: return (const void*)finalize_kernel<0>;
        case 1: return (const void*)finalize_kernel<1>;
        case 2: return (const void*)finalize_kernel<2>;
        case 3: return (const void*)finalize_kernel<3>;
        case 4: return (const void*)finalize_kernel<4>;
        case 5: return (const void*)finalize_kernel<5>;
        case 6: return (const void*)finalize_kernel<6>;
        case 7: return (const void*)finalize_kernel<7>;
        default: return nullptr;
    }
}

static size_t fin_lds_bytes(int max_n) { return sizeof(double4) * (size_t)(max_n > 0 ? max_n : 1); }

hipError_t launch_finalize(int model, const FinalizeArgs& a, int grid, hipStream_t st) {
    const void* k = fin_ptr(model);
    if (!k || grid <= 0 || a.max_n < 0 || a.max_n > FIN_MAX_OBS) return hipErrorInvalidValue;
    void* args[] = {const_cast<FinalizeArgs*>(&a)};
    return hipLaunchKernel(k, dim3(grid), dim3(64), args, fin_lds_bytes(a.max_n), st);
}

hipError_t finalize_occupancy(int model, int max_n, int* blocks_per_cu) {
    const void* k = fin_ptr(model);
    if (!k) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, 64, fin_lds_bytes(max_n));
}

}  // namespace lcf
