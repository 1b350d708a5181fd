// laplace_kernels.hip -- batched Laplace posterior approximation (FP64).
//
// One wavefront per (band, model) job.  The Hessian of the un-normalised
// negative log posterior over the shape parameters (sigma_x held at its MAP
// value) is assembled from evaluations spread over the lanes, each lane
// walking the band's LDS copy of the observations:
//   * LCF_LAPLACE_FD4 (the reference's laplace_fit, src/parametric.rs:2623-2693):
//     lane e evaluates f at one of the 1 + 4 * ns(ns+1)/2 <= 85 stencil points
//     theta +- h_i e_i +- h_j e_j, h_i = 1e-4 max(|theta_i|, 1);
//     H_ij = (f++ - f+- - f-+ + f--) / (4 h_i h_j);
//   * LCF_LAPLACE_GRAD_FD (forward over adjoint): lane e evaluates the
//     analytic gradient at theta +- h_i e_i (2 ns <= 12 points);
//     H_ij = sym((g_j(+i) - g_j(-i)) / (2 h_i)).
// Lane 0 then runs the cyclic-max Jacobi eigensolver of symmetric_eigen
// (:2517-2587) and forms log sigma = 1/2 ln diag(V diag(1/max(lambda,1e-6)) V^T)
// (:2590-2613); sigma_x's slot is ln(prior width), elbo = -f(MAP).
// neg_log_posterior (:2464-2511) uses the Gaussian likelihood with the 1/2
// and ln 2 pi terms and the prior_params priors; MetzgerKN is integrated on
// the 200-point grid of metzger_kn_eval_batch and rescaled to its peak.
#include "laplace_kernels.hpp"
#include "metzger_batch.hpp"
#include "models_grad.hpp"

namespace lcf {

// prior_params (src/parametric.rs:2072-2115), [model][param]
__constant__ double c_lp_c[8][7] = {
    {0, 0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 1.67, 0},
    {0, 0, 2.3, 0, 0, 0, 0}, {0, 0, 3.0, 2.3, 0, 0, 0}, {0, 0, 0.5, 1.0, 0, 0, 0},
    {0, 0, 1.0, 0.8, 2.0, 0, 0}, {-1.75, -1.25, 0.5, -0.5, 0.0, 0, 0}};
__constant__ double c_lp_w[8][7] = {
    {2, 2, 2, 2, 2, 2, 2}, {2, 2, 2, 2, 2, 2, 2}, {2, 2, 2, 2, 2, 1.0, 2},
    {2, 2, 1.0, 2, 2, 2, 2}, {2, 2, 1.5, 1.0, 2, 2, 2}, {2, 2, 1.0, 2.0, 2, 2, 2},
    {2, 2, 2.0, 1.0, 1.0, 2, 2}, {1.25, 0.75, 1.5, 1.5, 2.0, 2, 2}};

constexpr int LAP_MAX_PTS = 85;
constexpr double LAP_PI = 3.14159265358979323846;

template <int MODEL>
struct LaplaceJob {
    static constexpr int NP = ModelTraits<MODEL>::NP;
    static constexpr int NS = NP - 1;
    static constexpr int SE = NP - 1;

    // neg_log_posterior at p (one lane).
    __device__ double nlp(const double (&p)[NP], const double4* lds, int n) {
        const double se2 = ::exp(2.0 * p[SE]);
        double nll = 0.0;
        bool ok = true;
        if (MODEL == 7) {
            double gt[200], gl[200];
            double phmax = 0.01;
            for (int i = 0; i < n; i++) phmax = fmax(phmax, lds[i].x - p[3]);
            bool live = metzger_grid(p, phmax, gt, gl);
            double mx = -INFINITY;
            for (int i = 0; i < n; i++)
                if (lds[i].w == 0.0) mx = fmax(mx, live ? metzger_interp(gt, gl, lds[i].x - p[3]) : 0.0);
            double sc = 1.0;
            if (mx > 1e-10 && __builtin_isfinite(mx)) sc = fmin(fmax(1.0 / mx, 0.1), 10.0);
            for (int i = 0; i < n; i++) {
                double4 o = lds[i];
                double f = (live ? metzger_interp(gt, gl, o.x - p[3]) : 0.0) * sc;
                double tv = o.z + se2;
                if (o.w != 0.0) nll -= log_normal_cdf<double>((o.y - f) / ::sqrt(tv));
                else {
                    double d = f - o.y;
                    nll += 0.5 * (d * d / tv + ::log(2.0 * LAP_PI * tv));
                }
            }
        } else {
            Model<MODEL, double> mdl;
            mdl.setup(p);
            for (int i = 0; i < n; i++) {
                double4 o = lds[i];
                double f;
                ok = ok && mdl.pred(o.x, f) && __builtin_isfinite(f);
                double tv = o.z + se2;
                if (o.w != 0.0) nll -= log_normal_cdf<double>((o.y - f) / ::sqrt(tv));
                else {
                    double d = f - o.y;
                    nll += 0.5 * (d * d / tv + ::log(2.0 * LAP_PI * tv));
                }
            }
        }
        if (!ok) return 1e99;
        double lp = 0.0;
#pragma unroll
        for (int j = 0; j < NP; j++) {
            double d = p[j] - c_lp_c[MODEL][j], w = c_lp_w[MODEL][j];
            lp += 0.5 * d * d / (w * w);
        }
        return nll + lp;
    }

    // gradient of neg_log_posterior over the shape parameters (one lane).
    __device__ void nlp_grad(const double (&p)[NP], const double4* lds, int n, double (&gs)[NP]) {
        ModelGrad<MODEL, double> mdl;
        mdl.setup(p);
        const double se2 = ::exp(2.0 * p[SE]);
#pragma unroll
        for (int j = 0; j < NP; j++) gs[j] = 0.0;
        for (int i = 0; i < n; i++) {
            double4 o = lds[i];
            double f, g[NP];
            if (!mdl.pred_grad(o.x, f, g) || !__builtin_isfinite(f)) continue;
            double tv = o.z + se2, st = ::sqrt(tv);
            double w;
            if (o.w != 0.0) {
                double z = (o.y - f) / st;
                double ph = ::exp(-0.5 * z * z) * 0.39894228040143267794;
                double a = fabs(z * 0.70710678118654752440);
                double t = 1.0 / (1.0 + 0.3275911 * a);
                double poly = t * (0.254829592 + t * (-0.284496736 + t * (1.421413741 + t * (-1.453152027 + t * 1.061405429))));
                double val = 1.0 - poly * ::exp(-a * a);
                double cdf = fmax(0.5 * (1.0 + (z >= 0.0 ? val : -val)), 1e-300);
                w = ph / cdf / st;
            } else {
                w = (f - o.y) / tv;
            }
#pragma unroll
            for (int j = 0; j < NS; j++) gs[j] += w * g[j];
        }
#pragma unroll
        for (int j = 0; j < NS; j++) {
            double w = c_lp_w[MODEL][j];
            gs[j] += (p[j] - c_lp_c[MODEL][j]) / (w * w);
        }
    }

    __device__ void run(const LaplaceArgs& a, double4* lds, double* vals, int job) {
        const int lane = lane_id();
        const double4* g = reinterpret_cast<const double4*>(a.obs) + a.offsets[job];
        const int n = (int)(a.offsets[job + 1] - a.offsets[job]);
        __syncthreads();
        for (int i = lane; i < n; i += WAVE) lds[i] = g[i];
        __syncthreads();
        double map[NP], h[NP];
#pragma unroll
        for (int j = 0; j < NP; j++) {
            map[j] = a.map[(int64_t)job * 7 + j];
            h[j] = 1e-4 * fmax(fabs(map[j]), 1.0);
        }
        const bool fd4 = a.mode == 0;
        constexpr int NPAIR = NS * (NS + 1) / 2;
        const int E = fd4 ? 1 + 4 * NPAIR : 1 + 2 * NS;
        for (int base = 0; base < E; base += WAVE) {
            const int e = base + lane;
            if (e < E) {
                double p[NP];
#pragma unroll
                for (int j = 0; j < NP; j++) p[j] = map[j];
                if (e == 0) {
                    vals[0] = nlp(p, lds, n);
                } else if (fd4) {
                    int pr = (e - 1) >> 2, k = (e - 1) & 3;
                    int i = 0, rem = pr;
                    while (rem >= NS - i) { rem -= NS - i; i++; }
                    int j = i + rem;
                    p[i] += (k < 2) ? h[i] : -h[i];
                    p[j] += (k & 1) ? -h[j] : h[j];
                    vals[e] = nlp(p, lds, n);
                } else {
                    int i = (e - 1) >> 1;
                    p[i] += (e & 1) ? h[i] : -h[i];
                    double gs[NP];
                    nlp_grad(p, lds, n, gs);
#pragma unroll
                    for (int j = 0; j < NS; j++) vals[LAP_MAX_PTS + (e - 1) * NS + j] = gs[j];
                }
            }
        }
        __syncthreads();
        if (lane == 0) {
            double H[NS * NS];
            if (fd4) {
                int pr = 0;
                for (int i = 0; i < NS; i++)
                    for (int j = i; j < NS; j++, pr++) {
                        const double* f4 = vals + 1 + 4 * pr;
                        double hij = (f4[0] - f4[1] - f4[2] + f4[3]) / (4.0 * h[i] * h[j]);
                        H[i * NS + j] = hij;
                        H[j * NS + i] = hij;
                    }
            } else {
                // gradient slot 2i holds theta + h_i e_i, slot 2i+1 theta - h_i e_i
                const double* G = vals + LAP_MAX_PTS;
                for (int i = 0; i < NS; i++)
                    for (int j = 0; j < NS; j++) {
                        double gij = (G[(2 * i) * NS + j] - G[(2 * i + 1) * NS + j]) / (2.0 * h[i]);
                        double gji = (G[(2 * j) * NS + i] - G[(2 * j + 1) * NS + i]) / (2.0 * h[j]);
                        H[i * NS + j] = 0.5 * (gij + gji);
                    }
            }
            for (int i = 0; i < NS * NS; i++)
                if (a.out_hessian) a.out_hessian[(int64_t)job * 36 + i] = H[i];
            // cyclic-max Jacobi (symmetric_eigen)
            double V[NS * NS];
            for (int i = 0; i < NS * NS; i++) V[i] = 0.0;
            for (int i = 0; i < NS; i++) V[i * NS + i] = 1.0;
            for (int it = 0; it < 100 * NS * NS; it++) {
                double mv = 0.0;
                int P = 0, Q = 1;
                for (int i = 0; i < NS; i++)
                    for (int j = i + 1; j < NS; j++)
                        if (fabs(H[i * NS + j]) > mv) { mv = fabs(H[i * NS + j]); P = i; Q = j; }
                if (mv < 1e-15) break;
                double app = H[P * NS + P], aqq = H[Q * NS + Q], apq = H[P * NS + Q];
                double th = fabs(app - aqq) < 1e-30 ? 0.25 * LAP_PI : 0.5 * ::atan(2.0 * apq / (app - aqq));
                double c = ::cos(th), s = ::sin(th);
                for (int i = 0; i < NS; i++) {
                    double aip = H[i * NS + P], aiq = H[i * NS + Q];
                    H[i * NS + P] = c * aip + s * aiq;
                    H[i * NS + Q] = -s * aip + c * aiq;
                }
                for (int j = 0; j < NS; j++) {
                    double apj = H[P * NS + j], aqj = H[Q * NS + j];
                    H[P * NS + j] = c * apj + s * aqj;
                    H[Q * NS + j] = -s * apj + c * aqj;
                }
                H[P * NS + P] = c * c * app + 2.0 * c * s * apq + s * s * aqq;
                H[Q * NS + Q] = s * s * app - 2.0 * c * s * apq + c * c * aqq;
                H[P * NS + Q] = 0.0;
                H[Q * NS + P] = 0.0;
                for (int i = 0; i < NS; i++) {
                    double vip = V[i * NS + P], viq = V[i * NS + Q];
                    V[i * NS + P] = c * vip + s * viq;
                    V[i * NS + Q] = -s * vip + c * viq;
                }
            }
            double* ls = a.out_log_sigma + (int64_t)job * 7;
            for (int i = 0; i < NS; i++) {
                double sm = 0.0;
                for (int k = 0; k < NS; k++) sm += V[i * NS + k] * V[i * NS + k] / fmax(H[k * NS + k], 1e-6);
                ls[i] = 0.5 * ::log(fmax(sm, 1e-30));
            }
            ls[SE] = ::log(c_lp_w[MODEL][SE]);
            for (int i = NP; i < 7; i++) ls[i] = 0.0;
            a.out_elbo[job] = -vals[0];
        }
    }
};

template <int MODEL>
__global__ __launch_bounds__(64) void laplace_kernel(LaplaceArgs a) {
    extern __shared__ double4 lap_lds[];
    __shared__ double vals[LAP_MAX_PTS + 12 * 6];
    LaplaceJob<MODEL> job;
    for (;;) {
        const int t = wave_next_task(a.queue);
        if (t >= a.n_tasks) break;
        job.run(a, lap_lds, vals, a.order[t]);
        wave_reconverge();
    }
}

static const void* lap_ptr(int model) {
    switch (model) {
        case 0: return (const void*)laplace_kernel<0>;
        case 1: return (const void*)laplace_kernel<1>;
        case 2: return (const void*)laplace_kernel<2>;
        case 3: return (const void*)laplace_kernel<3>;
        case 4: return (const void*)laplace_kernel<4>;
        case 5: return (const void*)laplace_kernel<5>;
        case 6: return (const void*)laplace_kernel<6>;
        case 7: return (const void*)laplace_kernel<7>;
        default: return nullptr;
    }
}

hipError_t launch_laplace(int model, const LaplaceArgs& a, int grid, hipStream_t st) {
    const void* k = lap_ptr(model);
    if (!k || grid <= 0 || a.max_n < 0) return hipErrorInvalidValue;
    void* args[] = {const_cast<LaplaceArgs*>(&a)};
    return hipLaunchKernel(k, dim3(grid), dim3(64), args, sizeof(double4) * (size_t)(a.max_n > 0 ? a.max_n : 1), st);
}

hipError_t laplace_occupancy(int model, int max_n, int* blocks_per_cu) {
    const void* k = lap_ptr(model);
    if (!k) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, 64,
                                                         sizeof(double4) * (size_t)(max_n > 0 ? max_n : 1));
}

}  // namespace lcf
