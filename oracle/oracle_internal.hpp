// oracle_internal.hpp -- shared declarations inside the oracle library.
// TEST INFRASTRUCTURE ONLY.
#pragma once
#include <stdint.h>
#include <functional>

namespace orc {

int n_params(int m);
double eval_scalar(int m, const double* p, double t);
void eval_times(int m, const double* p, const double* t, int64_t n, double* out);
double metzger_single(const double* p, double t);
int eval_grad(int m, const double* p, double t, double* o);

double log_normal_cdf(double x);
double dlog_normal_cdf(double x);  // exact derivative of log_normal_cdf
double erf_approx(double x);

void pso_bounds_gpu(int m, double* lo, double* hi);
void pso_bounds_cpu(int m, double* lo, double* hi);
void population_priors(int m, double* c, double* w);
void prior_params(int m, double* c, double* w);

// xoshiro128** (Blackman and Vigna, "Scrambled linear pseudorandom number
// generators", ACM TOMS 2021): the per-particle stream of the swarms'
// velocity updates, seeded from one Philox4x32-10 block (an all-zero block,
// the one state the generator cannot leave, becomes {1, 0, 0, 0}).
struct Xoshiro128ss {
    uint32_t s[4];
    void seed(const uint32_t r[4]) {
        for (int i = 0; i < 4; i++) s[i] = r[i];
        if ((s[0] | s[1] | s[2] | s[3]) == 0u) s[0] = 1u;
    }
    static uint32_t rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
    uint32_t next() {
        const uint32_t out = rotl(s[1] * 5u, 7) * 9u;
        const uint32_t t = s[1] << 9;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 11);
        return out;
    }
};

// Run `fn(i)` for i in [0, n) on up to n_threads threads.
void parallel_for(int64_t n, int n_threads, const std::function<void(int64_t)>& fn);

}  // namespace orc
