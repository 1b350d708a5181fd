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
double erf_approx(double x);

void pso_bounds_gpu(int m, double* lo, double* hi);
void pso_bounds_cpu(int m, double* lo, double* hi);
void population_priors(int m, double* c, double* w);
void prior_params(int m, double* c, double* w);

// Run `fn(i)` for i in [0, n) on up to n_threads threads.
void parallel_for(int64_t n, int n_threads, const std::function<void(int64_t)>& fn);

}  // namespace orc
