// Upper tail of the chi-square distribution in FP64 (device and host): gamma_q / chisq_prob of
// src/lmm_engine.cpp:22-64 (the same routine is repeated in src/mvlmm_engine.cpp:12-52):
// series for x < a + 1, Lentz continued fraction otherwise, relative tolerance 1e-7, at most 100
// terms, 0 when neither converges; stat <= 0 gives 1.
#pragma once
#include <hip/hip_runtime.h>

namespace csamd {

__host__ __device__ inline double d_gamma_q(double a, double x) {
    if (x < 0.0 || a <= 0.0) return 1.0;
    if (x < a + 1.0) {
        double ap = a, sum = 1.0 / a, del = sum;
        for (int i = 0; i < 100; ++i) {
            ap += 1.0;
            del *= x / ap;
            sum += del;
            if (fabs(del) < fabs(sum) * 1e-7) return 1.0 - sum * exp(-x + a * log(x) - lgamma(a));
        }
    } else {
        double b = x + 1.0 - a, c = 1.0 / 1e-30, d = 1.0 / b, h = d;
        for (int i = 1; i <= 100; ++i) {
            const double an = -i * (i - a);
            b += 2.0;
            d = an * d + b;
            if (fabs(d) < 1e-30) d = 1e-30;
            c = b + an / c;
            if (fabs(c) < 1e-30) c = 1e-30;
            d = 1.0 / d;
            const double del = d * c;
            h *= del;
            if (fabs(del - 1.0) < 1e-7) return exp(-x + a * log(x) - lgamma(a)) * h;
        }
    }
    return 0.0;
}

__host__ __device__ inline double d_chisq_prob(double stat, int df) {
    if (stat <= 0) return 1.0;
    return d_gamma_q(df / 2.0, stat / 2.0);
}

}  // namespace csamd
