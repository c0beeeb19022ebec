// Pieces of the mu-rescaled Legendre recurrence shared by the latitudinal kernels:
// extended-exponent constants, seeds lambda_M as (mu, k) and the lane rescaling.
#pragma once

#include "common.h"

namespace gm {
namespace {

// Extended-exponent constants per arithmetic type: a lane whose seed lies below 2^thr
// starts with k < 0 and mu = lambda 2^(unit |k|) < 2^inactive; it is rescaled by 2^-unit
// whenever |mu| passes 2^inactive.  FP32 keeps inactive lanes below 2^-36, far under its
// 2^-24 resolution, and active seeds above 2^-100, inside its normal range.
template <typename F> struct ExtExp;
template <> struct ExtExp<double> {
    static constexpr double inactive_max = 0x1p-300, rescale = 0x1p-600;
    static constexpr int thr = -900, unit = 600;
};
template <> struct ExtExp<float> {
    static constexpr float inactive_max = 0x1p-36f, rescale = 0x1p-64f;
    static constexpr int thr = -100, unit = 64;
};

// lambda_M for order m at a ring as (mu, k): mant * 2^e from the host constant and
// the powers sin^(m+s)(theta/2) cos^|m-s|(theta/2), evaluated in FP64.
template <typename F>
__device__ inline void seed_lane(double smant, int sexp, int a, int b, double shalf, double chalf, F& mu, int& k) {
    double m1, m2;
    int e1, e2;
    pow_ext(shalf, a, m1, e1);
    pow_ext(chalf, b, m2, e2);
    const double v = smant * m1 * m2;
    const int E = sexp + e1 + e2;
    constexpr int thr = ExtExp<F>::thr, unit = ExtExp<F>::unit;
    if (E >= thr) {
        k = 0;
        mu = (F)ldexp(v, E);
    } else {
        const int kk = (thr - E + unit - 1) / unit;
        k = -kk;
        mu = (F)ldexp(v, E + unit * kk);
    }
}

// One rescaling step of an inactive lane (k < 0) whose |mu| has passed 2^inactive: a
// select of the factor and two multiplies (exact: the result stays a normal number).
template <typename F>
__device__ inline void rescale_lane(F& mu0, F& mu1, int& k) {
    const bool up = k < 0 && fabs(mu0) > ExtExp<F>::inactive_max;
    const F f = up ? ExtExp<F>::rescale : F(1);
    mu0 *= f;
    mu1 *= f;
    k += up ? 1 : 0;
}

// Rescaling of R lanes' recurrences; anyinact (wave-uniform) skips it once every lane is
// active.  Callers check every 4 degrees, or every 8 where m <= 12288: over 8 degrees from
// l0 > m a lane grows by at most prod_{k=1..8} (1 + alpha_{m+k}) < 2^55 (alpha_{m+k}^2 <
// 4 (m + k) / k), so an inactive FP64 lane found just above 2^-300 had stayed below
// 2^-245: an FP32 zero, and negligible in FP64.
template <int R, typename F>
__device__ inline void rescale(F (&mu0)[R], F (&mu1)[R], int (&kx)[R], bool& anyinact) {
    if (!anyinact) return;
    bool still = false;
#pragma unroll
    for (int t = 0; t < R; ++t) {
        rescale_lane(mu0[t], mu1[t], kx[t]);
        still = still || (kx[t] < 0);
    }
    anyinact = __any(still);
}

}  // namespace
}  // namespace gm
