// Mode-coupling matrices on gfx950, FP64: one thread per (l1, l2 >= l1) walks the
// Wigner-3j rows (l1 l2 l3; 0 0 0) and (l1 l2 l3; 2 -2 0) in l3 and accumulates all four
// MASTER blocks at once:
//     M00  = (2 l2 + 1) sum W'_l3 w0^2            M02  = (2 l2 + 1) sum W'_l3 w0 w2
//     M22+ = (2 l2 + 1) sum_{J even} W'_l3 w2^2    M22- = (2 l2 + 1) sum_{J odd} W'_l3 w2^2
// with W'_l = (2l + 1) W_l / (4 pi) (the caller passes W') and J = l1 + l2 + l3 (general
// coupling formula of gmaster/workspaces.py:1069-1077 with (s1, s2, n1, n2) = (0,0,0,0),
// (0,2,0,2), (2,2,2,2)).
//
// Both rows obey the Schulten-Gordon three-term recursion in l3 (m3 = 0):
//     Q(l3+1) f(l3+1) = 4 (2 l3 + 1) f(l3) [w2 only] - Q(l3) f(l3-1),
//     Q(l3)^2 = (l3^2 - (l2-l1)^2) ((l1+l2+1)^2 - l3^2)
// started from the closed forms at l3 = l2 - l1 (w0^2 = g(l1) g(d) / ((2 l2 + 1) g(l2)),
// w2 = w0 sqrt((l2+1)(l2+2) l1 (l1-1) / (l2 (l2-1)(l1+1)(l1+2)))), and for l1 = l2 the
// extra closed form (l l 1; 2 -2 0) = (-1)^l 2 / sqrt(l (l+1)(2l+1)).  In this direction
// the recursion is forward-stable for |m| <= 2 up to l3 = l1 + l2.
//
// Per pair of steps (J even, then J odd): w0 vanishes at odd J, so it advances two degrees
// at a time (w0(l3+2) = -Q(l3+1) w0(l3) / Q(l3+2)); the factors of Q^2 are exact integers
// updated incrementally and 1/Q comes from one rsqrt.  The row sums are symmetric, so each
// thread writes both (l1, l2) and (l2, l1).
#include "common.h"

namespace gm {
namespace {

__global__ void k_coupling(const double* __restrict__ Wp, int lmax_w, const double* __restrict__ g, int L,
                           double* __restrict__ m00, double* __restrict__ m02, double* __restrict__ m22p,
                           double* __restrict__ m22m) {
    const int l1 = blockIdx.y;
    const int l2 = l1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (l2 >= L) return;
    const int d = l2 - l1;
    const double S2 = (double)(l1 + l2 + 1) * (double)(l1 + l2 + 1);
    const double d2 = (double)d * (double)d;
    const int l3max = min(l1 + l2, lmax_w);

    double w0 = sqrt(g[l1] * g[d] / ((2.0 * l2 + 1.0) * g[l2]));
    if (l2 & 1) w0 = -w0;
    double w2 = 0.0;
    if (l1 >= 2)
        w2 = w0 * sqrt((double)(l2 + 1) * (l2 + 2) * l1 * (l1 - 1) /
                       ((double)l2 * (l2 - 1) * (l1 + 1) * (l1 + 2)));
    double a00 = 0.0, a02 = 0.0, aE = 0.0, aO = 0.0;
    // state at the first even step l3: (w0, w2) current, p2 = w2(l3 - 1), Qc = Q(l3)
    int l3 = d;
    double p2 = 0.0, Qc = 0.0;
    if (d == 0) {
        // l3 = 0 (even) and l3 = 1 (odd: w0 = 0) peeled; the l3 = 0 step is degenerate
        const double t0 = Wp[0];
        a00 = t0 * w0 * w0;
        a02 = t0 * w0 * w2;
        aE = t0 * w2 * w2;
        if (l3max < 1) goto done;
        const double w2_1 = (l1 >= 2) ? ((l1 & 1) ? -2.0 : 2.0) / sqrt((double)l1 * (l1 + 1) * (2.0 * l1 + 1.0)) : 0.0;
        aO = Wp[1] * w2_1 * w2_1;
        if (l3max < 2) goto done;
        const double Q1 = sqrt(1.0 * (S2 - 1.0));
        const double Q2 = sqrt(4.0 * (S2 - 4.0));
        const double w0_2 = -Q1 * w0 / Q2;
        const double w2_2 = (12.0 * w2_1 - Q1 * w2) / Q2;
        w0 = w0_2;
        p2 = w2_1;
        w2 = w2_2;
        Qc = Q2;
        l3 = 2;
    }
    if (l3 > l3max) goto done;  // window shorter than l2 - l1: no term
    {
        // A = l3'^2 - d^2 and B = S^2 - l3'^2 at l3' = l3 + 1, exact integers in FP64
        double A = (double)(l3 + 1) * (l3 + 1) - d2;
        double B = S2 - (double)(l3 + 1) * (l3 + 1);
        double c = 4.0 * (2.0 * l3 + 1.0);
        double step = 2.0 * l3 + 3.0;  // (l3+2)^2 - (l3+1)^2
        for (;;) {
            // ---- even step at l3
            const double te = Wp[l3];
            const double tw0 = te * w0;
            a00 = fma(tw0, w0, a00);
            a02 = fma(tw0, w2, a02);
            aE = fma(te * w2, w2, aE);
            if (l3 == l3max) break;
            // advance to l3 + 1 (odd; w0(l3+1) = 0)
            const double q1 = A * B;
            const double r1 = rsqrt(q1);          // 1 / Q(l3+1)
            const double Q1 = q1 * r1;
            const double w2o = fma(c, w2, -Qc * p2) * r1;
            A += step;
            B -= step;
            step += 2.0;
            c += 8.0;
            // ---- odd step at l3 + 1
            const double to = Wp[l3 + 1];
            aO = fma(to * w2o, w2o, aO);
            if (l3 + 1 == l3max) break;
            // advance to l3 + 2 (even)
            const double q2 = A * B;
            const double r2 = rsqrt(q2);          // 1 / Q(l3+2)
            const double Q2 = q2 * r2;
            const double w2e = fma(c, w2o, -Q1 * w2) * r2;
            w0 = -Q1 * w0 * r2;
            A += step;
            B -= step;
            step += 2.0;
            c += 8.0;
            p2 = w2o;
            w2 = w2e;
            Qc = Q2;
            l3 += 2;
        }
    }
done:
    const double c2 = 2.0 * l2 + 1.0, c1 = 2.0 * l1 + 1.0;
    const int64_t i12 = (int64_t)l1 * L + l2, i21 = (int64_t)l2 * L + l1;
    if (m00) { m00[i12] = c2 * a00; m00[i21] = c1 * a00; }
    if (m02) { m02[i12] = c2 * a02; m02[i21] = c1 * a02; }
    if (m22p) { m22p[i12] = c2 * aE; m22p[i21] = c1 * aE; }
    if (m22m) { m22m[i12] = c2 * aO; m22m[i21] = c1 * aO; }
}

}  // namespace
}  // namespace gm

using namespace gm;

extern "C" int gm_coupling(const double* window, int64_t lmax_window, const double* g, int64_t lmax, double* m00,
                           double* m02, double* m22p, double* m22m, void* stream) {
    if (!window || !g) return fail(GM_E_ARG, "gm_coupling: null window / g table");
    if (lmax < 0 || lmax > 65535) return fail(GM_E_ARG, "gm_coupling: bad lmax %lld", (long long)lmax);
    if (!m00 && !m02 && !m22p && !m22m) return fail(GM_E_ARG, "gm_coupling: no output requested");
    if (lmax_window < 0 || lmax_window > 2 * lmax) return fail(GM_E_ARG, "gm_coupling: need 0 <= lmax_window <= 2 lmax");
    const int L = (int)lmax + 1;
    const int bx = 128;
    hipLaunchKernelGGL(k_coupling, dim3((unsigned)((L + bx - 1) / bx), (unsigned)L), dim3(bx), 0,
                       (hipStream_t)stream, window, (int)lmax_window, g, L, m00, m02, m22p, m22m);
    return check_launch("gm_coupling");
}
