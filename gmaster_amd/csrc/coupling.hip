// Mode-coupling matrices on gfx950, FP64: one thread per (l1, l2 >= l1) walks the
// Wigner-3j rows (l1 l2 l3; 0 0 0) and (l1 l2 l3; 2 -2 0) in l3 and accumulates all four
// MASTER blocks at once:
//     M00  = (2 l2 + 1) sum W'_l3 w0^2            M02  = (2 l2 + 1) sum W'_l3 w0 w2
//     M22+ = (2 l2 + 1) sum_{J even} W'_l3 w2^2    M22- = (2 l2 + 1) sum_{J odd} W'_l3 w2^2
// with W'_l = (2l + 1) W_l / (4 pi) and J = l1 + l2 + l3 (general coupling formula of
// gmaster/workspaces.py:1069-1077 with (s1, s2, n1, n2) = (0,0,0,0), (0,2,0,2), (2,2,2,2)).
//
// Both rows obey the Schulten-Gordon three-term recursion in l3 (m3 = 0):
//     Q(l3+1) f(l3+1) = 4 (2 l3 + 1) f(l3) [w2 only] - Q(l3) f(l3-1),
//     Q(l3)^2 = (l3^2 - (l2-l1)^2) ((l1+l2+1)^2 - l3^2)
// started from the closed forms at l3 = l2 - l1 (w0^2 = g(l1) g(d) / ((2 l2 + 1) g(l2)),
// w2 = w0 sqrt((l2+1)(l2+2) l1 (l1-1) / (l2 (l2-1)(l1+1)(l1+2)))), and for l1 = l2 the
// extra closed form (l l 1; 2 -2 0) = (-1)^l 2 / sqrt(l (l+1)(2l+1)).  In this direction
// the recursion is forward-stable for |m| <= 2 up to l3 = l1 + l2.  The row sums are
// symmetric, so each thread writes both (l1, l2) and (l2, l1).
#include "common.h"

namespace gm {
namespace {

__global__ void k_coupling(const double* __restrict__ W, int lmax_w, const double* __restrict__ g, int L,
                           double* __restrict__ m00, double* __restrict__ m02, double* __restrict__ m22p,
                           double* __restrict__ m22m) {
    const int l1 = blockIdx.y;
    const int l2 = l1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (l2 >= L) return;
    const int d = l2 - l1;
    const double S = (double)(l1 + l2 + 1);
    const double dd = (double)d;
    const int l3max = min(l1 + l2, lmax_w);

    double w0 = sqrt(g[l1] * g[d] / ((2.0 * l2 + 1.0) * g[l2]));
    if (l2 & 1) w0 = -w0;
    double w2 = 0.0;
    if (l1 >= 2)
        w2 = w0 * sqrt((double)(l2 + 1) * (l2 + 2) * l1 * (l1 - 1) /
                       ((double)l2 * (l2 - 1) * (l1 + 1) * (l1 + 2)));
    double p0 = 0.0, p2 = 0.0, Qc = 0.0;
    double a00 = 0.0, a02 = 0.0, aE = 0.0, aO = 0.0;
    int l3 = d;
    bool even = ((l1 + l2 + l3) & 1) == 0;
    if (d == 0) {
        // l3 = 0 term, then restart the recursion at l3 = 1 (the l3 = 0 step is degenerate)
        if (l3 <= l3max) {
            const double t = W[0];
            a00 = t * w0 * w0;
            a02 = t * w0 * w2;
            aE = t * w2 * w2;
        }
        p0 = w0;
        p2 = w2;
        w0 = 0.0;
        w2 = (l1 >= 2) ? ((l1 & 1) ? -2.0 : 2.0) / sqrt((double)l1 * (l1 + 1) * (2.0 * l1 + 1.0)) : 0.0;
        l3 = 1;
        Qc = sqrt((1.0 - dd * dd) * (S * S - 1.0));
        even = !even;
    }
    for (; l3 <= l3max; ++l3) {
        const double t = W[l3] * (2.0 * l3 + 1.0);
        const double tw2 = t * w2;
        a00 = fma(t * w0, w0, a00);
        a02 = fma(t * w0, w2, a02);
        if (even) aE = fma(tw2, w2, aE);
        else aO = fma(tw2, w2, aO);
        even = !even;
        if (l3 == l3max) break;
        const double n = (double)(l3 + 1);
        const double Qn = sqrt((n - dd) * (n + dd) * (S - n) * (S + n));
        const double inv = 1.0 / Qn;
        const double n0 = -Qc * p0 * inv;
        const double n2 = (4.0 * (2.0 * l3 + 1.0) * w2 - Qc * p2) * inv;
        p0 = w0;
        p2 = w2;
        w0 = n0;
        w2 = n2;
        Qc = Qn;
    }
    // the l3 = 0 term above was added without its (2 l3 + 1) = 1 factor, which is exact
    const double c2 = (2.0 * l2 + 1.0) / (4.0 * M_PI), c1 = (2.0 * l1 + 1.0) / (4.0 * M_PI);
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
                           double* m02,
                           double* m22p, double* m22m, void* stream) {
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
