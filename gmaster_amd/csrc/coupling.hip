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

// Row sums for one pair l1 <= l2 (see above); results in s[0..3] = (a00, a02, aE, aO).
__device__ __forceinline__ void couple_pair(const double* __restrict__ Wp, int lmax_w, const double* __restrict__ g,
                                            int l1, int l2, double s[4]) {
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
    s[0] = a00;
    s[1] = a02;
    s[2] = aE;
    s[3] = aO;
}

__device__ __forceinline__ void store_pair(int l1, int l2, int L, const double s[4], double* __restrict__ m00,
                                           double* __restrict__ m02, double* __restrict__ m22p,
                                           double* __restrict__ m22m) {
    const double c2 = 2.0 * l2 + 1.0, c1 = 2.0 * l1 + 1.0;
    const int64_t i12 = (int64_t)l1 * L + l2, i21 = (int64_t)l2 * L + l1;
    if (m00) { m00[i12] = c2 * s[0]; m00[i21] = c1 * s[0]; }
    if (m02) { m02[i12] = c2 * s[1]; m02[i21] = c1 * s[1]; }
    if (m22p) { m22p[i12] = c2 * s[2]; m22p[i21] = c1 * s[2]; }
    if (m22m) { m22m[i12] = c2 * s[3]; m22m[i21] = c1 * s[3]; }
}

__global__ void k_coupling(const double* __restrict__ Wp, int lmax_w, const double* __restrict__ g, int L,
                           double* __restrict__ m00, double* __restrict__ m02, double* __restrict__ m22p,
                           double* __restrict__ m22m) {
    const int l1 = blockIdx.y;
    const int l2 = l1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (l2 >= L) return;
    double s[4];
    couple_pair(Wp, lmax_w, g, l1, l2, s);
    store_pair(l1, l2, L, s, m00, m02, m22p, m22m);
}

// The four row sums (without the column factor) of each listed pair, vals[k][0..3].
__global__ void k_coupling_pair_values(const double* __restrict__ Wp, int lmax_w, const double* __restrict__ g,
                                       const int32_t* __restrict__ lo, const int32_t* __restrict__ hi,
                                       int64_t npairs, double* __restrict__ vals) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npairs) return;
    double s[4];
    couple_pair(Wp, lmax_w, g, lo[k], hi[k], s);
    for (int q = 0; q < 4; ++q) vals[k * 4 + q] = s[q];
}

// Only the listed pairs (lo[k] <= hi[k]); every other element of the outputs is left as it was.
__global__ void k_coupling_pairs(const double* __restrict__ Wp, int lmax_w, const double* __restrict__ g, int L,
                                 const int32_t* __restrict__ lo, const int32_t* __restrict__ hi, int64_t npairs,
                                 double* __restrict__ m00, double* __restrict__ m02, double* __restrict__ m22p,
                                 double* __restrict__ m22m) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npairs) return;
    const int l1 = lo[k], l2 = hi[k];
    double s[4];
    couple_pair(Wp, lmax_w, g, l1, l2, s);
    store_pair(l1, l2, L, s, m00, m02, m22p, m22m);
}

// Toeplitz approximation of one coupling block (Louis et al. 2020, NaMaster's fill order as
// in gmaster/workspaces.py:1141-1194), on the master coefficients S = M / (2 l' + 1).  The
// reference applies a sequence of overlapping fills to the lower triangle and then mirrors
// it; here each output element evaluates the last fill that covers it:
//   pure(i, j) = c_t[(i - j + lt) mod n] sqrt(D_i D_j),  c_t[k] = S[k][lt] / sqrt(D_k D_lt)
//   j <= le or i - j <= dl -> exact;  i, j >= lt -> pure;  i == j -> exact;
//   1 <= i - j <= dl with j <= lt -> exact;
//   le <= j <= lt, i >= n - 1 - lt + j -> c_e[le + i - j] sqrt(D_i D_j),  c_e[k] = S[k][le] / sqrt(D_k D_le)
// for i = max(row, col), j = min(row, col); band and diagonal read S from the upper
// triangle exactly as the reference's diag(S, offset) does.
__global__ void k_toeplitz(const double* __restrict__ M, int n, int lt, int le, int dl, double* __restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    if (b >= n) return;
    const int i = a > b ? a : b, j = a > b ? b : a;
    auto S = [&](int r, int c) { return M[(int64_t)r * n + c] / (2.0 * c + 1.0); };
    auto D = [&](int k) { return fabs(S(k, k)); };
    const double Di = D(i), Dj = D(j);
    double v;
    {
        const int k = (i - j + lt) % n;
        const double den = sqrt(D(k) * D(lt));
        v = (den > 0.0 ? S(k, lt) / den : 0.0) * sqrt(Di * Dj);
    }
    if (j <= le) v = S(i, j);
    if (i - j <= dl) v = (i == j) ? S(i, i) : S(j, i);
    if (i >= lt && j >= lt) {
        const int k = (i - j + lt) % n;
        const double den = sqrt(D(k) * D(lt));
        v = (den > 0.0 ? S(k, lt) / den : 0.0) * sqrt(Di * Dj);
    }
    if (i == j) v = S(i, i);
    if (i - j >= 1 && i - j <= dl && j <= lt) v = S(j, i);
    if (j >= le && j <= lt && i >= n - 1 - lt + j) {
        const int k = le + i - j;
        const double Dk = D(k);
        const double den = sqrt(Dk * D(le));
        const double ce = Dk > 0.0 ? S(k, le) / den : 0.0;  // guarded by D_k only, as the reference
        v = ce * sqrt(Dj * Di);
    }
    out[(int64_t)a * n + b] = v * (2.0 * b + 1.0);
}

// Which fill of the Toeplitz approximation sets element (i >= j): 0 = an exact value,
// 1 = the pure Toeplitz extension, 2 = the corner extrapolation from column l_exact (the
// same precedence as k_toeplitz).
__device__ __forceinline__ int toeplitz_kind(int i, int j, int n, int lt, int le, int dl) {
    int kind = 1;
    if (j <= le) kind = 0;
    if (i - j <= dl) kind = 0;
    if (i >= lt && j >= lt) kind = 1;
    if (i == j) kind = 0;
    if (i - j >= 1 && i - j <= dl && j <= lt) kind = 0;
    if (j >= le && j <= lt && i >= n - 1 - lt + j) kind = 2;
    return kind;
}

// Learned correction of a Toeplitz-approximated block M (column factor included), in place:
// for every element not set exactly, with j = min(row, col) and d = |row - col|,
//     M += e(j, d) sqrt(D_j D_{j+d}) (2 col + 1),   D_k = |M_kk| / (2k + 1),
// e bilinear in (log j, log d) over the grid lg[0..nl) x dg[0..nd) (increasing integers)
// and zero outside it.
__global__ void k_toeplitz_correct(double* __restrict__ M, int n, int lt, int le, int dl,
                                   const int32_t* __restrict__ lg, int nl, const int32_t* __restrict__ dg, int nd,
                                   const float* __restrict__ e) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    if (b >= n) return;
    const int i = a > b ? a : b, j = a > b ? b : a, d = i - j;
    if (j < lg[0] || j > lg[nl - 1] || d < dg[0] || d > dg[nd - 1]) return;
    if (toeplitz_kind(i, j, n, lt, le, dl) == 0) return;
    int p = 0, q = 0;
    while (p + 2 < nl && lg[p + 1] <= j) ++p;
    while (q + 2 < nd && dg[q + 1] <= d) ++q;
    const double tl = nl > 1 ? (log((double)j) - log((double)lg[p])) / (log((double)lg[p + 1]) - log((double)lg[p])) : 0.0;
    const double td = nd > 1 ? (log((double)d) - log((double)dg[q])) / (log((double)dg[q + 1]) - log((double)dg[q])) : 0.0;
    const int p1 = nl > 1 ? p + 1 : p, q1 = nd > 1 ? q + 1 : q;
    const double v = (1.0 - tl) * ((1.0 - td) * e[p * nd + q] + td * e[p * nd + q1]) +
                     tl * ((1.0 - td) * e[p1 * nd + q] + td * e[p1 * nd + q1]);
    const double Di = fabs(M[(int64_t)i * n + i]) / (2.0 * i + 1.0);
    const double Dj = fabs(M[(int64_t)j * n + j]) / (2.0 * j + 1.0);
    M[(int64_t)a * n + b] += v * sqrt(Di * Dj) * (2.0 * b + 1.0);
}

}  // namespace
}  // namespace gm

using namespace gm;

extern "C" int gm_coupling_pair_values(const double* window, int64_t lmax_window, const double* g, int64_t lmax,
                                       const int32_t* lo, const int32_t* hi, int64_t npairs, double* vals,
                                       void* stream) {
    if (!window || !g || !vals || (npairs > 0 && (!lo || !hi)))
        return fail(GM_E_ARG, "gm_coupling_pair_values: null pointer");
    if (lmax < 0 || lmax > 65535) return fail(GM_E_ARG, "gm_coupling_pair_values: bad lmax %lld", (long long)lmax);
    if (lmax_window < 0 || lmax_window > 2 * lmax)
        return fail(GM_E_ARG, "gm_coupling_pair_values: need 0 <= lmax_window <= 2 lmax");
    if (npairs < 0) return fail(GM_E_ARG, "gm_coupling_pair_values: negative pair count");
    if (npairs == 0) return GM_OK;
    hipLaunchKernelGGL(k_coupling_pair_values, dim3((unsigned)((npairs + 127) / 128)), dim3(128), 0,
                       (hipStream_t)stream, window, (int)lmax_window, g, lo, hi, npairs, vals);
    return check_launch("gm_coupling_pair_values");
}

extern "C" int gm_toeplitz_correct(double* M, int64_t lmax, int64_t l_toeplitz, int64_t l_exact, int64_t dl_band,
                                   const int32_t* l_grid, int64_t n_l, const int32_t* d_grid, int64_t n_d,
                                   const float* e, void* stream) {
    if (!M || !l_grid || !d_grid || !e) return fail(GM_E_ARG, "gm_toeplitz_correct: null pointer");
    if (lmax < 1 || l_toeplitz <= 0 || l_exact <= 0 || dl_band < 0 || l_exact > l_toeplitz || l_toeplitz >= lmax ||
        dl_band >= lmax)
        return fail(GM_E_ARG, "gm_toeplitz_correct: need 0 < l_exact <= l_toeplitz < lmax and 0 <= dl_band < lmax");
    if (n_l < 1 || n_d < 1 || n_l > 4096 || n_d > 4096) return fail(GM_E_ARG, "gm_toeplitz_correct: bad grid sizes");
    const int n = (int)lmax + 1;
    hipLaunchKernelGGL(k_toeplitz_correct, dim3((unsigned)((n + 255) / 256), (unsigned)n), dim3(256), 0,
                       (hipStream_t)stream, M, n, (int)l_toeplitz, (int)l_exact, (int)dl_band, l_grid, (int)n_l,
                       d_grid, (int)n_d, e);
    return check_launch("gm_toeplitz_correct");
}

extern "C" int gm_toeplitz(const double* M, int64_t lmax, int64_t l_toeplitz, int64_t l_exact, int64_t dl_band,
                           double* out, void* stream) {
    if (!M || !out) return fail(GM_E_ARG, "gm_toeplitz: null pointer");
    if (M == out) return fail(GM_E_ARG, "gm_toeplitz: out must not alias M");
    if (lmax < 1 || l_toeplitz <= 0 || l_exact <= 0 || dl_band < 0 || l_exact > l_toeplitz ||
        l_toeplitz >= lmax || dl_band >= lmax)
        return fail(GM_E_ARG, "gm_toeplitz: need 0 < l_exact <= l_toeplitz < lmax and 0 <= dl_band < lmax");
    const int n = (int)lmax + 1;
    hipLaunchKernelGGL(k_toeplitz, dim3((unsigned)((n + 255) / 256), (unsigned)n), dim3(256), 0, (hipStream_t)stream,
                       M, n, (int)l_toeplitz, (int)l_exact, (int)dl_band, out);
    return check_launch("gm_toeplitz");
}

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

extern "C" int gm_coupling_pairs(const double* window, int64_t lmax_window, const double* g, int64_t lmax,
                                 const int32_t* lo, const int32_t* hi, int64_t npairs, double* m00, double* m02,
                                 double* m22p, double* m22m, void* stream) {
    if (!window || !g || (npairs > 0 && (!lo || !hi))) return fail(GM_E_ARG, "gm_coupling_pairs: null pointer");
    if (lmax < 0 || lmax > 65535) return fail(GM_E_ARG, "gm_coupling_pairs: bad lmax %lld", (long long)lmax);
    if (!m00 && !m02 && !m22p && !m22m) return fail(GM_E_ARG, "gm_coupling_pairs: no output requested");
    if (lmax_window < 0 || lmax_window > 2 * lmax)
        return fail(GM_E_ARG, "gm_coupling_pairs: need 0 <= lmax_window <= 2 lmax");
    if (npairs < 0) return fail(GM_E_ARG, "gm_coupling_pairs: negative pair count");
    if (npairs == 0) return GM_OK;
    const int L = (int)lmax + 1;
    const int bx = 128;
    hipLaunchKernelGGL(k_coupling_pairs, dim3((unsigned)((npairs + bx - 1) / bx)), dim3(bx), 0, (hipStream_t)stream,
                       window, (int)lmax_window, g, L, lo, hi, npairs, m00, m02, m22p, m22m);
    return check_launch("gm_coupling_pairs");
}
