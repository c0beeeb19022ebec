// Pure-E/B coupling blocks on gfx950, FP64, by Gauss-Legendre quadrature over the
// Wigner-d functions of the field and of the mask's spin-weighted derivatives
// (gmaster/workspaces.py:928-1052):
//   * Gauss-Legendre nodes and weights of order N by Newton iteration on P_N (one thread
//     per root, the three-term recurrence for P_N and P_{N-1});
//   * Wigner d^l_{mn}(beta) tables at the nodes, l-major (row l holds every node), from
//     the Jacobi-polynomial recurrence in l, optionally scaled per l;
//   * the mask correlations c(x) = sum_l d^l_{mn}(x) a_l and s(x) = sum_l (-1)^l d^l_{mn}(x) a_l
//     without a table;
//   * the combinations sum_k coef_k u_k[node] v_k[l] T_k[l][node] (the "left" tables, with
//     the quadrature weight folded in), whose products with the right tables are plain FP64
//     GEMMs; and the parity split of the spin-2 blocks.
#include <math.h>

#include "common.h"

namespace gm {
namespace {

// Gauss-Legendre root i (ascending nodes x_0 < ... < x_{N-1}) and its weight
__global__ void k_gauss_legendre(int N, double* __restrict__ x, double* __restrict__ w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    // root k (descending) from the standard asymptotic guess; stored ascending
    const int k = N - 1 - i;
    double z = cos(M_PI * (k + 0.75) / (N + 0.5));
    double dp = 1.0;
    // Newton converges quadratically from these guesses: once a step is below 1e-11 the
    // root is at rounding level, and one more step polishes it
    bool last = false;
    for (int it = 0; it < 50; ++it) {
        double p0 = 1.0, p1 = z;
        for (int n = 2; n <= N; ++n) {
            const double p2 = ((2.0 * n - 1.0) * z * p1 - (n - 1.0) * p0) / n;
            p0 = p1;
            p1 = p2;
        }
        // p1 = P_N(z), p0 = P_{N-1}(z)
        dp = N * (z * p1 - p0) / (z * z - 1.0);
        const double dz = p1 / dp;
        z -= dz;
        if (last) break;
        last = fabs(dz) < 1e-11;
    }
    {
        double p0 = 1.0, p1 = z;
        for (int n = 2; n <= N; ++n) {
            const double p2 = ((2.0 * n - 1.0) * z * p1 - (n - 1.0) * p0) / n;
            p0 = p1;
            p1 = p2;
        }
        dp = N * (z * p1 - p0) / (z * z - 1.0);
    }
    x[i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
}

// d^l_{mn}(beta_node), cos(beta_node) = x[node], for l = lmin .. lmax through the Jacobi
// recurrence in k = l - lmin with (mu, nu) = (|m - n|, |m + n|), returned one degree at a time.
struct DState {
    double pre, norm, prev, cur, x;
    int mu, nu, lmin;
};

__device__ inline DState d_start(double x, int m, int n) {
    DState s;
    s.mu = abs(m - n);
    s.nu = abs(m + n);
    s.lmin = max(abs(m), abs(n));
    const double sh = sqrt(fmax(0.5 * (1.0 - x), 0.0)), ch = sqrt(fmax(0.5 * (1.0 + x), 0.0));
    // (-1)^((n - m + mu) / 2) sin^mu(beta/2) cos^nu(beta/2)
    const int e = (n - m + s.mu) / 2;
    s.pre = ((e & 1) ? -1.0 : 1.0) * pow(sh, (double)s.mu) * pow(ch, (double)s.nu);
    // sqrt(k! (k+mu+nu)! / ((k+mu)! (k+nu)!)) at k = 0, i.e. sqrt((mu+nu)! / (mu! nu!))
    double f = 1.0;
    for (int i = 1; i <= s.mu + s.nu; ++i) f *= i;
    for (int i = 2; i <= s.mu; ++i) f /= i;
    for (int i = 2; i <= s.nu; ++i) f /= i;
    s.norm = sqrt(f);
    s.prev = 0.0;
    s.cur = 1.0;  // P_0
    s.x = x;
    return s;
}

// advance from P_k to P_{k+1} (k >= 0); the normalisation follows
__device__ inline void d_step(DState& s, int k) {
    const double mu = s.mu, nu = s.nu, t = mu + nu;
    const int kk = k + 1;  // the degree being formed
    double next;
    if (kk == 1) {
        next = 0.5 * ((mu - nu) + (t + 2.0) * s.x);
    } else {
        const double o = kk;
        next = ((2 * o + t - 1) * ((2 * o + t) * (2 * o + t - 2) * s.x + mu * mu - nu * nu) * s.cur -
                2 * (o + mu - 1) * (o + nu - 1) * (2 * o + t) * s.prev) /
               (2 * o * (o + t) * (2 * o + t - 2));
    }
    s.prev = s.cur;
    s.cur = next;
    s.norm *= sqrt((double)kk * (kk + t) / ((kk + mu) * (kk + nu)));
}

__global__ void k_wigner_d_table(const double* __restrict__ x, int nnodes, int m, int n, int lmax,
                                 const double* __restrict__ lscale, double* __restrict__ out) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= nnodes) return;
    DState s = d_start(x[node], m, n);
    for (int l = 0; l < s.lmin && l <= lmax; ++l) out[(int64_t)l * nnodes + node] = 0.0;
    for (int l = s.lmin; l <= lmax; ++l) {
        if (l > s.lmin) d_step(s, l - s.lmin - 1);
        const double d = s.pre * s.norm * s.cur;
        out[(int64_t)l * nnodes + node] = lscale ? d * lscale[l] : d;
    }
}

__global__ void k_wigner_d_corr(const double* __restrict__ x, int nnodes, int m, int n, int lmax,
                                const double* __restrict__ coef, double* __restrict__ c, double* __restrict__ sgn) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= nnodes) return;
    DState s = d_start(x[node], m, n);
    double cs = 0.0, ss = 0.0;
    for (int l = s.lmin; l <= lmax; ++l) {
        if (l > s.lmin) d_step(s, l - s.lmin - 1);
        const double v = s.pre * s.norm * s.cur * coef[l];
        cs += v;
        ss += (l & 1) ? -v : v;
    }
    c[node] = cs;
    if (sgn) sgn[node] = ss;
}

constexpr int kMaxTerms = 8;
struct Terms {
    const double* T[kMaxTerms];  // (L x nnodes) tables
    const double* u[kMaxTerms];  // per-node factors (NULL = 1)
    const double* v[kMaxTerms];  // per-degree factors (NULL = 1)
    double coef[kMaxTerms];
    int n;
};

__global__ void k_table_combine(Terms terms, const double* __restrict__ wnode, int L, int nnodes,
                                double* __restrict__ out) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    if (node >= nnodes) return;
    const int64_t i = (int64_t)l * nnodes + node;
    double s = 0.0;
    for (int k = 0; k < terms.n; ++k) {
        double t = terms.coef[k] * terms.T[k][i];
        if (terms.u[k]) t *= terms.u[k][node];
        if (terms.v[k]) t *= terms.v[k][l];
        s += t;
    }
    out[i] = wnode ? s * wnode[node] : s;
}

// even = (total + (-1)^(l1+l2) signed) / 2, odd = (total - (-1)^(l1+l2) signed) / 2
__global__ void k_parity_split(const double* __restrict__ total, const double* __restrict__ sgnd, int L,
                               double* __restrict__ even, double* __restrict__ odd) {
    const int l2 = blockIdx.x * blockDim.x + threadIdx.x;
    const int l1 = blockIdx.y;
    if (l2 >= L) return;
    const int64_t i = (int64_t)l1 * L + l2;
    const double s = ((l1 + l2) & 1) ? -sgnd[i] : sgnd[i];
    even[i] = 0.5 * (total[i] + s);
    odd[i] = 0.5 * (total[i] - s);
}

}  // namespace
}  // namespace gm

using namespace gm;

extern "C" int gm_gauss_legendre(int64_t order, double* x, double* w, void* stream) {
    if (order < 1 || order > (1 << 20)) return fail(GM_E_ARG, "gm_gauss_legendre: bad order");
    if (!x || !w) return fail(GM_E_ARG, "gm_gauss_legendre: null pointer");
    hipLaunchKernelGGL(k_gauss_legendre, dim3((unsigned)((order + 127) / 128)), dim3(128), 0, (hipStream_t)stream,
                       (int)order, x, w);
    return check_launch("gm_gauss_legendre");
}

extern "C" int gm_wigner_d_table(const double* x, int64_t nnodes, int64_t m, int64_t n, int64_t lmax,
                                 const double* lscale, double* out, void* stream) {
    if (!x || !out) return fail(GM_E_ARG, "gm_wigner_d_table: null pointer");
    if (nnodes < 1 || lmax < 0 || m < -16 || m > 16 || n < -16 || n > 16)
        return fail(GM_E_ARG, "gm_wigner_d_table: bad sizes / spins (|m|, |n| <= 16)");
    hipLaunchKernelGGL(k_wigner_d_table, dim3((unsigned)((nnodes + 127) / 128)), dim3(128), 0, (hipStream_t)stream, x,
                       (int)nnodes, (int)m, (int)n, (int)lmax, lscale, out);
    return check_launch("gm_wigner_d_table");
}

extern "C" int gm_wigner_d_corr(const double* x, int64_t nnodes, int64_t m, int64_t n, int64_t lmax,
                                const double* coef, double* c, double* s, void* stream) {
    if (!x || !coef || !c) return fail(GM_E_ARG, "gm_wigner_d_corr: null pointer");
    if (nnodes < 1 || lmax < 0 || m < -16 || m > 16 || n < -16 || n > 16)
        return fail(GM_E_ARG, "gm_wigner_d_corr: bad sizes / spins (|m|, |n| <= 16)");
    hipLaunchKernelGGL(k_wigner_d_corr, dim3((unsigned)((nnodes + 127) / 128)), dim3(128), 0, (hipStream_t)stream, x,
                       (int)nnodes, (int)m, (int)n, (int)lmax, coef, c, s);
    return check_launch("gm_wigner_d_corr");
}

extern "C" int gm_table_combine(int64_t nterms, const double* const* tables, const double* const* node_factors,
                                const double* const* degree_factors, const double* coef, const double* node_weight,
                                int64_t L, int64_t nnodes, double* out, void* stream) {
    if (nterms < 1 || nterms > kMaxTerms || !tables || !coef || !out)
        return fail(GM_E_ARG, "gm_table_combine: bad terms / null pointer");
    if (L < 1 || nnodes < 1 || L > 65535) return fail(GM_E_ARG, "gm_table_combine: bad sizes");
    Terms t{};
    t.n = (int)nterms;
    for (int k = 0; k < t.n; ++k) {
        if (!tables[k]) return fail(GM_E_ARG, "gm_table_combine: null table %d", k);
        t.T[k] = tables[k];
        t.u[k] = node_factors ? node_factors[k] : nullptr;
        t.v[k] = degree_factors ? degree_factors[k] : nullptr;
        t.coef[k] = coef[k];
    }
    hipLaunchKernelGGL(k_table_combine, dim3((unsigned)((nnodes + 255) / 256), (unsigned)L), dim3(256), 0,
                       (hipStream_t)stream, t, node_weight, (int)L, (int)nnodes, out);
    return check_launch("gm_table_combine");
}

extern "C" int gm_parity_split(const double* total, const double* sgnd, int64_t L, double* even, double* odd,
                               void* stream) {
    if (!total || !sgnd || !even || !odd) return fail(GM_E_ARG, "gm_parity_split: null pointer");
    if (L < 1 || L > 65535) return fail(GM_E_ARG, "gm_parity_split: bad size");
    hipLaunchKernelGGL(k_parity_split, dim3((unsigned)((L + 255) / 256), (unsigned)L), dim3(256), 0,
                       (hipStream_t)stream, total, sgnd, (int)L, even, odd);
    return check_launch("gm_parity_split");
}
