// Contaminant deprojection and E/B purification on gfx950, FP64: weighted map inner
// products (the template covariance and the template-map products), the subtraction of
// the best-fit templates, the per-multipole filter of alms by a matrix of spectra, and the
// spin-weighted map products of the purification.  All are HBM-bound streams over
// (nmaps x npix) maps or healpy-packed alms.
#include "common.h"

namespace gm {
namespace {

constexpr int kDotThreads = 256;
constexpr int kDotPerThread = 8;  // elements per thread and block pass

int dot_blocks(int64_t n) {
    const int64_t per = (int64_t)kDotThreads * kDotPerThread;
    int64_t nb = (n + per - 1) / per;
    if (nb < 1) nb = 1;
    if (nb > 1024) nb = 1024;
    return (int)nb;
}

// partial[pair][blk] = sum over this block's strided share of p of a_i[p] b_j[p] v[p mod period]
__global__ void k_dot_partial(const double* __restrict__ a, int64_t lda, const double* __restrict__ b,
                              int64_t ldb, int nb_vec, const double* __restrict__ v, int64_t period, int64_t n,
                              double* __restrict__ partial) {
    __shared__ double red[kDotThreads / 64];
    const int pair = blockIdx.y;
    const int i = pair / nb_vec, j = pair % nb_vec;
    const double* x = a + (int64_t)i * lda;
    const double* y = b + (int64_t)j * ldb;
    double s = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        const double t = x[p] * y[p];
        s += v ? t * v[p % period] : t;
    }
    for (int msk = 32; msk >= 1; msk >>= 1) s += __shfl_xor(s, msk);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kDotThreads / 64; ++w) t += red[w];
        partial[(int64_t)pair * gridDim.x + blockIdx.x] = t;
    }
}

// out[pair] = w * sum_blk partial[pair][blk], summed in a fixed order (deterministic)
__global__ void k_dot_final(const double* __restrict__ partial, int nblk, int npair, double w,
                            double* __restrict__ out) {
    const int pair = blockIdx.x * blockDim.x + threadIdx.x;
    if (pair >= npair) return;
    double s = 0.0;
    for (int k = 0; k < nblk; ++k) s += partial[(int64_t)pair * nblk + k];
    out[pair] = w * s;
}

// maps[p] -= sum_i alpha[i] t_i[p]
__global__ void k_deproject(double* __restrict__ maps, int64_t n, const double* __restrict__ t, int ntemp,
                            int64_t ldt, const double* __restrict__ alpha) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        double s = 0.0;
        for (int i = 0; i < ntemp; ++i) s += alpha[i] * t[(int64_t)i * ldt + p];
        maps[p] -= s;
    }
}

// out[o][(l, m)] = sum_k cl[k][o][l] alm[k][(l, m)]; one workgroup row per order m
__global__ void k_alm_filter(const double2* __restrict__ alm, int nin, int64_t astride, const double* __restrict__ cl,
                             int nout, int L, double2* __restrict__ out, int64_t ostride) {
    const int m = blockIdx.y;
    const int lmax = L - 1;
    for (int l = m + blockIdx.x * blockDim.x + threadIdx.x; l < L; l += gridDim.x * blockDim.x) {
        const int64_t ia = alm_index(l, m, lmax);
        for (int o = 0; o < nout; ++o) {
            double2 s = make_double2(0.0, 0.0);
            for (int k = 0; k < nin; ++k) {
                const double c = cl[((int64_t)k * nout + o) * L + l];
                const double2 a = alm[k * astride + ia];
                s.x += c * a.x;
                s.y += c * a.y;
            }
            out[o * ostride + ia] = s;
        }
    }
}

}  // namespace
}  // namespace gm

using namespace gm;

extern "C" int64_t gm_map_dots_work_size(int64_t na, int64_t nb, int64_t n) {
    if (na < 1 || nb < 1 || n < 1) return 0;
    return na * nb * (int64_t)dot_blocks(n);
}

extern "C" int gm_map_dots(const double* a, int64_t na, int64_t lda, const double* b, int64_t nb, int64_t ldb,
                           int64_t n, const double* weight_map, int64_t period, double w, double* out, double* work,
                           void* stream) {
    if (!a || !b || !out || !work) return fail(GM_E_ARG, "gm_map_dots: null pointer");
    if (na < 1 || nb < 1 || n < 1 || na * nb > 65535) return fail(GM_E_ARG, "gm_map_dots: bad sizes");
    if ((na > 1 && lda < n) || (nb > 1 && ldb < n)) return fail(GM_E_SHAPE, "gm_map_dots: leading dimension < n");
    if (weight_map && (period < 1 || n % period)) return fail(GM_E_SHAPE, "gm_map_dots: n must be a multiple of period");
    const int nblk = dot_blocks(n);
    const int npair = (int)(na * nb);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_dot_partial, dim3(nblk, npair), dim3(kDotThreads), 0, st, a, lda, b, ldb, (int)nb,
                       weight_map, period, n, work);
    hipLaunchKernelGGL(k_dot_final, dim3((npair + 63) / 64), dim3(64), 0, st, work, nblk, npair, w, out);
    return check_launch("gm_map_dots");
}

extern "C" int gm_map_deproject(double* maps, int64_t n, const double* templates, int64_t ntemp, int64_t ldt,
                                const double* alphas, void* stream) {
    if (!maps || (ntemp > 0 && (!templates || !alphas))) return fail(GM_E_ARG, "gm_map_deproject: null pointer");
    if (n < 0 || ntemp < 0 || ntemp > 4096) return fail(GM_E_ARG, "gm_map_deproject: bad sizes");
    if (ntemp > 1 && ldt < n) return fail(GM_E_SHAPE, "gm_map_deproject: template stride < n");
    if (n == 0 || ntemp == 0) return GM_OK;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_deproject, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, maps, n, templates,
                       (int)ntemp, ldt, alphas);
    return check_launch("gm_map_deproject");
}

extern "C" int gm_alm_filter(const double* alm, int64_t nin, int64_t alm_stride, const double* cl, int64_t nout,
                             int64_t lmax, double* out, int64_t out_stride, void* stream) {
    if (!alm || !cl || !out) return fail(GM_E_ARG, "gm_alm_filter: null pointer");
    if (nin < 1 || nout < 1 || nin > 16 || nout > 16 || lmax < 0) return fail(GM_E_ARG, "gm_alm_filter: bad sizes");
    const int64_t L = lmax + 1, nelem = L * (L + 1) / 2;
    if ((nin > 1 && alm_stride < nelem) || (nout > 1 && out_stride < nelem))
        return fail(GM_E_SHAPE, "gm_alm_filter: stride < nelem");
    if (out == alm) return fail(GM_E_ARG, "gm_alm_filter: out must not alias alm");
    const unsigned bx = (unsigned)((L + 255) / 256 < 8 ? (L + 255) / 256 : 8);
    hipLaunchKernelGGL(k_alm_filter, dim3(bx, (unsigned)L), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<const double2*>(alm), (int)nin, alm_stride, cl, (int)nout, (int)L,
                       reinterpret_cast<double2*>(out), out_stride);
    return check_launch("gm_alm_filter");
}

// ------------------------------------------------------------------------ purification
namespace gm {
namespace {

// out = conj(w) p per pixel for two-component (re, im) maps: (w0 p0 + w1 p1, w0 p1 - w1 p0)
__global__ void k_spin_mul(const double* __restrict__ w, const double* __restrict__ p, int64_t npix,
                           double* __restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += stride) {
        const double w0 = w[i], w1 = w[npix + i], p0 = p[i], p1 = p[npix + i];
        out[i] = w0 * p0 + w1 * p1;
        out[npix + i] = w0 * p1 - w1 * p0;
    }
}

}  // namespace
}  // namespace gm

extern "C" int gm_spin_mul(const double* w, const double* p, int64_t npix, double* out, void* stream) {
    if (!w || !p || !out) return fail(GM_E_ARG, "gm_spin_mul: null pointer");
    if (npix < 1) return fail(GM_E_ARG, "gm_spin_mul: npix must be positive");
    int64_t blocks = (npix + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_spin_mul, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, w, p, npix, out);
    return check_launch("gm_spin_mul");
}
