// Correlated Gaussian alms on gfx950 for synfast (gmaster/utils.py:806-927): per packed
// element (l, m), 2 ncomp standard normals from a counter-based generator, multiplied by the
// per-l square root of the component covariance.
//
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw 2011), key = the 64-bit seed,
// counter = (element index low, element index high, draw index, 0).  Each Philox output
// (four 32-bit words) gives two uniforms of 53 bits, u1 in (0, 1] and u2 in [0, 1), and two
// normals by Box-Muller, sqrt(-2 ln u1) (cos 2 pi u2, sin 2 pi u2).  Draw c of an element
// gives the real (first normal) and imaginary (second normal) parts of component c.  The
// stream depends only on (seed, element), not on the launch shape.
#include <math.h>

#include "common.h"

namespace gm {
namespace {

__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += W0;
            k1 += W1;
        }
        const uint64_t p0 = (uint64_t)M0 * c[0];
        const uint64_t p1 = (uint64_t)M1 * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0;
        const uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

__device__ inline void gauss_pair(uint64_t seed, uint64_t elem, uint32_t draw, double& z0, double& z1) {
    uint32_t c[4] = {(uint32_t)elem, (uint32_t)(elem >> 32), draw, 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t a = ((uint64_t)c[0] << 21) ^ (uint64_t)(c[1] >> 11);
    const uint64_t b = ((uint64_t)c[2] << 21) ^ (uint64_t)(c[3] >> 11);
    const double inv53 = 1.0 / 9007199254740992.0;  // 2^-53
    const double u1 = ((double)(a & ((1ull << 53) - 1)) + 1.0) * inv53;
    const double u2 = (double)(b & ((1ull << 53) - 1)) * inv53;
    const double rad = sqrt(-2.0 * log(u1));
    double s, co;
    sincospi(2.0 * u2, &s, &co);
    z0 = rad * co;
    z1 = rad * s;
}

constexpr int kMaxComp = 8;

// out[c][e] = sum_i root[l][c][i] (re_i + i im_i) * (1 or 1/sqrt2), im = 0 at m = 0
__global__ void k_gaussian_alms(const double* __restrict__ root, int lmax, int ncomp, uint64_t seed,
                                double2* __restrict__ out, int64_t nelem) {
    const int m = blockIdx.y;
    const int L = lmax + 1;
    for (int l = m + blockIdx.x * blockDim.x + threadIdx.x; l < L; l += gridDim.x * blockDim.x) {
        const int64_t e = alm_index(l, m, lmax);
        double re[kMaxComp], im[kMaxComp];
        for (int i = 0; i < ncomp; ++i) gauss_pair(seed, (uint64_t)e, (uint32_t)i, re[i], im[i]);
        const double sc = (m == 0) ? 1.0 : 0.70710678118654752440;
        const double* R = root + (int64_t)l * ncomp * ncomp;
        for (int c = 0; c < ncomp; ++c) {
            double a = 0.0, b = 0.0;
            for (int i = 0; i < ncomp; ++i) {
                a += R[c * ncomp + i] * re[i];
                b += R[c * ncomp + i] * im[i];
            }
            out[(int64_t)c * nelem + e] = make_double2(a * sc, m == 0 ? 0.0 : b * sc);
        }
    }
}

// raw draws for tests: out[2 (e * ndraw + d) + {0, 1}] = the normal pair of (element e, draw d)
__global__ void k_gauss_draws(uint64_t seed, int64_t nelem, int ndraw, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nelem * ndraw) return;
    const int64_t e = i / ndraw;
    const int d = (int)(i % ndraw);
    double z0, z1;
    gauss_pair(seed, (uint64_t)e, (uint32_t)d, z0, z1);
    out[2 * i] = z0;
    out[2 * i + 1] = z1;
}

}  // namespace
}  // namespace gm

using namespace gm;

extern "C" int gm_gaussian_alms(const double* root, int64_t lmax, int64_t ncomp, uint64_t seed, double* out,
                                void* stream) {
    if (!root || !out) return fail(GM_E_ARG, "gm_gaussian_alms: null pointer");
    if (lmax < 0 || lmax > 65535 || ncomp < 1 || ncomp > kMaxComp)
        return fail(GM_E_ARG, "gm_gaussian_alms: bad lmax / ncomp (1..%d)", kMaxComp);
    const int64_t L = lmax + 1, nelem = L * (L + 1) / 2;
    const unsigned bx = (unsigned)((L + 255) / 256 < 8 ? (L + 255) / 256 : 8);
    hipLaunchKernelGGL(k_gaussian_alms, dim3(bx, (unsigned)L), dim3(256), 0, (hipStream_t)stream, root, (int)lmax,
                       (int)ncomp, seed, reinterpret_cast<double2*>(out), nelem);
    return check_launch("gm_gaussian_alms");
}

extern "C" int gm_gauss_draws(uint64_t seed, int64_t nelem, int64_t ndraw, double* out, void* stream) {
    if (!out) return fail(GM_E_ARG, "gm_gauss_draws: null pointer");
    if (nelem < 1 || ndraw < 1 || ndraw > 1024 || nelem * ndraw > (1LL << 31))
        return fail(GM_E_ARG, "gm_gauss_draws: bad sizes");
    const int64_t n = nelem * ndraw;
    hipLaunchKernelGGL(k_gauss_draws, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, seed,
                       nelem, (int)ndraw, out);
    return check_launch("gm_gauss_draws");
}
