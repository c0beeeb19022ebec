// Shared device helpers for the gfx950 kernels: HEALPix ring geometry from closed forms,
// complex arithmetic on double2, extended-exponent seeds and error reporting.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gmaster_amd.h"

namespace gm {

// --------------------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);
int check_launch(const char* what);

// Layout of the padded per-order tables shared by the Legendre kernels: order m owns
// the entries l = m .. L - 1 + GM_TABLE_PAD, at base(m) + (l - m).
__host__ __device__ inline int64_t table_base(int64_t m, int64_t L) {
    return m * (L + GM_TABLE_PAD) - m * (m - 1) / 2;
}
__host__ __device__ inline int64_t table_size(int64_t L) { return table_base(L, L); }
// healpy m-major packing
__host__ __device__ inline int64_t alm_index(int64_t l, int64_t m, int64_t lmax) {
    return m * (2 * lmax + 1 - m) / 2 + l;
}

// --------------------------------------------------------------------------- geometry
struct Ring {
    int north;      // 1 .. 2 nside (the equator is 2 nside)
    int nphi;
    int64_t offset; // first pixel
    double z;       // cos(theta)
    double sin;     // sin(theta)
    double shalf;   // sin(theta/2)
    double chalf;   // cos(theta/2)
    int phi0_num;   // phi0 = pi * num / den
    int phi0_den;
};

// Ring r (0-based, 0 .. 4 nside - 2) of a HEALPix map in RING order.
__host__ __device__ inline Ring ring_info(int nside, int r) {
    Ring g;
    const int i = r + 1;
    const bool south = i > 2 * nside;
    const int north = south ? 4 * nside - i : i;
    const bool cap = north < nside;
    const double n = (double)nside;
    const double q = (double)north;
    double one_minus, z_north;
    if (cap) {
        one_minus = q * q / (3.0 * n * n);
        z_north = 1.0 - one_minus;
    } else {
        one_minus = (2.0 * q - n) / (3.0 * n);
        z_north = 2.0 * (2.0 * n - q) / (3.0 * n);
    }
    const double one_plus = 1.0 + z_north;
    g.north = north;
    g.nphi = cap ? 4 * north : 4 * nside;
    const int64_t N = nside;
    if (i < nside) {
        g.offset = 2LL * i * (i - 1);
    } else if (i <= 3 * nside) {
        g.offset = 2 * N * (N - 1) + (int64_t)(i - nside) * 4 * N;
    } else {
        const int64_t ni = 4 * N - i;
        g.offset = 12 * N * N - 2 * ni * (ni + 1);
    }
    g.sin = sqrt(one_minus * one_plus);
    const double sh = sqrt(0.5 * one_minus), ch = sqrt(0.5 * one_plus);
    g.z = south ? -z_north : z_north;
    g.shalf = south ? ch : sh;
    g.chalf = south ? sh : ch;
    g.phi0_den = cap ? 4 * north : 4 * nside;
    g.phi0_num = cap ? 1 : (((north - nside) % 2) == 0 ? 1 : 0);
    return g;
}

// e^{i sign pi * num * m / den} with exact integer range reduction.
__device__ inline double2 phase_pi(int64_t m, int num, int den, double sign) {
    if (num == 0) return make_double2(1.0, 0.0);
    const int64_t period = 2LL * den;
    int64_t t = (m * num) % period;
    if (t < 0) t += period;
    double s, c;
    sincospi(sign * (double)t / (double)den, &s, &c);
    return make_double2(c, s);
}

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

// x^n as mantissa * 2^e with mantissa in [0.5, 1): binary powering with renormalisation,
// so the relative error is O(log2 n) ulps at any n (no overflow or underflow).
__device__ inline void pow_ext(double x, int n, double& mant, int& e) {
    int ex;
    double b = frexp(x, &ex);  // b in [0.5, 1)
    double r = 1.0;
    int re = 0;
    int64_t be = ex;
    while (n) {
        if (n & 1) {
            r *= b;
            re += (int)be;
            int t;
            r = frexp(r, &t);
            re += t;
        }
        n >>= 1;
        if (n) {
            b *= b;
            be *= 2;
            int t;
            b = frexp(b, &t);
            be += t;
        }
    }
    mant = r;
    e = re;
}

}  // namespace gm
