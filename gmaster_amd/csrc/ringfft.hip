// Ring (azimuthal) stage of the HEALPix transforms on gfx950, FP64.
//
// Each HEALPix ring has nphi = 4 i (polar caps) or 4 nside (belt) pixels, so every
// length is even: a real ring of length n is transformed as one complex FFT of
// N = n / 2 points held entirely in LDS (N <= 8192 -> 128 KiB at Nside 4096), mixed
// radix 4/2/3/5 Stockham passes plus a direct "DFT pass" (Horner over the p inputs, one
// output per thread) for larger prime factors.  One workgroup per ring; rings are
// launched in size buckets so the dynamic LDS matches the largest ring of the bucket.
//
// The forward output is the *phased* spectrum  G[r][m] = w e^{-i m phi0} X[m mod n]
// (w = 4 pi / npix), exactly the right-hand side of the Legendre analysis, so no
// separate weight / phase pass runs.  The Jacobi residual is formed in this space by
// the fold kernel at the end of the file, with no FFT inside the refinement loop.
#include <stdarg.h>

#include <mutex>
#include <utility>
#include <vector>

#include "common.h"

namespace gm {

namespace {

__device__ inline double2 twiddle(int64_t num, int64_t den, double sign) {
    // e^{sign 2 pi i num / den}
    double s, c;
    sincospi(sign * 2.0 * (double)num / (double)den, &s, &c);
    return make_double2(c, s);
}

template <int P>
__device__ inline void dft_small(double2* v, double sign) {
    if constexpr (P == 2) {
        const double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (P == 4) {
        const double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        const double2 t2 = cadd(v[1], v[3]);
        const double2 d = csub(v[1], v[3]);
        const double2 t3 = make_double2(-sign * d.y, sign * d.x);  // d * (sign i)
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    } else if constexpr (P == 3) {
        const double h = 0.86602540378443864676;  // sqrt(3)/2
        const double2 t = cadd(v[1], v[2]);
        const double2 d = csub(v[1], v[2]);
        const double2 u = make_double2(-sign * h * d.y, sign * h * d.x);
        const double2 base = make_double2(v[0].x - 0.5 * t.x, v[0].y - 0.5 * t.y);
        v[0] = cadd(v[0], t);
        v[1] = cadd(base, u);
        v[2] = csub(base, u);
    } else {  // P == 5: direct 5-point DFT with exact constants
        const double c1 = 0.30901699437494742410, s1 = 0.95105651629515357212;
        const double c2 = -0.80901699437494742410, s2 = 0.58778525229247312917;
        const double2 w[5] = {make_double2(1.0, 0.0), make_double2(c1, sign * s1),
                              make_double2(c2, sign * s2), make_double2(c2, -sign * s2),
                              make_double2(c1, -sign * s1)};
        double2 out[5];
        for (int k = 0; k < 5; ++k) {
            double2 acc = v[0];
            for (int q = 1; q < 5; ++q) acc = cadd(acc, cmul(v[q], w[(q * k) % 5]));
            out[k] = acc;
        }
        for (int k = 0; k < 5; ++k) v[k] = out[k];
    }
}

// One Stockham pass of radix P (P in {2,3,4,5}) over buf[0..N), in place.  With
// blockDim.x >= N / 8 each thread owns at most ceil(8 / P) butterflies.
template <int P>
__device__ void radix_pass(double2* buf, int N, int Ns, double sign) {
    const int J = N / P;
    const int T = blockDim.x;
    constexpr int MB = (8 + P - 1) / P;
    double2 v[MB][P];
#pragma unroll
    for (int t = 0; t < MB; ++t) {
        const int j = threadIdx.x + t * T;
        if (j < J) {
            const int k = j % Ns;
#pragma unroll
            for (int q = 0; q < P; ++q) v[t][q] = buf[j + q * J];
            if (Ns > 1) {
                const double2 w1 = twiddle(k, (int64_t)Ns * P, sign);
                double2 w = w1;
#pragma unroll
                for (int q = 1; q < P; ++q) {
                    v[t][q] = cmul(v[t][q], w);
                    w = cmul(w, w1);
                }
            }
            dft_small<P>(v[t], sign);
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MB; ++t) {
        const int j = threadIdx.x + t * T;
        if (j < J) {
            const int k = j % Ns;
            const int base = (j / Ns) * Ns * P + k;
#pragma unroll
            for (int q = 0; q < P; ++q) buf[base + q * Ns] = v[t][q];
        }
    }
    __syncthreads();
}

// Direct DFT pass of any radix p (a prime above 5): output f = j + q J (J = N / p) is
// sum_u x[j + u J] w_f^u, w_f = e^{sign 2 pi i (j mod Ns + q Ns) / (Ns p)}, by Horner's rule.
// Each thread takes one residue j and up to 8 values of q (q = qb + t Tj with Tj = T / J
// threads per residue), so its 8 Horner chains read the same input: one LDS read per step
// feeds 8 independent complex multiply-adds.  When 8 chains per thread do not cover p
// (T < J p / 8 after rounding) each thread computes outputs f = tid + t T instead, one
// read per chain and step.
// a w + x with two fused multiply-adds per component
__device__ inline double2 horner_step(double2 a, double2 w, double2 x) {
    return make_double2(fma(a.x, w.x, fma(-a.y, w.y, x.x)), fma(a.x, w.y, fma(a.y, w.x, x.y)));
}

__device__ __forceinline__ void dft_pass(double2* buf, int N, int p, int Ns, double sign) {
    const int J = N / p;
    const int T = blockDim.x;
    const int Tj = T / J;
    double2 w[8], acc[8];
    if (Tj > 0 && 8 * Tj >= p) {
        const int j = threadIdx.x % J, qb = threadIdx.x / J;
        const bool active = qb < Tj;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int q = min(qb + t * Tj, p - 1);
            w[t] = twiddle(j % Ns + (int64_t)q * Ns, (int64_t)Ns * p, sign);
        }
        if (active) {
            const double2 x0 = buf[j + (p - 1) * J];
#pragma unroll
            for (int t = 0; t < 8; ++t) acc[t] = x0;
            for (int u = p - 2; u >= 0; --u) {
                const double2 x = buf[j + u * J];
#pragma unroll
                for (int t = 0; t < 8; ++t) acc[t] = horner_step(acc[t], w[t], x);
            }
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int q = qb + t * Tj;
                if (q < p) buf[(j / Ns) * Ns * p + j % Ns + q * Ns] = acc[t];
            }
        }
        __syncthreads();
        return;
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const int f = threadIdx.x + t * T;
        if (f < N) {
            const int j = f % J, q = f / J;
            const double2 wf = twiddle(j % Ns + (int64_t)q * Ns, (int64_t)Ns * p, sign);
            double2 a = buf[j + (p - 1) * J];
            for (int u = p - 2; u >= 0; --u) a = horner_step(a, wf, buf[j + u * J]);
            acc[t] = a;
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const int f = threadIdx.x + t * T;
        if (f < N) {
            const int j = f % J, q = f / J;
            buf[(j / Ns) * Ns * p + j % Ns + q * Ns] = acc[t];
        }
    }
    __syncthreads();
}

// ------------------------------------------------------------------ Bluestein pass
// A prime radix p >= kBluesteinMin as the last pass (rem == p, Ns = J = N / p) runs as a
// chirp-z convolution (Bluestein 1970) instead of Horner's O(N p): with
// c_k = e^{sign i pi k^2 / p}, the DFT X_q = sum_u y_u e^{sign 2 pi i q u / p} equals
// c_q sum_u (y_u c_u) conj(c_{q-u}), a circular convolution of length M >= 2p - 1 (a power
// of two) done by radix-4/2 FFTs in LDS.  y_u = x[j + u J] e^{sign 2 pi i j u / N} for
// residue j, and its outputs return to positions j + q J, as in the Horner pass.  The
// ring's N points are parked in registers (T >= N / 8: at most 8 per thread) so that LDS
// holds the M-point buffer; the filter spectrum FFT(conj c) is computed once per ring and
// kept in registers (M / T <= 8 per thread).
#ifndef GM_BLUESTEIN_MIN
#define GM_BLUESTEIN_MIN 1024
#endif
constexpr int kBluesteinMin = GM_BLUESTEIN_MIN;

// e^{sign i pi k^2 / p}, the phase reduced mod 2 p in integers
__device__ inline double2 chirp(int k, int p, double sign) {
    const int r = (k * k) % (2 * p);  // k < 8192
    double sn, cs;
    sincospi(sign * (double)r / (double)p, &sn, &cs);
    return make_double2(cs, sn);
}

// e^{sign 2 pi i k / M} for 0 <= k < M / 2 from tw[k] = e^{-2 pi i k / M}, k <= M / 8
// (the other octants by e^{-2 pi i (M/4 - k) / M} = i conj(...) and e^{-i pi / 2} = -i).
__device__ inline double2 tw_get(const double2* tw, int k, int M, double sign) {
    const int q = M >> 2, e = M >> 3;
    const bool upper = k >= q;  // times -i
    if (upper) k -= q;
    double2 w = k <= e ? tw[k] : make_double2(-tw[q - k].y, -tw[q - k].x);  // -i conj(tw[q - k])
    if (upper) w = make_double2(w.y, -w.x);
    if (sign > 0.0) w.y = -w.y;
    return w;
}

// One Stockham pass of radix P in {2, 4} over a power-of-two length M, twiddles from tw.
template <int P>
__device__ void radix_pass_tab(double2* buf, int M, int Ns, double sign, const double2* tw) {
    const int J = M / P;
    const int T = blockDim.x;
    constexpr int MB = 8 / P;
    const int step = M / (Ns * P);
    double2 v[MB][P];
#pragma unroll
    for (int t = 0; t < MB; ++t) {
        const int j = threadIdx.x + t * T;
        if (j < J) {
            const int k = j % Ns;
#pragma unroll
            for (int q = 0; q < P; ++q) v[t][q] = buf[j + q * J];
            if (Ns > 1) {
                const double2 w1 = tw_get(tw, k * step, M, sign);
                double2 w = w1;
#pragma unroll
                for (int q = 1; q < P; ++q) {
                    v[t][q] = cmul(v[t][q], w);
                    w = cmul(w, w1);
                }
            }
            dft_small<P>(v[t], sign);
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MB; ++t) {
        const int j = threadIdx.x + t * T;
        if (j < J) {
            const int k = j % Ns;
            const int base = (j / Ns) * Ns * P + k;
#pragma unroll
            for (int q = 0; q < P; ++q) buf[base + q * Ns] = v[t][q];
        }
    }
    __syncthreads();
}

// Power-of-two FFT of length M in LDS (radix 4 passes, one radix 2 when log2 M is odd).
__device__ __forceinline__ void fft_pow2(double2* buf, int M, double sign, const double2* tw) {
    int Ns = 1, rem = M;
    while (rem > 1) {
        if (rem % 4 == 0) {
            radix_pass_tab<4>(buf, M, Ns, sign, tw);
            Ns *= 4;
            rem /= 4;
        } else {
            radix_pass_tab<2>(buf, M, Ns, sign, tw);
            Ns *= 2;
            rem /= 2;
        }
    }
}

__device__ inline int bluestein_len(int p) {
    int M = 1;
    while (M < 2 * p - 1) M <<= 1;
    return M;
}

// Whether the last pass of radix p over N points runs as a convolution: p large enough,
// at least two residues (so M <= the next power of two of N, which the launch provides
// in LDS) and threads for the M-point FFT and the parked data.
__device__ inline bool use_bluestein(int N, int p) {
    if (p < kBluesteinMin || N / p < 2) return false;
    const int M = bluestein_len(p);
    const int T = blockDim.x;
    return M <= 8 * T && N <= 8 * T;
}

// tw: M / 8 + 1 points of LDS beside the M-point buffer for the twiddle table.
__device__ __forceinline__ void bluestein_pass(double2* buf, int N, int p, double sign, double2* tw) {
    const int T = blockDim.x, tid = threadIdx.x;
    const int J = N / p;
    const int M = bluestein_len(p);
    for (int k = tid; k <= M / 8; k += T) tw[k] = twiddle(k, M, -1.0);
    double2 park[8], bf[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int i = tid + k * T;
        park[k] = i < N ? buf[i] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    // filter conj(c_k) on k in (-p, p), wrapped into M points, and its spectrum
    for (int k = tid; k < M; k += T) {
        const int kk = k < p ? k : (k > M - p ? M - k : -1);
        buf[k] = kk >= 0 ? chirp(kk, p, -sign) : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_pow2(buf, M, -1.0, tw);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int i = tid + k * T;
        bf[k] = i < M ? buf[i] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const double inv_m = 1.0 / (double)M;
    for (int j = 0; j < J; ++j) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int i = tid + k * T;
            if (i < N && i % J == j) {
                const int u = i / J;
                buf[u] = cmul(cmul(park[k], twiddle((int64_t)j * u, N, sign)), chirp(u, p, sign));
            }
        }
        for (int k = p + tid; k < M; k += T) buf[k] = make_double2(0.0, 0.0);
        __syncthreads();
        fft_pow2(buf, M, -1.0, tw);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int i = tid + k * T;
            if (i < M) buf[i] = cmul(buf[i], bf[k]);
        }
        __syncthreads();
        fft_pow2(buf, M, 1.0, tw);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int i = tid + k * T;
            if (i < N && i % J == j) {
                const int q = i / J;
                park[k] = cscale(cmul(buf[q], chirp(q, p, sign)), inv_m);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int i = tid + k * T;
        if (i < N) buf[i] = park[k];
    }
    __syncthreads();
}

__device__ int smallest_factor(int n) {
    if (n % 2 == 0) return 2;
    for (int f = 3; f * f <= n; f += 2)
        if (n % f == 0) return f;
    return n;
}

// In-place complex FFT of length N in LDS, sign -1 forward, +1 inverse (unnormalised).
// Requires blockDim.x >= N / 8 (at most 8 outputs per thread per pass).  Inlined into each
// kernel, so the library makes no device-side calls: compiled as a called function with
// the FMA Horner step above, it returned spectra off by up to ~1e-7 relative on gfx950
// (rings that never reach the direct pass included), while the inlined form matches
// numpy to ~1e-16; the cause was not isolated.
// tw: LDS for a twiddle table of N / 8 + 1 points past the data buffer (or null); with it,
// power-of-two lengths take their twiddles from the table.  BS: prime passes
// >= kBluesteinMin may run as a Bluestein convolution (its table goes there too).
template <bool BS = false>
__device__ __forceinline__ void fft_lds(double2* buf, int N, double sign, double2* tw = nullptr) {
    if (tw != nullptr && N >= 8 && (N & (N - 1)) == 0) {  // power of two: twiddles from an LDS table
        for (int k = threadIdx.x; k <= N / 8; k += blockDim.x) tw[k] = twiddle(k, N, -1.0);
        __syncthreads();
        fft_pow2(buf, N, sign, tw);
        return;
    }
    int Ns = 1, rem = N;
    while (rem > 1) {
        int p;
        if (rem % 4 == 0) p = 4;
        else if (rem % 2 == 0) p = 2;
        else if (rem % 3 == 0) p = 3;
        else if (rem % 5 == 0) p = 5;
        else p = smallest_factor(rem);
        switch (p) {
            case 4: radix_pass<4>(buf, N, Ns, sign); break;
            case 2: radix_pass<2>(buf, N, Ns, sign); break;
            case 3: radix_pass<3>(buf, N, Ns, sign); break;
            case 5: radix_pass<5>(buf, N, Ns, sign); break;
            default:
                if (BS && rem == p && use_bluestein(N, p))
                    bluestein_pass(buf, N, p, sign, tw);
                else
                    dft_pass(buf, N, p, Ns, sign);
                break;
        }
        Ns *= p;
        rem /= p;
    }
}

// Ring index of block b within a size bucket: caps take north indices [qlo, qhi] in both
// hemispheres; the belt bucket (qlo == 0) takes rings nside-1 .. 3 nside-1.
__device__ inline int bucket_ring(int nside, int qlo, int qhi, int b) {
    if (qlo == 0) return nside - 1 + b;
    const int count = qhi - qlo + 1;
    if (b < count) return qlo + b - 1;
    const int q = qlo + (b - count);
    return 4 * nside - 1 - q;
}

// z[k] = x[2k] + i x[2k+1] -> X[k] (k = 0..N) in place; X[N] returned through *xn.
__device__ void r2c_post(double2* buf, int N, double2* xn) {
    const int n = 2 * N;
    for (int k = threadIdx.x; k <= N / 2; k += blockDim.x) {
        if (k == 0) {
            const double2 z = buf[0];
            buf[0] = make_double2(z.x + z.y, 0.0);
            *xn = make_double2(z.x - z.y, 0.0);
            continue;
        }
        const int kk = N - k;
        const double2 a = buf[k], b = buf[kk];
        // X[k] = (a + conj b)/2 - (i/2) e^{-2 pi i k / n} (a - conj b)
        const double2 e = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
        const double2 o = make_double2(0.5 * (a.x - b.x), 0.5 * (a.y + b.y));
        const double2 w = twiddle(k, n, -1.0);
        const double2 wo = cmul(w, o);
        const double2 xk = make_double2(e.x + wo.y, e.y - wo.x);
        if (kk != k) {
            const double2 e2 = make_double2(e.x, -e.y);   // (b + conj a)/2
            const double2 o2 = make_double2(-o.x, o.y);   // (b - conj a)/2
            const double2 w2 = twiddle(kk, n, -1.0);
            const double2 wo2 = cmul(w2, o2);
            buf[kk] = make_double2(e2.x + wo2.y, e2.y - wo2.x);
        }
        buf[k] = xk;
    }
    __syncthreads();
}

// X[k] (k = 0..N, X[N] in xn) -> Z[k] for the inverse packed FFT, in place.
__device__ void c2r_pre(double2* buf, int N, double2 xn) {
    const int n = 2 * N;
    for (int k = threadIdx.x; k <= N / 2; k += blockDim.x) {
        const int kk = N - k;
        const double2 a = buf[k];
        const double2 b = (k == 0) ? xn : buf[kk];
        // Z[k] = (X_k + conj X_{N-k}) + i e^{2 pi i k/n} (X_k - conj X_{N-k})
        const double2 s = make_double2(a.x + b.x, a.y - b.y);
        const double2 d = make_double2(a.x - b.x, a.y + b.y);
        const double2 wd = cmul(twiddle(k, n, 1.0), d);
        const double2 zk = make_double2(s.x - wd.y, s.y + wd.x);
        if (k != 0 && kk != k) {
            const double2 s2 = make_double2(s.x, -s.y);
            const double2 d2 = make_double2(-d.x, d.y);
            const double2 wd2 = cmul(twiddle(kk, n, 1.0), d2);
            buf[kk] = make_double2(s2.x - wd2.y, s2.y + wd2.x);
        }
        buf[k] = zk;
    }
    __syncthreads();
}

__device__ inline double2 half_spectrum(const double2* buf, double2 xn, int N, int kk) {
    // X[kk] for kk in [0, 2N) from the half spectrum X[0..N]
    if (kk < N) return buf[kk];
    if (kk == N) return xn;
    return cconj(buf[2 * N - kk]);
}

// ----------------------------------------------------------------------- forward kernels
template <bool BS>
__global__ void k_ring_fwd_s0(const double* __restrict__ maps, int64_t map_stride, int nside,
                              int L, int qlo, int qhi, double2* __restrict__ spec, int64_t ld,
                              int64_t spec_map_stride, int dcap) {
    extern __shared__ double2 buf[];
    __shared__ double2 xn;
    const int r = bucket_ring(nside, qlo, qhi, blockIdx.x);
    const Ring g = ring_info(nside, r);
    const int n = g.nphi, N = n / 2;
    const double* x = maps + blockIdx.y * map_stride + g.offset;
    for (int k = threadIdx.x; k < N; k += blockDim.x) buf[k] = make_double2(x[2 * k], x[2 * k + 1]);
    __syncthreads();
    fft_lds<BS>(buf, N, -1.0, buf + dcap);
    r2c_post(buf, N, &xn);
    const double w = 4.0 * M_PI / (12.0 * (double)nside * (double)nside);
    double2* out = spec + blockIdx.y * spec_map_stride + (int64_t)r * ld;
    // w e^{-i m phi0}, advanced by the fixed rotation of a thread's stride
    double2 ph = cscale(phase_pi(threadIdx.x, g.phi0_num, g.phi0_den, -1.0), w);
    const double2 rot = phase_pi(blockDim.x, g.phi0_num, g.phi0_den, -1.0);
    for (int m = threadIdx.x; m < L; m += blockDim.x) {
        const double2 X = half_spectrum(buf, xn, N, m % n);
        out[m] = cmul(ph, X);
        ph = cmul(ph, rot);
    }
}

template <bool BS>
__global__ void k_ring_fwd_s(const double* __restrict__ q, const double* __restrict__ u, int nside,
                             int L, int qlo, int qhi, double2* __restrict__ spec, int64_t ld, int dcap) {
    extern __shared__ double2 buf[];
    __shared__ double2 xn;
    const int r = bucket_ring(nside, qlo, qhi, blockIdx.x);
    const Ring g = ring_info(nside, r);
    const int n = g.nphi, N = n / 2;
    const double w = 4.0 * M_PI / (12.0 * (double)nside * (double)nside);
    double2* out = spec + (int64_t)r * ld;
    const int W = 2 * L - 1;
    for (int comp = 0; comp < 2; ++comp) {
        const double* x = (comp == 0 ? q : u) + g.offset;
        for (int k = threadIdx.x; k < N; k += blockDim.x) buf[k] = make_double2(x[2 * k], x[2 * k + 1]);
        __syncthreads();
        fft_lds<BS>(buf, N, -1.0, buf + dcap);
        r2c_post(buf, N, &xn);
        double2 ph = cscale(phase_pi((int)threadIdx.x - (L - 1), g.phi0_num, g.phi0_den, -1.0), w);
        const double2 rot = phase_pi(blockDim.x, g.phi0_num, g.phi0_den, -1.0);
        for (int j = threadIdx.x; j < W; j += blockDim.x) {
            const int m = j - (L - 1);
            int kk = m % n;
            if (kk < 0) kk += n;
            const double2 X = half_spectrum(buf, xn, N, kk);
            const double2 v = cmul(ph, X);
            ph = cmul(ph, rot);
            if (comp == 0) {
                out[j] = v;
            } else {
                const double2 o = out[j];  // written by this same thread in pass 0
                out[j] = make_double2(o.x - v.y, o.y + v.x);  // + i v
            }
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------------- inverse kernels
// Y(k) = sum_{m = k mod n, m in [mlo, mhi]} F[m] e^{i m phi0}; F indexed at m - mlo_index.
// phi0 = pi num / n on every ring (num = 0 or 1), so the aliases m0 + j n of one k share the
// phase of m0 up to the sign (-1)^(j num): one phase per k.
__device__ inline double2 fold_sum(const double2* F, int64_t col0, int mlo, int mhi, int k, int n,
                                   const Ring& g) {
    // first m >= mlo with m = k (mod n)
    int64_t t = ((int64_t)k - mlo) % n;
    if (t < 0) t += n;
    const int64_t m0 = mlo + t;
    double2 acc = make_double2(0.0, 0.0);
    bool neg = false;
    for (int64_t m = m0; m <= mhi; m += n) {
        const double2 v = F[m + col0];
        acc = neg ? csub(acc, v) : cadd(acc, v);
        if (g.phi0_num) neg = !neg;
    }
    return cmul(acc, phase_pi(m0, g.phi0_num, g.phi0_den, 1.0));
}

// fold_sum with the phase of residue k supplied: phk = e^{i pi k num / n} for 0 <= k < n, so
// that a caller stepping k by a fixed stride advances it by one rotation instead of a
// sincospi per k; the first alias m0 = k + j0 n carries phk (-1)^(j0 num).
__device__ inline double2 fold_sum_ph(const double2* F, int64_t col0, int mlo, int mhi, int k, int n, int num,
                                      double2 phk) {
    int64_t t = ((int64_t)k - mlo) % n;
    if (t < 0) t += n;
    const int64_t m0 = mlo + t;
    double2 acc = make_double2(0.0, 0.0);
    bool neg = false;
    for (int64_t m = m0; m <= mhi; m += n) {
        const double2 v = F[m + col0];
        acc = neg ? csub(acc, v) : cadd(acc, v);
        if (num) neg = !neg;
    }
    const int64_t j0 = (m0 - k) / n;  // exact: m0 = k (mod n)
    if (num && (j0 & 1)) phk = make_double2(-phk.x, -phk.y);
    return cmul(acc, phk);
}

// e^{i pi (n - k) num / n} = (-1)^num conj(e^{i pi k num / n}) for 1 <= k < n; 1 for k = 0
__device__ inline double2 mirror_phase(double2 phk, int k, int num) {
    if (k == 0) return make_double2(1.0, 0.0);
    const double2 c = make_double2(phk.x, -phk.y);
    return num ? make_double2(-c.x, -c.y) : c;
}

template <bool BS>
__global__ void k_ring_inv_s0(const double2* __restrict__ spec, int64_t ld, int64_t spec_map_stride,
                              int nside, int L, int qlo, int qhi, double* __restrict__ maps,
                              int64_t map_stride, int dcap) {
    extern __shared__ double2 buf[];
    __shared__ double2 xn;
    const int r = bucket_ring(nside, qlo, qhi, blockIdx.x);
    const Ring g = ring_info(nside, r);
    const int n = g.nphi, N = n / 2;
    const double2* F = spec + blockIdx.y * spec_map_stride + (int64_t)r * ld;
    // e^{i pi k num / n}, advanced by the rotation of the thread stride
    double2 phk = phase_pi(threadIdx.x, g.phi0_num, g.phi0_den, 1.0);
    const double2 rot = phase_pi(blockDim.x, g.phi0_num, g.phi0_den, 1.0);
    for (int k = threadIdx.x; k <= N; k += blockDim.x) {
        // X[k] = sum_{m=k} F e^{i m phi0} + conj(sum_{m=-k, m>=1} F e^{i m phi0})
        const double2 a = fold_sum_ph(F, 0, 0, L - 1, k, n, g.phi0_num, phk);
        const int kneg = (n - k) % n;
        const double2 b = fold_sum_ph(F, 0, 1, L - 1, kneg, n, g.phi0_num, mirror_phase(phk, k, g.phi0_num));
        phk = cmul(phk, rot);
        const double2 X = make_double2(a.x + b.x, a.y - b.y);
        if (k < N) buf[k] = X;
        else xn = X;
    }
    __syncthreads();
    c2r_pre(buf, N, xn);
    fft_lds<BS>(buf, N, 1.0, buf + dcap);
    double* x = maps + blockIdx.y * map_stride + g.offset;
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
        const double2 z = buf[k];
        x[2 * k] = z.x;
        x[2 * k + 1] = z.y;
    }
}

template <bool BS>
__global__ void k_ring_inv_s(const double2* __restrict__ spec, int64_t ld, int nside, int L, int qlo,
                             int qhi, double* __restrict__ q, double* __restrict__ u, int dcap) {
    extern __shared__ double2 buf[];
    __shared__ double2 xn;
    const int r = bucket_ring(nside, qlo, qhi, blockIdx.x);
    const Ring g = ring_info(nside, r);
    const int n = g.nphi, N = n / 2;
    const double2* F = spec + (int64_t)r * ld;
    const int64_t col0 = L - 1;  // F[m] lives at column m + L - 1
    // XQ = (Y(k) + conj Y(-k)) / 2 and XU = -(i/2)(Y(k) - conj Y(-k)); the folds are cheap
    // next to the FFT, so each component refolds instead of holding the other in registers.
    const double2 rot = phase_pi(blockDim.x, g.phi0_num, g.phi0_den, 1.0);
    for (int comp = 0; comp < 2; ++comp) {
        double2 phk = phase_pi(threadIdx.x, g.phi0_num, g.phi0_den, 1.0);
        for (int k = threadIdx.x; k <= N; k += blockDim.x) {
            const double2 yp = fold_sum_ph(F, col0, -(L - 1), L - 1, k, n, g.phi0_num, phk);
            const double2 ym = fold_sum_ph(F, col0, -(L - 1), L - 1, (n - k) % n, n, g.phi0_num,
                                           mirror_phase(phk, k, g.phi0_num));
            phk = cmul(phk, rot);
            double2 X;
            if (comp == 0) {
                X = make_double2(0.5 * (yp.x + ym.x), 0.5 * (yp.y - ym.y));
            } else {
                const double2 d = make_double2(yp.x - ym.x, yp.y + ym.y);
                X = make_double2(0.5 * d.y, -0.5 * d.x);
            }
            if (k < N) buf[k] = X;
            else xn = X;
        }
        __syncthreads();
        c2r_pre(buf, N, xn);
        fft_lds<BS>(buf, N, 1.0, buf + dcap);
        double* x = (comp == 0 ? q : u) + g.offset;
        for (int k = threadIdx.x; k < N; k += blockDim.x) {
            const double2 z = buf[k];
            x[2 * k] = z.x;
            x[2 * k + 1] = z.y;
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------------- rings above LDS
// A ring whose packed length N = nphi / 2 exceeds kLdsN (Nside 8192: the belt has N =
// 16384, caps up to 16382) is transformed in two halves of H = N / 2 points, one parity of
// the output index at a time (one radix-2 decimation-in-frequency step outside LDS):
//   Z[2k]   = FFT_H(z[j] + z[j+H])[k],   Z[2k+1] = FFT_H((z[j] - z[j+H]) e^{-2 pi i j/N})[k].
// The packed-real post-processing pairs K with N - K, which have the same parity, and the
// phased output m reads X[m mod nphi], whose parity is that of m, so each parity pass
// finishes its own outputs.  The inverse runs the same split backwards: y_p = IFFT_H of
// the parity-p half, then z[j] = y_0 + w^j y_1, z[j+H] = y_0 - w^j y_1 (w = e^{2 pi i/N}),
// the second pass updating in place what the same thread wrote in the first.
#ifndef GM_RING_LDS_N
#define GM_RING_LDS_N 8192
#endif
constexpr int kLdsN = GM_RING_LDS_N;  // complex FP64 points per workgroup in LDS (128 KiB)

// buf[k] = Z[2k+p] (k < H) -> X[2k+p] in place; for p = 0 also X[N] through *xn.
__device__ void r2c_post_half(double2* buf, int H, int p, double2* xn) {
    const int N = 2 * H, n = 2 * N;
    for (int k = threadIdx.x; k < H; k += blockDim.x) {
        const int kp = p ? H - 1 - k : (H - k) % H;  // slot of Z[N - K]
        if (kp < k) continue;
        if (p == 0 && k == 0) {
            const double2 z = buf[0];
            buf[0] = make_double2(z.x + z.y, 0.0);
            *xn = make_double2(z.x - z.y, 0.0);
            continue;
        }
        const int K = 2 * k + p, KK = 2 * kp + p;
        const double2 a = buf[k], b = buf[kp];
        const double2 e = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
        const double2 o = make_double2(0.5 * (a.x - b.x), 0.5 * (a.y + b.y));
        const double2 wo = cmul(twiddle(K, n, -1.0), o);
        const double2 xk = make_double2(e.x + wo.y, e.y - wo.x);
        if (kp != k) {
            const double2 e2 = make_double2(e.x, -e.y);
            const double2 o2 = make_double2(-o.x, o.y);
            const double2 wo2 = cmul(twiddle(KK, n, -1.0), o2);
            buf[kp] = make_double2(e2.x + wo2.y, e2.y - wo2.x);
        }
        buf[k] = xk;
    }
    __syncthreads();
}

// buf[k] = X[2k+p] (k < H; X[N] in xn) -> Z[2k+p] in place.
__device__ void c2r_pre_half(double2* buf, int H, int p, double2 xn) {
    const int N = 2 * H, n = 2 * N;
    for (int k = threadIdx.x; k < H; k += blockDim.x) {
        const int kp = p ? H - 1 - k : (H - k) % H;
        if (kp < k) continue;
        const int K = 2 * k + p, KK = 2 * kp + p;
        const double2 a = buf[k];
        const double2 b = (p == 0 && k == 0) ? xn : buf[kp];
        const double2 s = make_double2(a.x + b.x, a.y - b.y);
        const double2 d = make_double2(a.x - b.x, a.y + b.y);
        const double2 wd = cmul(twiddle(K, n, 1.0), d);
        const double2 zk = make_double2(s.x - wd.y, s.y + wd.x);
        if (!(p == 0 && k == 0) && kp != k) {
            const double2 s2 = make_double2(s.x, -s.y);
            const double2 d2 = make_double2(-d.x, d.y);
            const double2 wd2 = cmul(twiddle(KK, n, 1.0), d2);
            buf[kp] = make_double2(s2.x - wd2.y, s2.y + wd2.x);
        }
        buf[k] = zk;
    }
    __syncthreads();
}

// Half of the packed complex sequence of a real ring x (z[j] = x[2j] + i x[2j+1]) for
// parity pass p, into buf[0..H).
__device__ inline void load_half(double2* buf, const double* x, int H, int p) {
    const int N = 2 * H;
    for (int k = threadIdx.x; k < H; k += blockDim.x) {
        const double2 a = make_double2(x[2 * k], x[2 * k + 1]);
        const double2 b = make_double2(x[2 * (k + H)], x[2 * (k + H) + 1]);
        buf[k] = p ? cmul(csub(a, b), twiddle(k, N, -1.0)) : cadd(a, b);
    }
    __syncthreads();
}

// X[kk] (parity p) from the parity-p half spectrum in buf (X[N] in xn)
__device__ inline double2 half_parity_x(const double2* buf, double2 xn, int N, int p, int kk) {
    if (kk < N) return buf[(kk - p) >> 1];
    if (kk == N) return xn;
    return cconj(buf[(2 * N - kk - p) >> 1]);
}

__global__ void k_ring_fwd_s0_split(const double* __restrict__ maps, int64_t map_stride, int nside, int L, int qlo,
                                    int qhi, double2* __restrict__ spec, int64_t ld, int64_t spec_map_stride) {
    extern __shared__ double2 buf[];
    __shared__ double2 xn;
    const int r = bucket_ring(nside, qlo, qhi, blockIdx.x);
    const Ring g = ring_info(nside, r);
    const int n = g.nphi, N = n / 2, H = N / 2;
    const double* x = maps + blockIdx.y * map_stride + g.offset;
    const double w = 4.0 * M_PI / (12.0 * (double)nside * (double)nside);
    double2* out = spec + blockIdx.y * spec_map_stride + (int64_t)r * ld;
    for (int p = 0; p < 2; ++p) {
        load_half(buf, x, H, p);
        fft_lds(buf, H, -1.0);
        r2c_post_half(buf, H, p, &xn);
        double2 ph = cscale(phase_pi(p + 2 * (int)threadIdx.x, g.phi0_num, g.phi0_den, -1.0), w);
        const double2 rot = phase_pi(2 * blockDim.x, g.phi0_num, g.phi0_den, -1.0);
        for (int m = p + 2 * threadIdx.x; m < L; m += 2 * blockDim.x) {
            const double2 X = half_parity_x(buf, xn, N, p, m % n);
            out[m] = cmul(ph, X);
            ph = cmul(ph, rot);
        }
        __syncthreads();
    }
}

__global__ void k_ring_fwd_s_split(const double* __restrict__ q, const double* __restrict__ u, int nside, int L,
                                   int qlo, int qhi, double2* __restrict__ spec, int64_t ld) {
    extern __shared__ double2 buf[];
    __shared__ double2 xn;
    const int r = bucket_ring(nside, qlo, qhi, blockIdx.x);
    const Ring g = ring_info(nside, r);
    const int n = g.nphi, N = n / 2, H = N / 2;
    const double w = 4.0 * M_PI / (12.0 * (double)nside * (double)nside);
    double2* out = spec + (int64_t)r * ld;
    const int W = 2 * L - 1;
    for (int comp = 0; comp < 2; ++comp) {
        const double* x = (comp == 0 ? q : u) + g.offset;
        for (int p = 0; p < 2; ++p) {
            load_half(buf, x, H, p);
            fft_lds(buf, H, -1.0);
            r2c_post_half(buf, H, p, &xn);
            // outputs m = j - (L - 1) of parity p; the same thread owns j in both components
            const int j0 = ((L - 1 + p) & 1) + 2 * (int)threadIdx.x;
            double2 ph = cscale(phase_pi(j0 - (L - 1), g.phi0_num, g.phi0_den, -1.0), w);
            const double2 rot = phase_pi(2 * blockDim.x, g.phi0_num, g.phi0_den, -1.0);
            for (int j = j0; j < W; j += 2 * blockDim.x) {
                const int m = j - (L - 1);
                int kk = m % n;
                if (kk < 0) kk += n;
                const double2 X = half_parity_x(buf, xn, N, p, kk);
                const double2 v = cmul(ph, X);
                ph = cmul(ph, rot);
                if (comp == 0) {
                    out[j] = v;
                } else {
                    const double2 o = out[j];
                    out[j] = make_double2(o.x - v.y, o.y + v.x);  // + i v
                }
            }
            __syncthreads();
        }
    }
}

// z[j] and z[j+H] of the packed ring from the parity-p inverse half y_p in buf: pass 0
// stores y_0 in both, pass 1 adds +-w^j y_1 to what this thread stored in pass 0.
__device__ inline void store_half(double* x, const double2* buf, int H, int p) {
    const int N = 2 * H;
    for (int k = threadIdx.x; k < H; k += blockDim.x) {
        const double2 y = buf[k];
        if (p == 0) {
            x[2 * k] = y.x;
            x[2 * k + 1] = y.y;
            x[2 * (k + H)] = y.x;
            x[2 * (k + H) + 1] = y.y;
        } else {
            const double2 t = cmul(twiddle(k, N, 1.0), y);
            const double2 y0 = make_double2(x[2 * k], x[2 * k + 1]);
            x[2 * k] = y0.x + t.x;
            x[2 * k + 1] = y0.y + t.y;
            x[2 * (k + H)] = y0.x - t.x;
            x[2 * (k + H) + 1] = y0.y - t.y;
        }
    }
}

__global__ void k_ring_inv_s0_split(const double2* __restrict__ spec, int64_t ld, int64_t spec_map_stride, int nside,
                                    int L, int qlo, int qhi, double* __restrict__ maps, int64_t map_stride) {
    extern __shared__ double2 buf[];
    __shared__ double2 xn;
    const int r = bucket_ring(nside, qlo, qhi, blockIdx.x);
    const Ring g = ring_info(nside, r);
    const int n = g.nphi, N = n / 2, H = N / 2;
    const double2* F = spec + blockIdx.y * spec_map_stride + (int64_t)r * ld;
    double* x = maps + blockIdx.y * map_stride + g.offset;
    for (int p = 0; p < 2; ++p) {
        for (int kh = threadIdx.x; kh <= H; kh += blockDim.x) {
            if (kh == H && p == 1) continue;
            const int k = kh == H ? N : 2 * kh + p;  // kh == H carries X[N] (parity 0)
            const double2 a = fold_sum(F, 0, 0, L - 1, k, n, g);
            const double2 b = fold_sum(F, 0, 1, L - 1, (n - k) % n, n, g);
            const double2 X = make_double2(a.x + b.x, a.y - b.y);
            if (kh < H) buf[kh] = X;
            else xn = X;
        }
        __syncthreads();
        c2r_pre_half(buf, H, p, xn);
        fft_lds(buf, H, 1.0);
        store_half(x, buf, H, p);
        __syncthreads();
    }
}

__global__ void k_ring_inv_s_split(const double2* __restrict__ spec, int64_t ld, int nside, int L, int qlo, int qhi,
                                   double* __restrict__ q, double* __restrict__ u) {
    extern __shared__ double2 buf[];
    __shared__ double2 xn;
    const int r = bucket_ring(nside, qlo, qhi, blockIdx.x);
    const Ring g = ring_info(nside, r);
    const int n = g.nphi, N = n / 2, H = N / 2;
    const double2* F = spec + (int64_t)r * ld;
    const int64_t col0 = L - 1;
    for (int comp = 0; comp < 2; ++comp) {
        double* x = (comp == 0 ? q : u) + g.offset;
        for (int p = 0; p < 2; ++p) {
            for (int kh = threadIdx.x; kh <= H; kh += blockDim.x) {
                if (kh == H && p == 1) continue;
                const int k = kh == H ? N : 2 * kh + p;
                const double2 yp = fold_sum(F, col0, -(L - 1), L - 1, k, n, g);
                const double2 ym = fold_sum(F, col0, -(L - 1), L - 1, (n - k) % n, n, g);
                double2 X;
                if (comp == 0) {
                    X = make_double2(0.5 * (yp.x + ym.x), 0.5 * (yp.y - ym.y));
                } else {
                    const double2 d = make_double2(yp.x - ym.x, yp.y + ym.y);
                    X = make_double2(0.5 * d.y, -0.5 * d.x);
                }
                if (kh < H) buf[kh] = X;
                else xn = X;
            }
            __syncthreads();
            c2r_pre_half(buf, H, p, xn);
            fft_lds(buf, H, 1.0);
            store_half(x, buf, H, p);
            __syncthreads();
        }
    }
}

// ----------------------------------------------------------------------- fold residual
// One workgroup per ring, one thread per residue class c of m modulo nphi.
// Rings with fewer residue classes than threads (the polar caps: nphi = 4 i) give each class
// tpc = blockDim / nres threads: the class sum is split over them and reduced through LDS,
// and the writes of the class are shared the same way.  Larger rings keep one thread per
// class.  The branch depends only on the ring, so every barrier is reached by the whole
// workgroup.  Each output element k is read from G and written to R by the same thread, so
// R may alias G.
constexpr int kFoldThreads = 256;

__global__ void __launch_bounds__(kFoldThreads) k_fold_s0(const double2* __restrict__ F, const double2* G, double2* R,
                                                          int64_t ld, int64_t map_stride, int nside, int L) {
    __shared__ double2 part[kFoldThreads];
    const int r = blockIdx.x;
    const Ring g = ring_info(nside, r);
    const int n = g.nphi;
    const double sigma = g.phi0_num ? -1.0 : 1.0;
    const double scale = 4.0 * M_PI / (12.0 * (double)nside * (double)nside) * (double)n;
    const double2* f = F + blockIdx.y * map_stride + (int64_t)r * ld;
    const double2* gin = G + blockIdx.y * map_stride + (int64_t)r * ld;
    double2* out = R + blockIdx.y * map_stride + (int64_t)r * ld;
    const int nres = n < L ? n : L;
    const int tpc = nres < kFoldThreads ? kFoldThreads / nres : 1;
    if (tpc == 1) {
        for (int c = threadIdx.x; c < nres; c += blockDim.x) {
            double2 s = make_double2(0.0, 0.0);
            // m = c + j n over (-L, L); F[-m] = conj F[m]
            const int64_t jlo = -(((int64_t)L - 1 + c) / n);
            for (int64_t j = jlo; c + j * n < L; ++j) {
                const int64_t m = c + j * n;
                double2 v = (m >= 0) ? f[m] : cconj(f[-m]);
                if (sigma < 0 && (j & 1)) v = make_double2(-v.x, -v.y);
                s = cadd(s, v);
            }
            s = cscale(s, scale);
            for (int64_t j = 0; c + j * n < L; ++j) {
                const int64_t k = c + j * n;
                const double sg = (sigma < 0 && (j & 1)) ? -1.0 : 1.0;
                const double2 gk = gin[k];
                out[k] = make_double2(sg * s.x - gk.x, sg * s.y - gk.y);
            }
        }
        return;
    }
    const int c = threadIdx.x / tpc, u = threadIdx.x % tpc;
    double2 s = make_double2(0.0, 0.0);
    if (c < nres) {
        const int64_t jlo = -(((int64_t)L - 1 + c) / n);
        for (int64_t j = jlo + u; c + j * n < L; j += tpc) {
            const int64_t m = c + j * n;
            double2 v = (m >= 0) ? f[m] : cconj(f[-m]);
            if (sigma < 0 && (j & 1)) v = make_double2(-v.x, -v.y);
            s = cadd(s, v);
        }
    }
    part[threadIdx.x] = s;
    __syncthreads();
    if (c < nres) {
        double2 t = make_double2(0.0, 0.0);
        for (int i = 0; i < tpc; ++i) t = cadd(t, part[c * tpc + i]);  // fixed order: deterministic
        t = cscale(t, scale);
        for (int64_t j = u; c + j * n < L; j += tpc) {
            const int64_t k = c + j * n;
            const double sg = (sigma < 0 && (j & 1)) ? -1.0 : 1.0;
            const double2 gk = gin[k];
            out[k] = make_double2(sg * t.x - gk.x, sg * t.y - gk.y);
        }
    }
}

__global__ void __launch_bounds__(kFoldThreads) k_fold_s(const double2* __restrict__ F, const double2* G, double2* R,
                                                         int64_t ld, int nside, int L) {
    __shared__ double2 part[kFoldThreads];
    const int r = blockIdx.x;
    const Ring g = ring_info(nside, r);
    const int n = g.nphi;
    const double sigma = g.phi0_num ? -1.0 : 1.0;
    const double scale = 4.0 * M_PI / (12.0 * (double)nside * (double)nside) * (double)n;
    const int64_t col0 = L - 1;
    const double2* f = F + (int64_t)r * ld + col0;
    const double2* gin = G + (int64_t)r * ld + col0;
    double2* out = R + (int64_t)r * ld + col0;
    const int tpc = n < kFoldThreads ? kFoldThreads / n : 1;
    if (tpc == 1) {
        for (int c = threadIdx.x; c < n; c += blockDim.x) {
            // all m = c + j n with |m| < L
            const int64_t jlo = -(((int64_t)L - 1 + c) / n);
            if (c + jlo * n >= L) continue;
            double2 s = make_double2(0.0, 0.0);
            for (int64_t j = jlo; c + j * n < L; ++j) {
                double2 v = f[c + j * n];
                if (sigma < 0 && (j & 1)) v = make_double2(-v.x, -v.y);
                s = cadd(s, v);
            }
            s = cscale(s, scale);
            for (int64_t j = jlo; c + j * n < L; ++j) {
                const int64_t k = c + j * n;
                const double sg = (sigma < 0 && (j & 1)) ? -1.0 : 1.0;
                const double2 gk = gin[k];
                out[k] = make_double2(sg * s.x - gk.x, sg * s.y - gk.y);
            }
        }
        return;
    }
    const int c = threadIdx.x / tpc, u = threadIdx.x % tpc;
    const int64_t jlo = (c < n) ? -(((int64_t)L - 1 + c) / n) : 0;
    double2 s = make_double2(0.0, 0.0);
    if (c < n) {
        for (int64_t j = jlo + u; c + j * n < L; j += tpc) {
            double2 v = f[c + j * n];
            if (sigma < 0 && (j & 1)) v = make_double2(-v.x, -v.y);
            s = cadd(s, v);
        }
    }
    part[threadIdx.x] = s;
    __syncthreads();
    if (c < n) {
        double2 t = make_double2(0.0, 0.0);
        for (int i = 0; i < tpc; ++i) t = cadd(t, part[c * tpc + i]);
        t = cscale(t, scale);
        for (int64_t j = jlo + u; c + j * n < L; j += tpc) {
            const int64_t k = c + j * n;
            const double sg = (sigma < 0 && (j & 1)) ? -1.0 : 1.0;
            const double2 gk = gin[k];
            out[k] = make_double2(sg * t.x - gk.x, sg * t.y - gk.y);
        }
    }
}

// ----------------------------------------------------------------------- launch helpers
struct Bucket {
    int qlo, qhi;  // caps: north indices; belt: qlo == 0
    int nblocks;
    int N;         // largest N = nphi / 2 in the bucket
};

// Size buckets: the small caps together (N = 2q <= 512: one launch instead of one per
// power of two, each of which would hold a handful of workgroups), then caps with
// N = 2q in (2^(b-1), 2^b], then the belt.
int make_buckets(int nside, Bucket* out) {
    int nb = 0;
    int q = 1;
    if (nside - 1 >= 1) {
        const int qhi = nside - 1 < 256 ? nside - 1 : 256;
        out[nb].qlo = 1;
        out[nb].qhi = qhi;
        out[nb].nblocks = 2 * qhi;
        out[nb].N = 2 * qhi;
        ++nb;
        q = qhi + 1;
    }
    while (q <= nside - 1) {
        int N = 2 * q;
        int pow2 = 1;
        while (pow2 < N) pow2 *= 2;
        int qhi = pow2 / 2;
        if (qhi > nside - 1) qhi = nside - 1;
        out[nb].qlo = q;
        out[nb].qhi = qhi;
        out[nb].nblocks = 2 * (qhi - q + 1);
        out[nb].N = 2 * qhi;
        ++nb;
        q = qhi + 1;
    }
    out[nb].qlo = 0;
    out[nb].qhi = 0;
    out[nb].nblocks = 2 * nside + 1;
    out[nb].N = 2 * nside;
    return nb + 1;
}

// At least N / 8 threads (the passes' assumption), and at least 256 so that the output
// loop over up to 2 lmax + 1 phased coefficients of a small ring is not left to one wave.
inline int threads_for(int N) {
    int t = N / 8;
    if (t < 256) t = 256;
    t = (t + 63) / 64 * 64;
    if (t > 1024) t = 1024;
    return t;
}

constexpr int kMaxNside = 8192;

// Dynamic LDS of a bucket whose rings hold up to NL points: the next power of two (the
// Bluestein pass's M-point buffer never exceeds it), at most kLdsN points.
inline size_t lds_bytes(int NL) {
    int P = 1;
    while (P < NL) P <<= 1;
    if (P > kLdsN) P = kLdsN;
    if (P < NL) P = NL;
    return (size_t)P * sizeof(double2);
}

// Caps whose rings can reach a Bluestein pass (a prime p >= kBluesteinMin with N >= 2 p)
// run the kernels built with it.  Every non-split launch carries the twiddle table
// (dcap / 8 + 1 points: power-of-two rings and Bluestein convolutions use it).
inline bool bucket_bluestein(const Bucket& b, bool split) { return !split && b.qlo > 0 && b.N >= 2 * kBluesteinMin; }
inline size_t table_bytes(int dcap) { return (size_t)(dcap / 8 + 1) * sizeof(double2); }

// Dynamic LDS above 64 KiB needs the kernel's limit raised.  It is raised per kernel and
// device to at least kLdsN points (the largest plain ring buffer) and further when a launch
// needs more (a Bluestein bucket's twiddle table), so that once every bucket size has been
// launched no other runtime call is made (HIP graph capture records launches only).
template <typename K>
int set_lds(K kernel, size_t bytes) {
    if (bytes <= 64 * 1024) return GM_OK;
    constexpr size_t kBase = (size_t)kLdsN * sizeof(double2);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(GM_E_LAUNCH, "hipGetDevice failed");
    static std::mutex mu;
    struct Limit {
        const void* fn;
        int dev;
        size_t bytes;
    };
    static std::vector<Limit> done;
    const void* fn = reinterpret_cast<const void*>(kernel);
    std::lock_guard<std::mutex> lock(mu);
    Limit* cur = nullptr;
    for (auto& d : done)
        if (d.fn == fn && d.dev == dev) cur = &d;
    if (cur && cur->bytes >= bytes) return GM_OK;
    const size_t want = bytes > kBase ? bytes : kBase;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)want);
    if (e != hipSuccess) return fail(GM_E_LAUNCH, "hipFuncSetAttribute(%zu bytes): %s", want, hipGetErrorString(e));
    if (cur)
        cur->bytes = want;
    else
        done.push_back({fn, dev, want});
    return GM_OK;
}

int check_ring_args(int64_t nside, int64_t lmax) {
    if (nside < 1 || (nside & (nside - 1)) != 0) return fail(GM_E_ARG, "nside must be a power of 2 (got %lld)", (long long)nside);
    if (nside > kMaxNside) return fail(GM_E_UNSUPPORTED, "ring FFT supports nside <= %d (got %lld)", kMaxNside, (long long)nside);
    if (lmax < 0) return fail(GM_E_ARG, "lmax must be >= 0");
    return GM_OK;
}

}  // namespace

}  // namespace gm

using namespace gm;

extern "C" int gm_ring_fft_fwd_s0(const double* maps, int64_t nmaps, int64_t map_stride, int64_t nside,
                                  int64_t lmax, double* spec, int64_t ld, int64_t spec_map_stride,
                                  void* stream) {
    if (int e = check_ring_args(nside, lmax)) return e;
    if (!maps || !spec) return fail(GM_E_ARG, "gm_ring_fft_fwd_s0: null pointer");
    const int64_t L = lmax + 1, npix = 12 * nside * nside, nring = 4 * nside - 1;
    if (nmaps < 1) return fail(GM_E_ARG, "gm_ring_fft_fwd_s0: nmaps must be >= 1");
    if (ld < L) return fail(GM_E_SHAPE, "gm_ring_fft_fwd_s0: ld (%lld) < lmax + 1", (long long)ld);
    if (nmaps > 1 && (map_stride < npix || spec_map_stride < nring * ld))
        return fail(GM_E_SHAPE, "gm_ring_fft_fwd_s0: map strides too small");
    Bucket b[32];
    const int nb = make_buckets((int)nside, b);
    for (int i = 0; i < nb; ++i) {
        const bool split = b[i].N > kLdsN;
        const int NL = split ? b[i].N / 2 : b[i].N;  // points held in LDS
        const size_t lds = lds_bytes(NL);
        const int dcap = (int)(lds / sizeof(double2));
        const dim3 grid(b[i].nblocks, (unsigned)nmaps), blk(threads_for(b[i].N));
        double2* sp = reinterpret_cast<double2*>(spec);
        if (split) {
            if (int e = set_lds(k_ring_fwd_s0_split, lds)) return e;
            hipLaunchKernelGGL(k_ring_fwd_s0_split, grid, blk, lds, (hipStream_t)stream, maps, map_stride, (int)nside,
                               (int)L, b[i].qlo, b[i].qhi, sp, ld, spec_map_stride);
        } else if (bucket_bluestein(b[i], split)) {
            const size_t lt = lds + table_bytes(dcap);
            if (int e = set_lds(k_ring_fwd_s0<true>, lt)) return e;
            hipLaunchKernelGGL(k_ring_fwd_s0<true>, grid, blk, lt, (hipStream_t)stream, maps, map_stride, (int)nside,
                               (int)L, b[i].qlo, b[i].qhi, sp, ld, spec_map_stride, dcap);
        } else {
            const size_t lt = lds + table_bytes(dcap);
            if (int e = set_lds(k_ring_fwd_s0<false>, lt)) return e;
            hipLaunchKernelGGL(k_ring_fwd_s0<false>, grid, blk, lt, (hipStream_t)stream, maps, map_stride, (int)nside,
                               (int)L, b[i].qlo, b[i].qhi, sp, ld, spec_map_stride, dcap);
        }
    }
    return check_launch("gm_ring_fft_fwd_s0");
}

extern "C" int gm_ring_fft_fwd_s(const double* q, const double* u, int64_t nside, int64_t lmax, double* spec,
                                 int64_t ld, void* stream) {
    if (int e = check_ring_args(nside, lmax)) return e;
    if (!q || !u || !spec) return fail(GM_E_ARG, "gm_ring_fft_fwd_s: null pointer");
    const int64_t L = lmax + 1;
    if (ld < 2 * L - 1) return fail(GM_E_SHAPE, "gm_ring_fft_fwd_s: ld (%lld) < 2 lmax + 1", (long long)ld);
    Bucket b[32];
    const int nb = make_buckets((int)nside, b);
    for (int i = 0; i < nb; ++i) {
        const bool split = b[i].N > kLdsN;
        const int NL = split ? b[i].N / 2 : b[i].N;  // points held in LDS
        const size_t lds = lds_bytes(NL);
        const int dcap = (int)(lds / sizeof(double2));
        const dim3 grid(b[i].nblocks), blk(threads_for(b[i].N));
        double2* sp = reinterpret_cast<double2*>(spec);
        if (split) {
            if (int e = set_lds(k_ring_fwd_s_split, lds)) return e;
            hipLaunchKernelGGL(k_ring_fwd_s_split, grid, blk, lds, (hipStream_t)stream, q, u, (int)nside, (int)L,
                               b[i].qlo, b[i].qhi, sp, ld);
        } else if (bucket_bluestein(b[i], split)) {
            const size_t lt = lds + table_bytes(dcap);
            if (int e = set_lds(k_ring_fwd_s<true>, lt)) return e;
            hipLaunchKernelGGL(k_ring_fwd_s<true>, grid, blk, lt, (hipStream_t)stream, q, u, (int)nside, (int)L,
                               b[i].qlo, b[i].qhi, sp, ld, dcap);
        } else {
            const size_t lt = lds + table_bytes(dcap);
            if (int e = set_lds(k_ring_fwd_s<false>, lt)) return e;
            hipLaunchKernelGGL(k_ring_fwd_s<false>, grid, blk, lt, (hipStream_t)stream, q, u, (int)nside, (int)L,
                               b[i].qlo, b[i].qhi, sp, ld, dcap);
        }
    }
    return check_launch("gm_ring_fft_fwd_s");
}

extern "C" int gm_ring_fft_inv_s0(const double* spec, int64_t nmaps, int64_t ld, int64_t spec_map_stride,
                                  int64_t nside, int64_t lmax, double* maps, int64_t map_stride, void* stream) {
    if (int e = check_ring_args(nside, lmax)) return e;
    if (!maps || !spec) return fail(GM_E_ARG, "gm_ring_fft_inv_s0: null pointer");
    const int64_t L = lmax + 1, npix = 12 * nside * nside, nring = 4 * nside - 1;
    if (nmaps < 1) return fail(GM_E_ARG, "gm_ring_fft_inv_s0: nmaps must be >= 1");
    if (ld < L) return fail(GM_E_SHAPE, "gm_ring_fft_inv_s0: ld < lmax + 1");
    if (nmaps > 1 && (map_stride < npix || spec_map_stride < nring * ld))
        return fail(GM_E_SHAPE, "gm_ring_fft_inv_s0: map strides too small");
    Bucket b[32];
    const int nb = make_buckets((int)nside, b);
    for (int i = 0; i < nb; ++i) {
        const bool split = b[i].N > kLdsN;
        const int NL = split ? b[i].N / 2 : b[i].N;  // points held in LDS
        const size_t lds = lds_bytes(NL);
        const int dcap = (int)(lds / sizeof(double2));
        const dim3 grid(b[i].nblocks, (unsigned)nmaps), blk(threads_for(b[i].N));
        const double2* sp = reinterpret_cast<const double2*>(spec);
        if (split) {
            if (int e = set_lds(k_ring_inv_s0_split, lds)) return e;
            hipLaunchKernelGGL(k_ring_inv_s0_split, grid, blk, lds, (hipStream_t)stream, sp, ld, spec_map_stride,
                               (int)nside, (int)L, b[i].qlo, b[i].qhi, maps, map_stride);
        } else if (bucket_bluestein(b[i], split)) {
            const size_t lt = lds + table_bytes(dcap);
            if (int e = set_lds(k_ring_inv_s0<true>, lt)) return e;
            hipLaunchKernelGGL(k_ring_inv_s0<true>, grid, blk, lt, (hipStream_t)stream, sp, ld, spec_map_stride,
                               (int)nside, (int)L, b[i].qlo, b[i].qhi, maps, map_stride, dcap);
        } else {
            const size_t lt = lds + table_bytes(dcap);
            if (int e = set_lds(k_ring_inv_s0<false>, lt)) return e;
            hipLaunchKernelGGL(k_ring_inv_s0<false>, grid, blk, lt, (hipStream_t)stream, sp, ld, spec_map_stride,
                               (int)nside, (int)L, b[i].qlo, b[i].qhi, maps, map_stride, dcap);
        }
    }
    return check_launch("gm_ring_fft_inv_s0");
}

extern "C" int gm_ring_fft_inv_s(const double* spec, int64_t ld, int64_t nside, int64_t lmax, double* q,
                                 double* u, void* stream) {
    if (int e = check_ring_args(nside, lmax)) return e;
    if (!spec || !q || !u) return fail(GM_E_ARG, "gm_ring_fft_inv_s: null pointer");
    const int64_t L = lmax + 1;
    if (ld < 2 * L - 1) return fail(GM_E_SHAPE, "gm_ring_fft_inv_s: ld < 2 lmax + 1");
    Bucket b[32];
    const int nb = make_buckets((int)nside, b);
    for (int i = 0; i < nb; ++i) {
        const bool split = b[i].N > kLdsN;
        const int NL = split ? b[i].N / 2 : b[i].N;  // points held in LDS
        const size_t lds = lds_bytes(NL);
        const int dcap = (int)(lds / sizeof(double2));
        const dim3 grid(b[i].nblocks), blk(threads_for(b[i].N));
        const double2* sp = reinterpret_cast<const double2*>(spec);
        if (split) {
            if (int e = set_lds(k_ring_inv_s_split, lds)) return e;
            hipLaunchKernelGGL(k_ring_inv_s_split, grid, blk, lds, (hipStream_t)stream, sp, ld, (int)nside, (int)L,
                               b[i].qlo, b[i].qhi, q, u);
        } else if (bucket_bluestein(b[i], split)) {
            const size_t lt = lds + table_bytes(dcap);
            if (int e = set_lds(k_ring_inv_s<true>, lt)) return e;
            hipLaunchKernelGGL(k_ring_inv_s<true>, grid, blk, lt, (hipStream_t)stream, sp, ld, (int)nside, (int)L,
                               b[i].qlo, b[i].qhi, q, u, dcap);
        } else {
            const size_t lt = lds + table_bytes(dcap);
            if (int e = set_lds(k_ring_inv_s<false>, lt)) return e;
            hipLaunchKernelGGL(k_ring_inv_s<false>, grid, blk, lt, (hipStream_t)stream, sp, ld, (int)nside, (int)L,
                               b[i].qlo, b[i].qhi, q, u, dcap);
        }
    }
    return check_launch("gm_ring_fft_inv_s");
}

extern "C" int gm_ring_fold_residual(const double* F, const double* G, double* R, int64_t nmaps, int64_t ld,
                                     int64_t map_stride, int64_t nside, int64_t lmax, void* stream) {
    if (nside < 1 || lmax < 0) return fail(GM_E_ARG, "gm_ring_fold_residual: bad nside / lmax");
    if (!F || !G || !R) return fail(GM_E_ARG, "gm_ring_fold_residual: null pointer");
    const int64_t L = lmax + 1, nring = 4 * nside - 1;
    if (ld < L) return fail(GM_E_SHAPE, "gm_ring_fold_residual: ld < lmax + 1");
    if (nmaps < 1 || (nmaps > 1 && map_stride < nring * ld))
        return fail(GM_E_SHAPE, "gm_ring_fold_residual: bad nmaps / map_stride");
    hipLaunchKernelGGL(k_fold_s0, dim3((unsigned)nring, (unsigned)nmaps), dim3(kFoldThreads), 0, (hipStream_t)stream,
                       reinterpret_cast<const double2*>(F), reinterpret_cast<const double2*>(G),
                       reinterpret_cast<double2*>(R), ld, map_stride, (int)nside, (int)L);
    return check_launch("gm_ring_fold_residual");
}

extern "C" int gm_ring_fold_residual_c(const double* F, const double* G, double* R, int64_t ld, int64_t nside,
                                       int64_t lmax, void* stream) {
    if (nside < 1 || lmax < 0) return fail(GM_E_ARG, "gm_ring_fold_residual_c: bad nside / lmax");
    if (!F || !G || !R) return fail(GM_E_ARG, "gm_ring_fold_residual_c: null pointer");
    const int64_t L = lmax + 1, nring = 4 * nside - 1;
    if (ld < 2 * L - 1) return fail(GM_E_SHAPE, "gm_ring_fold_residual_c: ld < 2 lmax + 1");
    hipLaunchKernelGGL(k_fold_s, dim3((unsigned)nring), dim3(kFoldThreads), 0, (hipStream_t)stream,
                       reinterpret_cast<const double2*>(F), reinterpret_cast<const double2*>(G),
                       reinterpret_cast<double2*>(R), ld, (int)nside, (int)L);
    return check_launch("gm_ring_fold_residual_c");
}
