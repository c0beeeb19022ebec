// C1 / C2 mask apodization on gfx950 (gmaster/utils.py:61-147): every unmasked pixel whose
// chord-squared distance d = 0.5 |v - v_q|^2 = 1 - cos(angle) to the nearest masked pixel q
// (mask <= 0) is below 1 - cos(aposize) is multiplied by f(x), x = sqrt(d / (1 - cos aposize)),
// f = x - sin(2 pi x) / (2 pi) (C1) or (1 - cos(pi x)) / 2 (C2).
//
// The nearest-neighbour search uses the ring structure instead of a k-d tree:
//   * per ring, every pixel stores the ring-local index of the cyclically previous and
//     next masked pixel (one workgroup per ring, block-wide scans);
//   * on a ring, the chord distance to a fixed pixel grows with |phi - phi_p|, so the
//     nearest masked pixel of ring j is the previous or next masked pixel of the pixel of
//     ring j nearest in phi: two candidates per ring;
//   * rings are visited outward from the pixel's own ring while their smallest possible
//     distance (the chord to the same phi on ring j) stays below the best so far and the
//     threshold.
// The result equals the exhaustive search for every pixel with d below the threshold, the
// only ones the apodization changes.  Kind 3 instead zeroes every pixel within the scale of
// a masked pixel (masked pixels included): the first step of the "Smooth" apodization.
#include <math.h>

#include "common.h"

namespace gm {
namespace {

struct RingT {
    double z, s, a0;  // cos(theta), sin(theta), phi0 / pi
    int nphi;
    int pad;
    int64_t off;
};

__global__ void k_ring_table(int nside, int nring, RingT* __restrict__ t) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nring) return;
    const Ring g = ring_info(nside, r);
    RingT o;
    o.z = g.z;
    o.s = g.sin;
    o.a0 = (double)g.phi0_num / (double)g.phi0_den;
    o.nphi = g.nphi;
    o.pad = 0;
    o.off = g.offset;
    t[r] = o;
}

constexpr int kLinkThreads = 1024;

// prev[i] / next[i]: ring-local index of the cyclically previous / next masked pixel at or
// before / at or after i; -1 when the ring has none.
__global__ void __launch_bounds__(kLinkThreads) k_masked_links(const double* __restrict__ mask,
                                                               const RingT* __restrict__ rings,
                                                               int* __restrict__ prev, int* __restrict__ next) {
    __shared__ int lastv[kLinkThreads];
    __shared__ int firstv[kLinkThreads];
    const RingT g = rings[blockIdx.x];
    const int n = g.nphi;
    const int per = (n + kLinkThreads - 1) / kLinkThreads;
    const int t = threadIdx.x;
    const int b = t * per, e = min(n, b + per);
    const double* m = mask + g.off;
    int last = -1, first = 0x7fffffff;
    for (int i = b; i < e; ++i) {
        if (m[i] <= 0.0) {
            last = i;
            if (first == 0x7fffffff) first = i;
        }
    }
    lastv[t] = last;
    firstv[t] = first;
    __syncthreads();
    // inclusive max-scan of lastv (left to right) and min-scan of firstv (right to left)
    for (int d = 1; d < kLinkThreads; d <<= 1) {
        const int a = (t >= d) ? lastv[t - d] : -1;
        const int c = (t + d < kLinkThreads) ? firstv[t + d] : 0x7fffffff;
        __syncthreads();
        lastv[t] = max(lastv[t], a);
        firstv[t] = min(firstv[t], c);
        __syncthreads();
    }
    const int ring_last = lastv[kLinkThreads - 1];
    const int ring_first = firstv[0];
    int* pv = prev + g.off;
    int* nx = next + g.off;
    if (ring_last < 0) {
        for (int i = b; i < e; ++i) {
            pv[i] = -1;
            nx[i] = -1;
        }
        return;
    }
    int run = (t > 0) ? lastv[t - 1] : -1;
    if (run < 0) run = ring_last;  // wrap around
    for (int i = b; i < e; ++i) {
        if (m[i] <= 0.0) run = i;
        pv[i] = run;
    }
    run = (t + 1 < kLinkThreads) ? firstv[t + 1] : 0x7fffffff;
    if (run == 0x7fffffff) run = ring_first;
    for (int i = e - 1; i >= b; --i) {
        if (m[i] <= 0.0) run = i;
        nx[i] = run;
    }
}

// 0.5 |v_p - v_q|^2 between (theta_p, phi_p) and (theta_q, phi_q) without cancellation:
// 0.5 [(z_p - z_q)^2 + (s_p - s_q)^2] + 2 s_p s_q sin^2((phi_q - phi_p) / 2), da = (phi_q - phi_p) / pi
__device__ inline double chord2(double base, double ss, double da) {
    const double h = sinpi(0.5 * da);
    return fma(2.0 * ss * h, h, base);
}

// smallest d from the pixel over ring j's candidates, below `best`; base = the ring's
// same-phi term, ss = s_p s_j
__device__ inline double ring_best(const RingT& q, const int* __restrict__ prev, const int* __restrict__ next,
                                   double ap, double base, double ss, double best) {
    // pixel of ring q nearest in phi: k = round((phi_p - phi0) nphi / (2 pi)) mod nphi
    double t = (ap - q.a0) * 0.5 * q.nphi;
    int64_t k = (int64_t)floor(t + 0.5);
    k %= q.nphi;
    if (k < 0) k += q.nphi;
    const int c0 = prev[q.off + k];
    if (c0 < 0) return best;
    const int c1 = next[q.off + k];
    best = fmin(best, chord2(base, ss, q.a0 + 2.0 * (double)c0 / (double)q.nphi - ap));
    if (c1 != c0) best = fmin(best, chord2(base, ss, q.a0 + 2.0 * (double)c1 / (double)q.nphi - ap));
    return best;
}

__global__ void k_apodize(const double* __restrict__ mask, const RingT* __restrict__ rings, int nring,
                          const int* __restrict__ prev, const int* __restrict__ next, double thr, int kind,
                          double* __restrict__ out) {
    const int r = blockIdx.x;
    const RingT g = rings[r];
    for (int k = threadIdx.x; k < g.nphi; k += blockDim.x) {
        const int64_t p = g.off + k;
        const double mv = mask[p];
        if (!(mv > 0.0)) {
            out[p] = kind == 3 ? 0.0 : mv;
            continue;
        }
        const double ap = g.a0 + 2.0 * (double)k / (double)g.nphi;
        double best = thr;
        for (int j = r; j >= 0; --j) {
            const RingT q = rings[j];
            const double ds = q.s - g.s, dz = q.z - g.z;
            const double base = 0.5 * (ds * ds + dz * dz);
            if (base >= best) break;
            best = ring_best(q, prev, next, ap, base, q.s * g.s, best);
        }
        for (int j = r + 1; j < nring; ++j) {
            const RingT q = rings[j];
            const double ds = q.s - g.s, dz = q.z - g.z;
            const double base = 0.5 * (ds * ds + dz * dz);
            if (base >= best) break;
            best = ring_best(q, prev, next, ap, base, q.s * g.s, best);
        }
        double o = mv;
        if (best < thr && kind == 3) {
            o = 0.0;
        } else if (best < thr) {
            const double x = sqrt(fmax(best, 0.0) / thr);
            double f;
            if (kind == 1) {
                f = x - sinpi(2.0 * x) / (2.0 * M_PI);
            } else {
                f = 0.5 * (1.0 - cospi(x));
            }
            o = mv * f;
        }
        out[p] = o;
    }
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace
}  // namespace gm

using namespace gm;

extern "C" int64_t gm_apodize_work_size(int64_t nside) {
    if (nside < 1 || nside > 8192 || (nside & (nside - 1))) return 0;
    const int64_t npix = 12 * nside * nside, nring = 4 * nside - 1;
    return (int64_t)(align256(nring * sizeof(RingT)) + 2 * align256(npix * sizeof(int)));
}

extern "C" int gm_mask_apodize(const double* mask, int64_t nside, double aposize_rad, int64_t kind, double* out,
                               void* work, void* stream) {
    if (!mask || !out || !work) return fail(GM_E_ARG, "gm_mask_apodize: null pointer");
    if (nside < 1 || nside > 8192 || (nside & (nside - 1)))
        return fail(GM_E_ARG, "gm_mask_apodize: nside must be a power of 2 <= 8192");
    if (kind < 1 || kind > 3) return fail(GM_E_ARG, "gm_mask_apodize: kind must be 1 (C1), 2 (C2) or 3 (zero)");
    if (!(aposize_rad > 0.0) || aposize_rad > M_PI) return fail(GM_E_ARG, "gm_mask_apodize: aposize out of (0, pi]");
    const int64_t npix = 12 * nside * nside, nring = 4 * nside - 1;
    char* w = (char*)work;
    RingT* rings = (RingT*)w;
    int* prev = (int*)(w + align256(nring * sizeof(RingT)));
    int* next = (int*)(w + align256(nring * sizeof(RingT)) + align256(npix * sizeof(int)));
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_ring_table, dim3((unsigned)((nring + 255) / 256)), dim3(256), 0, st, (int)nside, (int)nring,
                       rings);
    hipLaunchKernelGGL(k_masked_links, dim3((unsigned)nring), dim3(kLinkThreads), 0, st, mask, rings, prev, next);
    const double thr = 1.0 - cos(aposize_rad);
    hipLaunchKernelGGL(k_apodize, dim3((unsigned)nring), dim3(256), 0, st, mask, rings, (int)nring, prev, next, thr,
                       (int)kind, out);
    return check_launch("gm_mask_apodize");
}
