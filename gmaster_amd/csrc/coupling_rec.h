// The l3 recursion of the mode-coupling row sums for one pair l1 <= l2, without square
// roots (host and device).
//
// Both Wigner rows f0 = (l1 l2 l3; 0 0 0) and f2 = (l1 l2 l3; 2 -2 0) obey, for m3 = 0,
//     Q(l3+1) f(l3+1) = c(l3) f(l3) - Q(l3) f(l3-1),
//     Q(l3)^2 = q(l3) = (l3^2 - d^2) ((l1+l2+1)^2 - l3^2),   d = l2 - l1,
// with c = 4 (2 l3 + 1) for f2 and c = 0 for f0, started from the closed forms at l3 = d
// and run forward (stable for |m| <= 2 up to l3 = l1 + l2).  Scaled by
// pi(l3) = Q(d+1) ... Q(l3), the values F = f pi obey
//     F(l3+1) = c(l3) F(l3) - q(l3) F(l3-1),
// which needs only the integer products q, so no square root or division per step.  The
// sums sum_l3 W'_l3 f^2 = sum W' F^2 / pi^2 are accumulated in Horner form,
// G <- G * (pi_new / pi_old)^2 + W' F^2, with pi^2 tracked as P (every even and every odd
// degree has its own chain: the 0-0, 0-2 and 2-2+ sums take even J = l1 + l2 + l3, the
// 2-2- sum odd J), and divided out once at the end.  P grows by at most S^4 / 4 < 2^57
// per degree (S = l1 + l2 + 1 < 2^15); every 4 pairs of degrees F, P and G are rescaled
// by powers of two (exact), so nothing leaves the FP64 range.
//
// d = 0: the unnormalised recursion (Schulten-Gordon) is empty at l3 = 0, but the step
// F2(1) = c(0) F2(0) = 4 f2(0) is still right: with (l l 0; 2 -2 0) = (-1)^l / sqrt(2l+1),
// (l l 1; 2 -2 0) = (-1)^l 2 / sqrt(l (l+1)(2l+1)) and Q(1) = sqrt((2l+1)^2 - 1), both sides
// equal 4 (-1)^l / sqrt(2l+1).
#pragma once

#include <hip/hip_runtime.h>

namespace gm {

struct CoupleRun {
    double F2, F2p, F0;      // F2 at the current even degree, at the one before it, F0 at the even degree
    double P;                // pi^2 at the current even degree
    double G00, G02, GE, GO;  // Horner sums: 0-0, 0-2, 2-2 even J (scale P), 2-2 odd J (scale of the odd chain)
    double A, B, st, c;      // A = (l3+1)^2 - d^2, B = S^2 - (l3+1)^2, st = 2 l3 + 3, c = 4 (2 l3 + 1)
    double qe, qo, me;       // q at the current even degree, at the odd degree after it; P ratio of the last even step

    // g[p] = Gamma(p + 1/2) / (sqrt(pi) Gamma(p + 1)), p <= l2.
    __host__ __device__ void init(const double* __restrict__ g, int l1, int l2) {
        const int d = l2 - l1;
        const double S2 = (double)(l1 + l2 + 1) * (double)(l1 + l2 + 1);
        double w0 = sqrt(g[l1] * g[d] / ((2.0 * l2 + 1.0) * g[l2]));
        if (l2 & 1) w0 = -w0;
        double w2 = 0.0;
        if (l1 >= 2)
            w2 = w0 * sqrt((double)(l2 + 1) * (l2 + 2) * l1 * (l1 - 1) / ((double)l2 * (l2 - 1) * (l1 + 1) * (l1 + 2)));
        F0 = w0;
        F2 = w2;
        F2p = 0.0;
        P = 1.0;
        G00 = G02 = GE = GO = 0.0;
        A = 2.0 * d + 1.0;
        B = S2 - (double)(d + 1) * (d + 1);
        st = 2.0 * d + 3.0;
        c = 4.0 * (2.0 * d + 1.0);
        qe = 0.0;
        qo = 0.0;
        me = 1.0;
    }

    // Terms of the even degree l3 (weight we), step to l3 + 1, term of l3 + 1 (weight wo).
    __host__ __device__ __forceinline__ void half_a(double we, double wo) {
        const double t0 = we * F0, t2 = we * F2;
        G00 = fma(G00, me, t0 * F0);
        G02 = fma(G02, me, t0 * F2);
        GE = fma(GE, me, t2 * F2);
        qo = A * B;
        A += st;
        B -= st;
        st += 2.0;
        const double F2o = fma(c, F2, -(qe * F2p));
        c += 8.0;
        GO = fma(GO, qe * qo, (wo * F2o) * F2o);
        F2p = F2o;
    }

    // Step from the odd degree to the next even one.
    __host__ __device__ __forceinline__ void half_b() {
        const double qn = A * B;
        A += st;
        B -= st;
        st += 2.0;
        const double F2n = fma(c, F2p, -(qo * F2));
        c += 8.0;
        F0 = -qo * F0;
        me = qo * qn;
        P *= me;
        F2 = F2n;
        qe = qn;
    }

    // Terms of the last (even) degree.
    __host__ __device__ __forceinline__ void even_term(double we) {
        const double t0 = we * F0, t2 = we * F2;
        G00 = fma(G00, me, t0 * F0);
        G02 = fma(G02, me, t0 * F2);
        GE = fma(GE, me, t2 * F2);
    }

    // Exact power-of-two rescale bringing P into [1/2, 2).
    __host__ __device__ __forceinline__ void rescale() {
        int e;
        (void)frexp(P, &e);
        const int h = e >> 1;
        F2 = ldexp(F2, -h);
        F2p = ldexp(F2p, -h);
        F0 = ldexp(F0, -h);
        P = ldexp(P, -2 * h);
        G00 = ldexp(G00, -2 * h);
        G02 = ldexp(G02, -2 * h);
        GE = ldexp(GE, -2 * h);
        GO = ldexp(GO, -2 * h);
    }

    // Row sums (a00, a02, aE, aO).  ended_even: the last degree was even (then the odd chain
    // ended one degree earlier, at pi^2 = P / qe); otherwise it ended at P * qo.
    __host__ __device__ void finish(bool ended_even, double s[4]) const {
        const double ri = 1.0 / P;
        s[0] = G00 * ri;
        s[1] = G02 * ri;
        s[2] = GE * ri;
        s[3] = ended_even ? GO * qe * ri : GO * ri / qo;
    }
};

// Degrees l3 = d .. min(l1 + l2, lmax_w) of one pair, W' read from Wp; s = (a00, a02, aE, aO).
__host__ __device__ inline void couple_pair(const double* __restrict__ Wp, int lmax_w, const double* __restrict__ g,
                                            int l1, int l2, double s[4]) {
    const int d = l2 - l1;
    const int n = min(l1 + l2, lmax_w) - d + 1;  // degrees with a term
    if (n <= 0) {
        s[0] = s[1] = s[2] = s[3] = 0.0;
        return;
    }
    CoupleRun r;
    r.init(g, l1, l2);
    const int np = n >> 1;
    const bool odd = n & 1;
    const int nfull = odd ? np : np - 1;
    int l3 = d;
    for (int j = 0; j < nfull; ++j) {
        r.half_a(Wp[l3], Wp[l3 + 1]);
        r.half_b();
        l3 += 2;
        if ((j & 3) == 3) r.rescale();
    }
    if (odd)
        r.even_term(Wp[l3]);
    else
        r.half_a(Wp[l3], Wp[l3 + 1]);
    r.finish(odd, s);
}

// ---------------------------------------------------------------- full matrix, one l1 per wave
// Work items are (l1, chunk j): the wave takes l2 = l1 + 64 j + lane.  Items run in
// descending l1 (longest rows first): rows l1 in [L - 64 (t+1), L - 64 t - 1] have t + 1
// chunks each, so 32 t (t + 1) items precede group t and row-major order inside it follows.
__host__ __device__ inline int64_t couple_items(int L) {
    int64_t n = 0;
    for (int t = 0; 64 * t < L; ++t) n += (int64_t)min(64, L - 64 * t) * (t + 1);
    return n;
}

__host__ __device__ inline void couple_decode(int64_t item, int L, int& l1, int& j) {
    int t = (int)((sqrt(1.0 + (double)item / 8.0) - 1.0) * 0.5);
    while (32LL * (t + 1) * (t + 2) <= item) ++t;
    while (t > 0 && 32LL * t * (t + 1) > item) --t;
    const int64_t r = item - 32LL * t * (t + 1);
    l1 = L - 1 - 64 * t - (int)(r / (t + 1));
    j = (int)(r % (t + 1));
}

// One lane of a wave's item: row l1, columns l2 = l2_0 + lane (l2_0 >= l1).  All lanes run
// the degree count of lane 0 (the smallest d, so the longest row): a lane with fewer
// degrees in the window reads zeros past lmax_w (the window is zero-padded), and every lane
// stays inside its triangle (l3 <= l1 + l2).  Win supplies W' per lane: stage(k0) loads
// W'[d0 + k0 + t], t < 128, d0 = l2_0 - l1 (zero past lmax_w), and at(t) returns entry t;
// lane i at degree l3 = d0 + i + k reads entry i + k - k0, with k0 = 64 floor(k / 64).
// Lanes with l2 >= L run the row of l2 = L - 1; the caller stores only the lanes it owns.
template <class Win>
__host__ __device__ inline void couple_wave_lane(Win& win, int lmax_w, const double* __restrict__ g, int L, int l1,
                                                 int l2_0, int lane, double s[4]) {
    const int d0 = l2_0 - l1;
    const int n = min(l1 + l2_0, lmax_w) - d0 + 1;  // lane 0's degrees (all lanes run these)
    const int l2 = min(l2_0 + lane, L - 1);
    if (n <= 0) {
        s[0] = s[1] = s[2] = s[3] = 0.0;
        return;
    }
    CoupleRun r;
    r.init(g, l1, l2);
    const int np = n >> 1;
    const bool odd = n & 1;
    const int nfull = odd ? np : np - 1;
    win.stage(0);
    int p = 0;  // pair index: degrees d + 2p, d + 2p + 1
    for (; p + 4 <= nfull; p += 4) {
        if (p && (p & 31) == 0) win.stage(2 * p);
        const int o = lane + 2 * (p & 31);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            r.half_a(win.at(o + 2 * u), win.at(o + 2 * u + 1));
            r.half_b();
        }
        r.rescale();
    }
    for (; p < nfull; ++p) {
        if (p && (p & 31) == 0) win.stage(2 * p);
        const int o = lane + 2 * (p & 31);
        r.half_a(win.at(o), win.at(o + 1));
        r.half_b();
    }
    if (p && (p & 31) == 0) win.stage(2 * p);
    const int o = lane + 2 * (p & 31);
    if (odd)
        r.even_term(win.at(o));
    else
        r.half_a(win.at(o), win.at(o + 1));
    r.finish(odd, s);
}

}  // namespace gm
