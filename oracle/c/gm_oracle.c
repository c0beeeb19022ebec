/*
 * FP64 C/OpenMP restatement of the two dominant stages of the MASTER path -- ORACLE /
 * CPU baseline only (test infrastructure; never linked by the product library).
 *
 *  gmo_lat_ana_s   spin-s Legendre analysis over all rings, direct and mirror channels:
 *                  a_D[l] = sum_r lambda_lm(theta_r) G[r][m],
 *                  a_M[l] = sum_r lambda_lm(theta_r) G[mirror r][-m]
 *                  (restates gmaster/_native/cuda/march_v2.cu:155-285 in FP64, with the
 *                  recurrence and seeds of docs/notes/latitudinal_march_maths.md:28-128)
 *  gmo_lat_syn_s   spin-s Legendre synthesis per order m over all rings, direct and mirror:
 *                  D[r] = sum_l cd[l] lambda_lm(theta_r),
 *                  Mi[r] = sum_l cm[l] lambda_lm(pi - theta_r)
 *                  (restates gmaster/_native/cuda/march_v2.cu:291-407 in FP64; the same
 *                  natural-form recurrence as the analysis); gmo_lat_syn_s_ext, the same in
 *                  x87 extended precision
 *  gmo_coupling    the four MASTER coupling sums per (l1, l2 >= l1) from the Wigner-3j
 *                  rows (l1 l2 l3; 0 0 0), (l1 l2 l3; 2 -2 0) (gmaster/workspaces.py:1069-1077)
 *
 * Both are plain scalar loops (one order / one row per OpenMP iteration).  The Legendre
 * recurrence is written in its natural form lambda_l = (A x + C) lambda_{l-1} - B lambda_{l-2}
 * with an extended exponent rescaled every 8 steps, independently of the HIP kernels'
 * mu-rescaled form.  Pinned against the numpy oracle in tests/test_oracle_c.py.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

typedef struct {
    double z, shalf, chalf;
} ring_t;

static ring_t ring_geom(int nside, int r) {
    ring_t g;
    const int i = r + 1;
    const int south = i > 2 * nside;
    const int north = south ? 4 * nside - i : i;
    const double n = nside, q = north;
    double om, zn;
    if (north < nside) {
        om = q * q / (3.0 * n * n);
        zn = 1.0 - om;
    } else {
        om = (2.0 * q - n) / (3.0 * n);
        zn = 2.0 * (2.0 * n - q) / (3.0 * n);
    }
    const double sh = sqrt(0.5 * om), ch = sqrt(0.5 * (1.0 + zn));
    g.z = south ? -zn : zn;
    g.shalf = south ? ch : sh;
    g.chalf = south ? sh : ch;
    return g;
}

/* log2 |lambda_M| constant: 0.5 log2((2M+1)/4pi) + 0.5 log2 C(2M, M + min(m, s)) */
static double seed_log2_const(int m, int s) {
    const int M = m > s ? m : s, k = m < s ? m : s;
    const double lb = (lgamma(2.0 * M + 1.0) - lgamma(M + k + 1.0) - lgamma(M - k + 1.0)) / log(2.0);
    return 0.5 * log2((2.0 * M + 1.0) / (4.0 * M_PI)) + 0.5 * lb;
}

/* Natural-form coefficients of order m, spin s:
 * lambda_{l+1} = (A[l+1] z + C[l+1]) lambda_l - B[l+1] lambda_{l-1}. */
static void rec_coefficients(int L, int m, int s, double* A, double* C, double* B) {
    const int M = m > s ? m : s;
    for (int lp = M + 1; lp < L; ++lp) {
        const int l = lp - 1;
        const double K = (l + 1.0) * (2.0 * l + 1.0) /
                         sqrt(((l + 1.0) * (l + 1.0) - (double)m * m) * ((l + 1.0) * (l + 1.0) - (double)s * s));
        A[lp] = sqrt((2.0 * l + 3.0) / (2.0 * l + 1.0)) * K;
        C[lp] = l > 0 ? A[lp] * m * s / ((double)l * (l + 1.0)) : 0.0;
        B[lp] = l > 0 ? sqrt((2.0 * l + 3.0) / (2.0 * l - 1.0)) * K *
                            sqrt(((double)l * l - (double)m * m) * ((double)l * l - (double)s * s)) /
                            ((double)l * (2.0 * l + 1.0))
                      : 0.0;
    }
}

/* One order m, synthesis: for every ring q, lambda_lm(theta_q) for l = M..L-1 feeds the
 * direct sum of ring q and the mirror sum of ring nring-1-q (whose reflection is q).
 * out[4 r + (dre, dim, mre, mim)]; cd, cm complex (re, im) per degree.  The working type
 * REAL is double (the FP64 restatement) or long double (x87 extended, 64-bit mantissa:
 * the recurrence's own rounding at the polar rings, ~2e-10 relative for m = 0 at
 * L = 12288 in FP64, drops by 2^-11, which makes it a reference for FP64 results). */
#define DEFINE_SYN_ONE(NAME, REAL, FABS, LDEXP, FREXP, EXP2)                                        \
    static void NAME(int nside, int lmax, int s, int m, const double* cd, const double* cm, double* out) { \
        const int L = lmax + 1;                                                                    \
        const int nring = 4 * nside - 1;                                                           \
        const int M = m > s ? m : s;                                                               \
        for (int k = 0; k < 4 * nring; ++k) out[k] = 0.0;                                          \
        if (M >= L) return;                                                                        \
        double* A = (double*)malloc(sizeof(double) * 3 * L);                                       \
        double* C = A + L;                                                                         \
        double* B = A + 2 * L;                                                                     \
        rec_coefficients(L, m, s, A, C, B);                                                        \
        const double c0 = seed_log2_const(m, s);                                                   \
        const REAL sign = ((m + s) & 1) ? -1.0 : 1.0;                                              \
        for (int q = 0; q < nring; ++q) {                                                          \
            const ring_t g = ring_geom(nside, q);                                                  \
            const int qm = nring - 1 - q;                                                          \
            REAL dre = 0, dim = 0, mre = 0, mim = 0;                                               \
            const double f = c0 + (m + s) * log2(g.shalf) + abs(m - s) * log2(g.chalf);            \
            int e = (int)floor(f);                                                                 \
            REAL v = sign * EXP2((REAL)(f - e)), vp = 0.0;                                         \
            REAL scale = (e > -1000) ? LDEXP((REAL)1.0, e) : (REAL)0.0;                            \
            for (int l = M; l < L; ++l) {                                                          \
                if (scale != 0.0) {                                                                \
                    const REAL lam = v * scale;                                                    \
                    dre += lam * cd[2 * l];                                                        \
                    dim += lam * cd[2 * l + 1];                                                    \
                    mre += lam * cm[2 * l];                                                        \
                    mim += lam * cm[2 * l + 1];                                                    \
                }                                                                                  \
                if (l + 1 < L) {                                                                   \
                    const REAL vn = ((REAL)A[l + 1] * (REAL)g.z + (REAL)C[l + 1]) * v - (REAL)B[l + 1] * vp; \
                    vp = v;                                                                        \
                    v = vn;                                                                        \
                    if (((l - M) & 7) == 7) {                                                      \
                        const REAL big = FABS(v) > FABS(vp) ? FABS(v) : FABS(vp);                  \
                        if (big > 0x1p256 || (big < 0x1p-256 && big > 0)) {                        \
                            int t;                                                                 \
                            FREXP(big, &t);                                                        \
                            v = LDEXP(v, -t);                                                      \
                            vp = LDEXP(vp, -t);                                                    \
                            e += t;                                                                \
                            scale = (e > -1000) ? LDEXP((REAL)1.0, e) : (REAL)0.0;                 \
                        }                                                                          \
                    }                                                                              \
                }                                                                                  \
            }                                                                                      \
            out[4 * q + 0] = (double)dre;                                                          \
            out[4 * q + 1] = (double)dim;                                                          \
            out[4 * qm + 2] = (double)mre;                                                         \
            out[4 * qm + 3] = (double)mim;                                                         \
        }                                                                                          \
        free(A);                                                                                   \
    }

DEFINE_SYN_ONE(syn_one, double, fabs, ldexp, frexp, exp2)
DEFINE_SYN_ONE(syn_one_ext, long double, fabsl, ldexpl, frexpl, exp2l)

/* One order m: acc[4 l + (dre, dim, mre, mim)] = sum_r lambda_lm(theta_r) (direct, mirror),
 * the direct value of ring r at dbase + 2 r dstride, the mirror value of ring r (the
 * spectrum at -m of the ring reflected through the equator) at mbase + 2 r mstride. */
static void ana_one(int nside, int lmax, int s, int m, const double* dbase, int64_t dstride, const double* mbase,
                    int64_t mstride, double* acc) {
    const int L = lmax + 1;
    const int nring = 4 * nside - 1;
    const int M = m > s ? m : s;
    for (int l = 0; l < 4 * L; ++l) acc[l] = 0.0;
    if (M >= L) return;
    double* A = (double*)malloc(sizeof(double) * 3 * L);
    double* C = A + L;
    double* B = A + 2 * L;
    rec_coefficients(L, m, s, A, C, B);
    const double c0 = seed_log2_const(m, s);
    const double sign = ((m + s) & 1) ? -1.0 : 1.0;
    for (int r = 0; r < nring; ++r) {
        const ring_t g = ring_geom(nside, r);
        const int rm = nring - 1 - r;
        const double* gd = dbase + 2 * (int64_t)r * dstride;
        const double* gm = mbase + 2 * (int64_t)rm * mstride;
        const double dre = gd[0], dim = gd[1];
        const double mre = m ? gm[0] : 0.0, mim = m ? gm[1] : 0.0;
        const double f = c0 + (m + s) * log2(g.shalf) + abs(m - s) * log2(g.chalf);
        int e = (int)floor(f);
        double v = sign * exp2(f - e), vp = 0.0;
        double scale = (e > -1000) ? ldexp(1.0, e) : 0.0;
        for (int l = M; l < L; ++l) {
            if (scale != 0.0) {
                const double lam = v * scale;
                acc[4 * l + 0] += lam * dre;
                acc[4 * l + 1] += lam * dim;
                acc[4 * l + 2] += lam * mre;
                acc[4 * l + 3] += lam * mim;
            }
            if (l + 1 < L) {
                const double vn = (A[l + 1] * g.z + C[l + 1]) * v - B[l + 1] * vp;
                vp = v;
                v = vn;
                if (((l - M) & 7) == 7) {
                    const double big = fabs(v) > fabs(vp) ? fabs(v) : fabs(vp);
                    if (big > 0x1p256 || (big < 0x1p-256 && big > 0)) {
                        int t;
                        frexp(big, &t);
                        v = ldexp(v, -t);
                        vp = ldexp(vp, -t);
                        e += t;
                        scale = (e > -1000) ? ldexp(1.0, e) : 0.0;
                    }
                }
            }
        }
    }
    free(A);
}

void gmo_lat_ana_s(int nside, int lmax, int spin, const double* spec, int64_t ld, const int* orders,
                   int norders, double* out) {
    const int L = lmax + 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int io = 0; io < norders; ++io) {
        const int m = orders[io];
        ana_one(nside, lmax, spin, m, spec + 2 * ((int64_t)(L - 1) + m), ld, spec + 2 * ((int64_t)(L - 1) - m), ld,
                out + (int64_t)io * L * 4);
    }
}

/* The same from per-order columns: dcol[io][r] = spec[r][L-1+m], mcol[io][r] = spec[r][L-1-m]
 * (complex, nring per order), for geometries whose full spectrum is too large for the host. */
void gmo_lat_ana_s_cols(int nside, int lmax, int spin, const double* dcol, const double* mcol, const int* orders,
                        int norders, double* out) {
    const int L = lmax + 1;
    const int64_t nring = 4 * (int64_t)nside - 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int io = 0; io < norders; ++io)
        ana_one(nside, lmax, spin, orders[io], dcol + 2 * io * nring, 1, mcol + 2 * io * nring, 1,
                out + (int64_t)io * L * 4);
}

/* Synthesis per sampled order: cd[io][l], cm[io][l] complex (L per order) ->
 * out[io][r][4] = (direct re, im, mirror re, im) over all 4 nside - 1 rings. */
void gmo_lat_syn_s(int nside, int lmax, int spin, const double* cd, const double* cm, const int* orders,
                   int norders, double* out) {
    const int L = lmax + 1;
    const int64_t nring = 4 * (int64_t)nside - 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int io = 0; io < norders; ++io)
        syn_one(nside, lmax, spin, orders[io], cd + 2 * (int64_t)io * L, cm + 2 * (int64_t)io * L,
                out + (int64_t)io * nring * 4);
}

/* The same in x87 extended precision (recurrence, seeds' mantissa and sums). */
void gmo_lat_syn_s_ext(int nside, int lmax, int spin, const double* cd, const double* cm, const int* orders,
                       int norders, double* out) {
    const int L = lmax + 1;
    const int64_t nring = 4 * (int64_t)nside - 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int io = 0; io < norders; ++io)
        syn_one_ext(nside, lmax, spin, orders[io], cd + 2 * (int64_t)io * L, cm + 2 * (int64_t)io * L,
                    out + (int64_t)io * nring * 4);
}

void gmo_coupling(int lmax, const double* W, int lmax_w, const double* g, const int* rows, int nrows,
                  double* out) {
    const int L = lmax + 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int ir = 0; ir < nrows; ++ir) {
        const int l1 = rows[ir];
        double* o = out + (int64_t)ir * L * 4;
        for (int k = 0; k < 4 * L; ++k) o[k] = 0.0;
        for (int l2 = l1; l2 < L; ++l2) {
            const int d = l2 - l1;
            const double S = l1 + l2 + 1.0, dd = d;
            const int l3max = (l1 + l2 < lmax_w) ? l1 + l2 : lmax_w;
            double w0 = sqrt(g[l1] * g[d] / ((2.0 * l2 + 1.0) * g[l2]));
            if (l2 & 1) w0 = -w0;
            double w2 = l1 >= 2 ? w0 * sqrt((double)(l2 + 1) * (l2 + 2) * l1 * (l1 - 1) /
                                             ((double)l2 * (l2 - 1) * (l1 + 1) * (l1 + 2)))
                                : 0.0;
            double p0 = 0, p2 = 0, Qc = 0, a00 = 0, a02 = 0, aE = 0, aO = 0;
            int l3 = d;
            if (d == 0) {
                if (l3max >= 0) {
                    a00 = W[0] * w0 * w0;
                    a02 = W[0] * w0 * w2;
                    aE = W[0] * w2 * w2;
                }
                p0 = w0;
                p2 = w2;
                w0 = 0.0;
                w2 = l1 >= 2 ? ((l1 & 1) ? -2.0 : 2.0) / sqrt((double)l1 * (l1 + 1) * (2.0 * l1 + 1.0)) : 0.0;
                l3 = 1;
                Qc = sqrt((1.0 - dd * dd) * (S * S - 1.0));
            }
            for (; l3 <= l3max; ++l3) {
                const double t = W[l3] * (2.0 * l3 + 1.0);
                a00 += t * w0 * w0;
                a02 += t * w0 * w2;
                if (((l1 + l2 + l3) & 1) == 0) aE += t * w2 * w2;
                else aO += t * w2 * w2;
                if (l3 == l3max) break;
                const double nn = l3 + 1.0;
                const double Qn = sqrt((nn - dd) * (nn + dd) * (S - nn) * (S + nn));
                const double n0 = -Qc * p0 / Qn;
                const double n2 = (4.0 * (2.0 * l3 + 1.0) * w2 - Qc * p2) / Qn;
                p0 = w0;
                p2 = w2;
                w0 = n0;
                w2 = n2;
                Qc = Qn;
            }
            o[4 * l2 + 0] = a00;
            o[4 * l2 + 1] = a02;
            o[4 * l2 + 2] = aE;
            o[4 * l2 + 3] = aO;
        }
    }
}
