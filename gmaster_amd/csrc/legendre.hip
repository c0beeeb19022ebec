// Latitudinal (Legendre / Wigner-d) stage of the HEALPix transforms on gfx950, FP64.
//
// For each order m the normalised rows lambda_lm(theta) = sqrt((2l+1)/4pi) d^l_{m,-s}
// are generated on the fly by the three-term recurrence in l, rescaled so that every
// step is two FP64 VALU operations:
//     mu_l = (alpha_l x + gamma_l) mu_{l-1} - mu_{l-2},   lambda_l = P_l mu_l
// (alpha, gamma, P per (l, m) come from gm_lat_tables; gamma = 0 for spin 0).  Lanes are
// rings: every lane carries R rings in registers (recurrence state + ring Fourier
// coefficients), a workgroup of NW waves covers NW*64*R rings of one order m, and the
// contraction over rings is reduced through LDS once per block of 4 degrees.
//
// Extended exponent: a ring whose seed lambda_M lies below 2^-900 starts with k < 0 and
// mu = lambda 2^(600|k|) <= 2^-300; whenever |mu| passes 2^-300 the lane is rescaled by
// 2^-600 and k moves towards 0.  Lanes with k < 0 therefore hold |mu| < 2^-268 and add
// nothing measurable to any sum, so the contraction needs no per-lane select.
//
// Rings are ordered from the equator to the poles and the orders m whose polar-skip
// bound mlim(theta) (ducc0's rule, docs/notes/latitudinal_march_maths.md:245-250) is
// exceeded are simply not visited: nactive[m] is a prefix length.
#include <type_traits>

#include "common.h"
#include "recurrence.h"

namespace gm {

namespace {

template <typename F> struct cpx { F x, y; };

template <typename F>
__device__ inline cpx<F> load_c(const double2* p) {
    const double2 v = *p;
    return cpx<F>{(F)v.x, (F)v.y};
}

// ------------------------------------------------------------------ block reduction
// acc[NV] per thread -> out[0..NV) = sum over all NW*64 threads; one barrier per call
// (double-buffered LDS: `buf` alternates between calls).  Threads o*TPO .. (o+1)*TPO-1
// sum output o, TPO = 64*NW / NV, each over NV thread slots strided by TPO (bank-conflict
// free), then combine with TPO-lane shuffles.  The result is stored in FP64.
template <int NV, int NW, typename F>
__device__ inline void block_reduce(const F (&acc)[NV], F* lds, int buf, double* out) {
    constexpr int T = 64 * NW;
    constexpr int TPO = T / NV;
    F* base = lds + buf * (NV * T);
    const int tid = threadIdx.x;
#pragma unroll
    for (int v = 0; v < NV; ++v) base[v * T + tid] = acc[v];
    __syncthreads();
    const int o = tid / TPO, sl = tid % TPO;
    // thread slot sl sums lanes sl, sl + TPO, ... : consecutive threads read consecutive words
    const F* src = base + o * T + sl;
    F s = 0;
#pragma unroll
    for (int u = 0; u < NV; ++u) s += src[u * TPO];
#pragma unroll
    for (int msk = 1; msk < TPO; msk <<= 1) s += __shfl_xor(s, msk);
    if (sl == 0) out[o] = (double)s;
}

// ------------------------------------------------------------------ spin-0 analysis
// Folded lanes: j = 0 is the equator ring, j > 0 the pair (north r = 2n-1-j, south 2n-1+j).
// The recurrence always runs in FP64; F is the arithmetic type of the contraction: each
// step's mu is rounded to F once and multiplied into F accumulators (spectra are read as
// FP64 and converted, partial sums stored in FP64).
template <int K, int R, int NW, typename F>
__global__ __launch_bounds__(64 * NW) void k_ana_s0(const double2* __restrict__ spec, int64_t ld,
                                                    int64_t smap, int nside, int L,
                                                    const double* __restrict__ alpha,
                                                    const double* __restrict__ smant,
                                                    const int* __restrict__ sexp,
                                                    const int* __restrict__ nact,
                                                    double* __restrict__ partial, int64_t tsize, int chunk0) {
    constexpr int NV = 4 * 2 * K;  // 4 degrees x (re, im) x maps
    constexpr int T = 64 * NW;
    constexpr int CH = T * R;
    __shared__ F lds[2 * NV * T];
    // ring chunk chunk0 + blockIdx.x, written to partial slot blockIdx.x
    const int c = chunk0 + blockIdx.x, m = blockIdx.y;
    const int na = nact[m];
    if (c * CH >= na) return;

    double x[R], mu0[R], mu1[R];
    int kx[R];
    cpx<F> gE[R][K], gO[R][K];
    const double sm = smant[m];
    const int se = sexp[m];
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int j = c * CH + t * T + threadIdx.x;
        mu1[t] = 0.0;
        if (j < na) {
            const int r = 2 * nside - 1 - j, rp = 2 * nside - 1 + j;
            const Ring g = ring_info(nside, r);
            x[t] = g.z;
            seed_lane<double>(sm, se, m, m, g.shalf, g.chalf, mu0[t], kx[t]);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const double2 a = spec[k * smap + (int64_t)r * ld + m];
                if (j == 0) {
                    gE[t][k] = cpx<F>{(F)a.x, (F)a.y};
                    gO[t][k] = cpx<F>{0, 0};
                } else {
                    const double2 b = spec[k * smap + (int64_t)rp * ld + m];
                    gE[t][k] = cpx<F>{(F)(a.x + b.x), (F)(a.y + b.y)};
                    gO[t][k] = cpx<F>{(F)(a.x - b.x), (F)(a.y - b.y)};
                }
            }
        } else {
            x[t] = 0;
            mu0[t] = 0;
            kx[t] = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) gE[t][k] = gO[t][k] = cpx<F>{0, 0};
        }
    }
    bool anyinact = false;
#pragma unroll
    for (int t = 0; t < R; ++t) anyinact = anyinact || (kx[t] < 0);
    anyinact = __any(anyinact);

    const double* al = alpha + table_base(m, L) - m;
    double* part = partial + ((int64_t)blockIdx.x * tsize + table_base(m, L) - m) * (2 * K);
    // mu0 holds degree l (even step), mu1 degree l - 1; each step overwrites the older one,
    // so a block of 4 degrees returns the roles to where they started (no register moves).
    double an[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) an[i] = al[m + i + 1];
    int buf = 0;
    for (int l0 = m; l0 < L; l0 += 4, buf ^= 1) {
        double a[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            a[i] = an[i];
            an[i] = al[l0 + 4 + i + 1];  // prefetch the next block's coefficients
        }
        F acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            // l0 + m is even, so degree l0 + i pairs with the even (E) combination for even i
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const F cur = (F)((i & 1) ? mu1[t] : mu0[t]);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const cpx<F> gv = (i & 1) ? gO[t][k] : gE[t][k];
                    acc[i * 2 * K + 2 * k] = fma(cur, gv.x, acc[i * 2 * K + 2 * k]);
                    acc[i * 2 * K + 2 * k + 1] = fma(cur, gv.y, acc[i * 2 * K + 2 * k + 1]);
                }
            }
#pragma unroll
            for (int t = 0; t < R; ++t) {
                if (i & 1) mu0[t] = fma(x[t] * a[i], mu1[t], -mu0[t]);
                else mu1[t] = fma(x[t] * a[i], mu0[t], -mu1[t]);
            }
        }
        rescale<R, double>(mu0, mu1, kx, anyinact);
        block_reduce<NV, NW, F>(acc, lds, buf, part + (int64_t)l0 * (2 * K));
    }
}

// ------------------------------------------------------------------ spin-0 synthesis
template <int K, int R, int NW, typename F>
__global__ __launch_bounds__(64 * NW) void k_syn_s0(const double2* __restrict__ coef, int nside, int L,
                                                    const double* __restrict__ alpha,
                                                    const double* __restrict__ smant,
                                                    const int* __restrict__ sexp,
                                                    const int* __restrict__ nact,
                                                    double2* __restrict__ spec, int64_t ld, int64_t smap) {
    constexpr int T = 64 * NW;
    constexpr int CH = T * R;
    const int c = blockIdx.x, m = blockIdx.y;
    const int na = nact[m];
    const int NF = 2 * nside;

    double x[R], mu0[R], mu1[R];
    int kx[R];
    cpx<F> aE[R][K], aO[R][K];
    const double sm = smant[m];
    const int se = sexp[m];
    bool anyact = false;
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int j = c * CH + t * T + threadIdx.x;
        mu1[t] = 0;
        kx[t] = 0;
        x[t] = 0;
        mu0[t] = 0;
        if (j < na) {
            const Ring g = ring_info(nside, 2 * nside - 1 - j);
            x[t] = g.z;
            seed_lane<double>(sm, se, m, m, g.shalf, g.chalf, mu0[t], kx[t]);
            anyact = true;
        }
#pragma unroll
        for (int k = 0; k < K; ++k) aE[t][k] = aO[t][k] = cpx<F>{0, 0};
    }
    if (__any(anyact)) {
        bool anyinact = false;
#pragma unroll
        for (int t = 0; t < R; ++t) anyinact = anyinact || (kx[t] < 0);
        anyinact = __any(anyinact);
        const double* al = alpha + table_base(m, L) - m;
        const double2* cf = coef + (table_base(m, L) - m) * K;
        double an[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) an[i] = al[m + i + 1];
        // one or two maps: the next block's coefficients are prefetched; for batches the
        // 4K uniform values would not fit the scalar registers twice, so they load in place
        constexpr int KP = K <= 2 ? K : 1;
        cpx<F> cn[4][KP];
        if constexpr (K <= 2) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int k = 0; k < K; ++k) cn[i][k] = load_c<F>(cf + (int64_t)(m + i) * K + k);
        }
        for (int l0 = m; l0 < L; l0 += 4) {
            double a[4];
            cpx<F> cb[4][K];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a[i] = an[i];
                an[i] = al[l0 + 4 + i + 1];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if constexpr (K <= 2) {
                        cb[i][k] = cn[i][k];
                        cn[i][k] = load_c<F>(cf + (int64_t)(l0 + 4 + i) * K + k);  // next block (zero padded)
                    } else {
                        cb[i][k] = load_c<F>(cf + (int64_t)(l0 + i) * K + k);
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                cpx<F> cv[K];
#pragma unroll
                for (int k = 0; k < K; ++k) cv[k] = cb[i][k];
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    const F cur = (F)((i & 1) ? mu1[t] : mu0[t]);
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        cpx<F>& acc = (i & 1) ? aO[t][k] : aE[t][k];
                        acc.x = fma(cur, cv[k].x, acc.x);
                        acc.y = fma(cur, cv[k].y, acc.y);
                    }
                }
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    if (i & 1) mu0[t] = fma(x[t] * a[i], mu1[t], -mu0[t]);
                    else mu1[t] = fma(x[t] * a[i], mu0[t], -mu1[t]);
                }
            }
            rescale<R, double>(mu0, mu1, kx, anyinact);
        }
    }
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int j = c * CH + t * T + threadIdx.x;
        if (j >= NF) continue;
        const int r = 2 * nside - 1 - j, rp = 2 * nside - 1 + j;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            // F_N = E + O, F_S = E - O (zero for rings beyond the polar-skip bound)
            const double ex = aE[t][k].x, ey = aE[t][k].y, ox = aO[t][k].x, oy = aO[t][k].y;
            spec[k * smap + (int64_t)r * ld + m] = make_double2(ex + ox, ey + oy);
            if (j > 0) spec[k * smap + (int64_t)rp * ld + m] = make_double2(ex - ox, ey - oy);
        }
    }
}

// ------------------------------------------------------------------ spin-s kernels
// All rings, ordered |z| ascending: slot 0 = equator, 2u-1 = north ring 2n-1-u, 2u = south
// ring 2n-1+u.  Channels per degree: direct (re, im) against G[r][m], mirror (re, im)
// against G[mirror(r)][-m].
__device__ inline int slot_ring(int nside, int j) {
    if (j == 0) return 2 * nside - 1;
    const int u = (j + 1) / 2;
    return (j & 1) ? 2 * nside - 1 - u : 2 * nside - 1 + u;
}

template <int R, int NW, typename F>
__global__ __launch_bounds__(64 * NW) void k_ana_s(const double2* __restrict__ spec, int64_t ld, int nside,
                                                   int L, int s, const double* __restrict__ alpha,
                                                   const double* __restrict__ gamma,
                                                   const double* __restrict__ smant,
                                                   const int* __restrict__ sexp,
                                                   const int* __restrict__ nact,
                                                   double* __restrict__ partial, int64_t tsize, int chunk0) {
    constexpr int NV = 16;  // 4 degrees x 4 channels
    constexpr int T = 64 * NW;
    constexpr int CH = T * R;
    __shared__ F lds[2 * NV * T];
    // ring chunk chunk0 + blockIdx.x, written to partial slot blockIdx.x
    const int c = chunk0 + blockIdx.x, m = blockIdx.y;
    const int na = nact[m];
    if (c * CH >= na) return;
    const int M = m > s ? m : s;
    const int nring = 4 * nside - 1;

    double x[R], mu0[R], mu1[R];
    int kx[R];
    cpx<F> gD[R], gM[R];
    const double sm = smant[m];
    const int se = sexp[m];
    const int a_exp = m + s, b_exp = m > s ? m - s : s - m;
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int j = c * CH + t * T + threadIdx.x;
        mu1[t] = 0.0;
        if (j < na) {
            const int r = slot_ring(nside, j);
            const Ring g = ring_info(nside, r);
            x[t] = g.z;
            seed_lane<double>(sm, se, a_exp, b_exp, g.shalf, g.chalf, mu0[t], kx[t]);
            gD[t] = load_c<F>(spec + (int64_t)r * ld + (L - 1) + m);
            gM[t] = m ? load_c<F>(spec + (int64_t)(nring - 1 - r) * ld + (L - 1) - m) : cpx<F>{0, 0};
        } else {
            x[t] = 0;
            mu0[t] = 0;
            kx[t] = 0;
            gD[t] = gM[t] = cpx<F>{0, 0};
        }
    }
    bool anyinact = false;
#pragma unroll
    for (int t = 0; t < R; ++t) anyinact = anyinact || (kx[t] < 0);
    anyinact = __any(anyinact);

    const double* al = alpha + table_base(m, L) - m;
    const double* ga = gamma + table_base(m, L) - m;
    double* part = partial + ((int64_t)blockIdx.x * tsize + table_base(m, L) - m) * 4;
    double an[4], gn[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        an[i] = al[M + i + 1];
        gn[i] = ga[M + i + 1];
    }
    int buf = 0;
    for (int l0 = M; l0 < L; l0 += 4, buf ^= 1) {
        double a[4], gg[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            a[i] = an[i];
            gg[i] = gn[i];
            an[i] = al[l0 + 4 + i + 1];
            gn[i] = ga[l0 + 4 + i + 1];
        }
        F acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const F cur = (F)((i & 1) ? mu1[t] : mu0[t]);
                acc[4 * i + 0] = fma(cur, gD[t].x, acc[4 * i + 0]);
                acc[4 * i + 1] = fma(cur, gD[t].y, acc[4 * i + 1]);
                acc[4 * i + 2] = fma(cur, gM[t].x, acc[4 * i + 2]);
                acc[4 * i + 3] = fma(cur, gM[t].y, acc[4 * i + 3]);
            }
#pragma unroll
            for (int t = 0; t < R; ++t) {
                if (i & 1) mu0[t] = fma(fma(x[t], a[i], gg[i]), mu1[t], -mu0[t]);
                else mu1[t] = fma(fma(x[t], a[i], gg[i]), mu0[t], -mu1[t]);
            }
        }
        rescale<R, double>(mu0, mu1, kx, anyinact);
        block_reduce<NV, NW, F>(acc, lds, buf, part + (int64_t)l0 * 4);
    }
}

template <int R, int NW, typename F>
__global__ __launch_bounds__(64 * NW) void k_syn_s(const double2* __restrict__ coef, int nside, int L, int s,
                                                   const double* __restrict__ alpha,
                                                   const double* __restrict__ gamma,
                                                   const double* __restrict__ smant,
                                                   const int* __restrict__ sexp,
                                                   const int* __restrict__ nact,
                                                   double2* __restrict__ spec, int64_t ld) {
    constexpr int T = 64 * NW;
    constexpr int CH = T * R;
    const int c = blockIdx.x, m = blockIdx.y;
    const int na = nact[m];
    const int nring = 4 * nside - 1;
    const int M = m > s ? m : s;

    double x[R], mu0[R], mu1[R];
    int kx[R];
    cpx<F> aD[R], aM[R];
    const double sm = smant[m];
    const int se = sexp[m];
    const int a_exp = m + s, b_exp = m > s ? m - s : s - m;
    bool anyact = false;
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int j = c * CH + t * T + threadIdx.x;
        mu1[t] = 0;
        kx[t] = 0;
        x[t] = 0;
        mu0[t] = 0;
        aD[t] = aM[t] = cpx<F>{0, 0};
        if (j < na && M < L) {
            const Ring g = ring_info(nside, slot_ring(nside, j));
            x[t] = g.z;
            seed_lane<double>(sm, se, a_exp, b_exp, g.shalf, g.chalf, mu0[t], kx[t]);
            anyact = true;
        }
    }
    if (__any(anyact)) {
        bool anyinact = false;
#pragma unroll
        for (int t = 0; t < R; ++t) anyinact = anyinact || (kx[t] < 0);
        anyinact = __any(anyinact);
        const double* al = alpha + table_base(m, L) - m;
        const double* ga = gamma + table_base(m, L) - m;
        const double2* cf = coef + (table_base(m, L) - m) * 2;
        double an[4], gn[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            an[i] = al[M + i + 1];
            gn[i] = ga[M + i + 1];
        }
        // (the coefficients are not prefetched here: 16 more scalar doubles spill SGPRs)
        for (int l0 = M; l0 < L; l0 += 4) {
            double a[4], gg[4];
            cpx<F> cdb[4], cmb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a[i] = an[i];
                gg[i] = gn[i];
                an[i] = al[l0 + 4 + i + 1];
                gn[i] = ga[l0 + 4 + i + 1];
                cdb[i] = load_c<F>(cf + (int64_t)(l0 + i) * 2);
                cmb[i] = load_c<F>(cf + (int64_t)(l0 + i) * 2 + 1);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const cpx<F> cd = cdb[i], cm = cmb[i];
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    const F cur = (F)((i & 1) ? mu1[t] : mu0[t]);
                    aD[t].x = fma(cur, cd.x, aD[t].x);
                    aD[t].y = fma(cur, cd.y, aD[t].y);
                    aM[t].x = fma(cur, cm.x, aM[t].x);
                    aM[t].y = fma(cur, cm.y, aM[t].y);
                }
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    if (i & 1) mu0[t] = fma(fma(x[t], a[i], gg[i]), mu1[t], -mu0[t]);
                    else mu1[t] = fma(fma(x[t], a[i], gg[i]), mu0[t], -mu1[t]);
                }
            }
            rescale<R, double>(mu0, mu1, kx, anyinact);
        }
    }
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int j = c * CH + t * T + threadIdx.x;
        if (j >= nring) continue;
        const int r = slot_ring(nside, j);
        spec[(int64_t)r * ld + (L - 1) + m] = make_double2(aD[t].x, aD[t].y);
        if (m) spec[(int64_t)(nring - 1 - r) * ld + (L - 1) - m] = make_double2(aM[t].x, aM[t].y);
    }
}

// ------------------------------------------------------------------ matrix-core kernels
// Mixed-precision synthesis on the matrix unit: the recurrence stays in FP64 on the vector
// ALU and the contraction runs as FP32 MFMA (v_mfma_f32_4x4x1f32: 16 independent 4x4
// outer products per wave, block b = lanes 4b..4b+3; A row r and B column c come from
// lanes 4b+r and 4b+c, and the result for (r, c) lands in register r of lane 4b+c).
// A = lambda of the lane's own ring at one degree, B = the degree's coefficient of channel
// (lane & 3), so one instruction adds a degree to a 4 ring x 4 channel tile per block:
// 8 cycles where four FP32 FMAs take 16-18.  On MI355X the matrix unit and the vector ALU
// of a SIMD do not run concurrently (measured: 8 MFMA + 16 FP64 FMA take the sum of both
// when grouped, and a third more when interleaved), so each tile runs the vector work for
// all its degrees first and then the matrix work, with scheduling barriers keeping the two
// apart.  The analysis stays on the vector ALU: there the rings are the contraction index,
// and bringing a ring's degrees across the quad costs as many matrix cycles as it saves.
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline f4 mfma4(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0); }

// Spin-0 synthesis on the matrix unit: one MFMA per degree and ring set and channel group.
template <int K, int R, int NW>
__global__ __launch_bounds__(64 * NW) void k_syn_s0_mx(const double2* __restrict__ coef, int nside, int L,
                                                       const double* __restrict__ alpha,
                                                       const double* __restrict__ smant,
                                                       const int* __restrict__ sexp, const int* __restrict__ nact,
                                                       double2* __restrict__ spec, int64_t ld, int64_t smap) {
    constexpr int C = 2 * K;
    constexpr int NG = (C + 3) / 4;
    constexpr int T = 64 * NW;
    constexpr int CH = T * R;
    const int c = blockIdx.x, m = blockIdx.y;
    const int na = nact[m];
    const int NF = 2 * nside;
    const int q4 = threadIdx.x & 3;
    const int qbase = threadIdx.x & ~3;

    double x[R], mu0[R], mu1[R];
    int kx[R];
    f4 acc[2][R][NG];  // [parity E, O]
    const double sm = smant[m];
    const int se = sexp[m];
    bool anyact = false;
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int j = c * CH + t * T + threadIdx.x;
        mu1[t] = 0;
        kx[t] = 0;
        x[t] = 0;
        mu0[t] = 0;
        if (j < na) {
            const Ring g = ring_info(nside, 2 * nside - 1 - j);
            x[t] = g.z;
            seed_lane<double>(sm, se, m, m, g.shalf, g.chalf, mu0[t], kx[t]);
            anyact = true;
        }
#pragma unroll
        for (int q = 0; q < NG; ++q) acc[0][t][q] = acc[1][t][q] = f4{0.f, 0.f, 0.f, 0.f};
    }
    if (__any(anyact)) {
        bool anyinact = false;
#pragma unroll
        for (int t = 0; t < R; ++t) anyinact = anyinact || (kx[t] < 0);
        anyinact = __any(anyinact);
        const double* al = alpha + table_base(m, L) - m;
        // coefficient of channel ch at degree l: cf[l * C + ch] (zero padded past lmax)
        const double* cf = reinterpret_cast<const double*>(coef + (table_base(m, L) - m) * K);
        double an[4], cn[4][NG];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            an[i] = al[m + i + 1];
#pragma unroll
            for (int q = 0; q < NG; ++q) {
                const int ch = 4 * q + q4;
                cn[i][q] = ch < C ? cf[(int64_t)(m + i) * C + ch] : 0.0;
            }
        }
        for (int l0 = m; l0 < L; l0 += 4) {
            double a[4];
            float cb[4][NG];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a[i] = an[i];
                an[i] = al[l0 + 4 + i + 1];
#pragma unroll
                for (int q = 0; q < NG; ++q) {
                    const int ch = 4 * q + q4;
                    cb[i][q] = (float)cn[i][q];
                    cn[i][q] = ch < C ? cf[(int64_t)(l0 + 4 + i) * C + ch] : 0.0;  // next tile
                }
            }
            // vector burst (recurrence) then matrix burst: the two units do not overlap
            // and every switch between them costs issue cycles
            float cur[R][4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
#pragma unroll
                for (int t = 0; t < R; ++t) cur[t][i] = (float)((i & 1) ? mu1[t] : mu0[t]);
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    if (i & 1) mu0[t] = fma(x[t] * a[i], mu1[t], -mu0[t]);
                    else mu1[t] = fma(x[t] * a[i], mu0[t], -mu1[t]);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int t = 0; t < R; ++t)
#pragma unroll
                    for (int q = 0; q < NG; ++q) acc[i & 1][t][q] = mfma4(cur[t][i], cb[i][q], acc[i & 1][t][q]);
            __builtin_amdgcn_sched_barrier(0);
            rescale<R, double>(mu0, mu1, kx, anyinact);
        }
    }
    // acc[p][t][q][v] at lane (b, q4): ring qbase + v of set t, channel 4q + q4
#pragma unroll
    for (int t = 0; t < R; ++t) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = c * CH + t * T + qbase + v;
            if (j >= NF) continue;
            const int r = 2 * nside - 1 - j, rp = 2 * nside - 1 + j;
#pragma unroll
            for (int q = 0; q < NG; ++q) {
                const int ch = 4 * q + q4;
                if (ch >= C) continue;
                const int k = ch >> 1, comp = ch & 1;
                const double ev = acc[0][t][q][v], ov = acc[1][t][q][v];
                reinterpret_cast<double*>(spec + k * smap + (int64_t)r * ld + m)[comp] = ev + ov;
                if (j > 0) reinterpret_cast<double*>(spec + k * smap + (int64_t)rp * ld + m)[comp] = ev - ov;
            }
        }
    }
}

// Spin-s synthesis on the matrix unit.
template <int R, int NW>
__global__ __launch_bounds__(64 * NW) void k_syn_s_mx(const double2* __restrict__ coef, int nside, int L, int s,
                                                      const double* __restrict__ alpha,
                                                      const double* __restrict__ gamma,
                                                      const double* __restrict__ smant,
                                                      const int* __restrict__ sexp, const int* __restrict__ nact,
                                                      double2* __restrict__ spec, int64_t ld) {
    constexpr int T = 64 * NW;
    constexpr int CH = T * R;
    const int c = blockIdx.x, m = blockIdx.y;
    const int na = nact[m];
    const int nring = 4 * nside - 1;
    const int M = m > s ? m : s;
    const int q4 = threadIdx.x & 3;
    const int qbase = threadIdx.x & ~3;

    double x[R], mu0[R], mu1[R];
    int kx[R];
    f4 acc[R];
    const double sm = smant[m];
    const int se = sexp[m];
    const int a_exp = m + s, b_exp = m > s ? m - s : s - m;
    bool anyact = false;
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int j = c * CH + t * T + threadIdx.x;
        mu1[t] = 0;
        kx[t] = 0;
        x[t] = 0;
        mu0[t] = 0;
        acc[t] = f4{0.f, 0.f, 0.f, 0.f};
        if (j < na && M < L) {
            const Ring g = ring_info(nside, slot_ring(nside, j));
            x[t] = g.z;
            seed_lane<double>(sm, se, a_exp, b_exp, g.shalf, g.chalf, mu0[t], kx[t]);
            anyact = true;
        }
    }
    if (__any(anyact)) {
        bool anyinact = false;
#pragma unroll
        for (int t = 0; t < R; ++t) anyinact = anyinact || (kx[t] < 0);
        anyinact = __any(anyinact);
        const double* al = alpha + table_base(m, L) - m;
        const double* ga = gamma + table_base(m, L) - m;
        // channel q4 of degree l: (direct re, im, mirror re, im) at cf[4 l + q4]
        const double* cf = reinterpret_cast<const double*>(coef + (table_base(m, L) - m) * 2);
        double an[4], gn[4], cn[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            an[i] = al[M + i + 1];
            gn[i] = ga[M + i + 1];
            cn[i] = cf[(int64_t)(M + i) * 4 + q4];
        }
        for (int l0 = M; l0 < L; l0 += 4) {
            double a[4], gg[4];
            float cb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a[i] = an[i];
                gg[i] = gn[i];
                cb[i] = (float)cn[i];
                an[i] = al[l0 + 4 + i + 1];
                gn[i] = ga[l0 + 4 + i + 1];
                cn[i] = cf[(int64_t)(l0 + 4 + i) * 4 + q4];
            }
            float cur[R][4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
#pragma unroll
                for (int t = 0; t < R; ++t) cur[t][i] = (float)((i & 1) ? mu1[t] : mu0[t]);
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    if (i & 1) mu0[t] = fma(fma(x[t], a[i], gg[i]), mu1[t], -mu0[t]);
                    else mu1[t] = fma(fma(x[t], a[i], gg[i]), mu0[t], -mu1[t]);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int t = 0; t < R; ++t) acc[t] = mfma4(cur[t][i], cb[i], acc[t]);
            __builtin_amdgcn_sched_barrier(0);
            rescale<R, double>(mu0, mu1, kx, anyinact);
        }
    }
    // acc[t][v] at lane (b, q4): ring qbase + v of set t, channel q4
#pragma unroll
    for (int t = 0; t < R; ++t) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = c * CH + t * T + qbase + v;
            if (j >= nring) continue;
            const int r = slot_ring(nside, j);
            if (q4 < 2) reinterpret_cast<double*>(spec + (int64_t)r * ld + (L - 1) + m)[q4] = acc[t][v];
            else if (m) reinterpret_cast<double*>(spec + (int64_t)(nring - 1 - r) * ld + (L - 1) - m)[q4 - 2] = acc[t][v];
        }
    }
}

// ------------------------------------------------------------------ tables
__global__ void k_lat_tables(int L, int s, double* __restrict__ alpha, double* __restrict__ gamma,
                             double* __restrict__ P) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L) return;
    const int M = m > s ? m : s;
    const int64_t base = table_base(m, L) - m;  // entry l at base + l
    const int lend = L + GM_TABLE_PAD;
    for (int l = m; l <= M && l < lend; ++l) {
        alpha[base + l] = 0.0;
        if (gamma) gamma[base + l] = 0.0;
        P[base + l] = 1.0;
    }
    double Pm2 = 1.0, Pm1 = 1.0;  // P_{l-2}, P_{l-1}
    const double mm = (double)m * m, ss = (double)s * s;
    for (int l = M + 1; l < L; ++l) {
        const int lp = l - 1;
        const double lp1 = lp + 1.0;
        const double K = lp1 * (2.0 * lp + 1.0) / sqrt((lp1 * lp1 - mm) * (lp1 * lp1 - ss));
        const double A = sqrt((2.0 * lp + 3.0) / (2.0 * lp + 1.0)) * K;
        const double C = lp > 0 ? A * (double)m * (double)s / ((double)lp * lp1) : 0.0;
        double B = 0.0;
        if (lp > 0)
            B = sqrt((2.0 * lp + 3.0) / (2.0 * lp - 1.0)) * K *
                sqrt(((double)lp * lp - mm) * ((double)lp * lp - ss)) / ((double)lp * (2.0 * lp + 1.0));
        const double Pl = (l >= M + 2) ? B * Pm2 : 1.0;
        alpha[base + l] = A * Pm1 / Pl;
        if (gamma) gamma[base + l] = C * Pm1 / Pl;
        P[base + l] = Pl;
        Pm2 = Pm1;
        Pm1 = Pl;
    }
    for (int l = (M + 1 > L ? M + 1 : L); l < lend; ++l) {
        alpha[base + l] = 0.0;
        if (gamma) gamma[base + l] = 0.0;
        P[base + l] = 0.0;
    }
}

// ------------------------------------------------------------------ finish / prep
// Chunk layout of the analysis kernels: CH rings per workgroup.
// nchunks ring chunks of CH rings; an analysis pass covers nslots of them (its partial
// buffer holds nslots chunk slots).  Up to Nside 4096 one pass takes every chunk; above
// that the passes bound the partial buffer (4 slots of ~9.7 GB at Nside 8192).
constexpr int kOnePassNside = 4096, kSlotsPerPass = 4;
struct LatConfig {
    int R, NW, CH, nchunks, nslots;
};

// mode 0 stores, 1 subtracts, 2 adds; the partial slots hold ring chunks chunk0 .. chunk0 +
// nchunk - 1, of which those below ceil(nact / CH) were written
__device__ inline int used_slots(int na, int CH, int chunk0, int nchunk) {
    int used = (na + CH - 1) / CH - chunk0;
    if (used < 0) used = 0;
    return used < nchunk ? used : nchunk;
}

__global__ void k_finish_s0(const double2* __restrict__ partial, int K, int nchunk, int CH,
                            const int* __restrict__ nact, const double* __restrict__ P, int L, int64_t tsize,
                            double2* __restrict__ alm, int64_t astride, int mode, int chunk0) {
    const int m = blockIdx.y;
    const int nu = used_slots(nact[m], CH, chunk0, nchunk);
    const int64_t pb = table_base(m, L) - m;
    const int lmax = L - 1;
    for (int l = m + blockIdx.x * blockDim.x + threadIdx.x; l < L; l += gridDim.x * blockDim.x) {
        const double p = P[pb + l];
        for (int k = 0; k < K; ++k) {
            double2 s = make_double2(0.0, 0.0);
            for (int c = 0; c < nu; ++c) s = cadd(s, partial[((int64_t)c * tsize + pb + l) * K + k]);
            s = cscale(s, p);
            double2& a = alm[k * astride + alm_index(l, m, lmax)];
            a = mode == 1 ? csub(a, s) : mode == 2 ? cadd(a, s) : s;
        }
    }
}

__global__ void k_prep_s0(const double2* __restrict__ alm, int K, int64_t astride, const double* __restrict__ P,
                          int L, double2* __restrict__ coef) {
    const int m = blockIdx.y;
    const int64_t pb = table_base(m, L) - m;
    const int lmax = L - 1;
    for (int l = m + blockIdx.x * blockDim.x + threadIdx.x; l < L + GM_TABLE_PAD; l += gridDim.x * blockDim.x) {
        for (int k = 0; k < K; ++k) {
            double2 v = make_double2(0.0, 0.0);
            if (l < L) v = cscale(alm[k * astride + alm_index(l, m, lmax)], P[pb + l]);
            coef[(pb + l) * K + k] = v;
        }
    }
}

__global__ void k_finish_s(const double2* __restrict__ partial, int nchunk, int CH, const int* __restrict__ nact,
                           const double* __restrict__ P, int L, int s, int64_t tsize, double2* __restrict__ ae,
                           double2* __restrict__ ab, int mode, int chunk0) {
    const int m = blockIdx.y;
    const int nu = used_slots(nact[m], CH, chunk0, nchunk);
    const int64_t pb = table_base(m, L) - m;
    const int lmax = L - 1;
    const double sgn_s = (s & 1) ? -1.0 : 1.0;
    const double sgn_m = (m & 1) ? -1.0 : 1.0;
    for (int l = m + blockIdx.x * blockDim.x + threadIdx.x; l < L; l += gridDim.x * blockDim.x) {
        double2 E = make_double2(0.0, 0.0), B = E;
        if (l >= s) {
            double2 d = make_double2(0.0, 0.0), mi = d;
            for (int c = 0; c < nu; ++c) {
                const double2* p = partial + ((int64_t)c * tsize + pb + l) * 2;
                d = cadd(d, p[0]);
                mi = cadd(mi, p[1]);
            }
            const double pl = P[pb + l];
            const double2 plus = cscale(d, sgn_s * pl);  // a+_{l,m}
            double2 am;                                  // a+_{l,-m}
            if (m) am = cscale(mi, sgn_s * pl * (((l + s) & 1) ? -1.0 : 1.0));
            else am = plus;
            const double2 minus = cscale(cconj(am), sgn_m);  // (-1)^m conj a+_{l,-m}
            E = make_double2(-0.5 * (plus.x + minus.x), -0.5 * (plus.y + minus.y));
            // B = (i/2)(plus - minus)
            B = make_double2(-0.5 * (plus.y - minus.y), 0.5 * (plus.x - minus.x));
        }
        const int64_t ia = alm_index(l, m, lmax);
        if (mode == 1) {
            ae[ia] = csub(ae[ia], E);
            ab[ia] = csub(ab[ia], B);
        } else if (mode == 2) {
            ae[ia] = cadd(ae[ia], E);
            ab[ia] = cadd(ab[ia], B);
        } else {
            ae[ia] = E;
            ab[ia] = B;
        }
    }
}

__global__ void k_prep_s(const double2* __restrict__ ae, const double2* __restrict__ ab,
                         const double* __restrict__ P, int L, int s, double2* __restrict__ coef) {
    const int m = blockIdx.y;
    const int64_t pb = table_base(m, L) - m;
    const int lmax = L - 1;
    const double sgn_s = (s & 1) ? -1.0 : 1.0;
    const double sgn_m = (m & 1) ? -1.0 : 1.0;
    for (int l = m + blockIdx.x * blockDim.x + threadIdx.x; l < L + GM_TABLE_PAD; l += gridDim.x * blockDim.x) {
        double2 cd = make_double2(0.0, 0.0), cm = cd;
        if (l < L && l >= s) {
            const int64_t ia = alm_index(l, m, lmax);
            const double2 e = ae[ia], b = ab[ia];
            // a+_{l,m} = -(E + iB); a+_{l,-m} = -(-1)^m (conj E + i conj B)
            const double2 plus = make_double2(-(e.x - b.y), -(e.y + b.x));
            const double2 minus = make_double2(-sgn_m * (e.x + b.y), -sgn_m * (-e.y + b.x));
            const double pl = P[pb + l] * sgn_s;
            cd = cscale(plus, pl);
            cm = cscale(minus, pl * (((l + s) & 1) ? -1.0 : 1.0));
        }
        coef[(pb + l) * 2] = cd;
        coef[(pb + l) * 2 + 1] = cm;
    }
}

// ------------------------------------------------------------------ configs
// Large geometries use 4-wave workgroups with R rings per lane, per precision and pass
// (the mixed kernels hold half the registers per ring; the mixed syntheses keep their
// accumulators in MFMA tiles).  Chosen by timing R = 3..10 on MI355X at Nside 4096:
// occupancy steps make the best R differ between passes.
#ifndef GM_ANA_R_S0
#define GM_ANA_R_S0 6
#endif
#ifndef GM_SYN_R_S0
#define GM_SYN_R_S0 4
#endif
#ifndef GM_ANA_R_S
#define GM_ANA_R_S 8
#endif
#ifndef GM_SYN_R_S
#define GM_SYN_R_S 6
#endif
#ifndef GM_MIX_ANA_R_S0
#define GM_MIX_ANA_R_S0 5
#endif
#ifndef GM_MIX_SYN_R_S0
#define GM_MIX_SYN_R_S0 5
#endif
#ifndef GM_MIX_ANA_R_S
#define GM_MIX_ANA_R_S 6
#endif
#ifndef GM_MIX_SYN_R_S
#define GM_MIX_SYN_R_S 6
#endif

// Batched spin-0 maps (K = 4 or 8 sharing one recurrence, mixed precision only) hold 4K
// FP32 coefficients per ring, so they take 3 (K = 4) or 2 (K = 8) rings per lane.
constexpr int large_r(bool syn, bool mixed, bool spin0, int K = 1) {
    return (spin0 && K >= 8) ? 2
           : (spin0 && K >= 4) ? 3
           : syn ? (mixed ? (spin0 ? GM_MIX_SYN_R_S0 : GM_MIX_SYN_R_S) : (spin0 ? GM_SYN_R_S0 : GM_SYN_R_S))
                 : (mixed ? (spin0 ? GM_MIX_ANA_R_S0 : GM_MIX_ANA_R_S) : (spin0 ? GM_ANA_R_S0 : GM_ANA_R_S));
}

// Ring counts up to GM_MID_RINGS use 2-wave workgroups of 4 rings per lane: more, shorter
// workgroups for the smaller geometries.  The spin-0 analyses and the spin-s syntheses keep
// them up to kMidRingsWide rings (measured per pass at Nside 512 and 1024: 25-35% faster
// there, while the spin-s analyses and spin-0 syntheses are slower with them above 512).
#ifndef GM_MID_RINGS
#define GM_MID_RINGS 512
#endif
// (the spin-s syntheses measured slower with them at 4095 rings, Nside 1024: 12.1 vs 10.2 ms)
#ifndef GM_MID_RINGS_SYN_S
#define GM_MID_RINGS_SYN_S 2048
#endif
constexpr int64_t kMidRingsWide = 2048;

LatConfig make_config(int64_t nside, int64_t spin, bool mixed, bool syn, int K = 1) {
    const int64_t nrings = spin == 0 ? 2 * nside : 4 * nside - 1;
    LatConfig cfg;
    // the launchers select the instantiation by NW: NW 1 -> R 2, NW 2 -> R 4, NW 4 -> the
    // large R; CH = 64 NW R must match the kernel's own chunk size
    const bool wide = syn ? spin != 0 : spin == 0;
    const int64_t wide_rings = (syn && spin != 0) ? GM_MID_RINGS_SYN_S : kMidRingsWide;
    const int64_t mid = wide && wide_rings > GM_MID_RINGS ? wide_rings : GM_MID_RINGS;
    if (nrings <= 128) cfg = {2, 1, 128, 0, 0};
    else if (nrings <= mid) cfg = {4, 2, 512, 0, 0};
    else {
        const int r = large_r(syn, mixed, spin == 0, K);
        cfg = {r, 4, 256 * r, 0, 0};
    }
    cfg.nchunks = (int)((nrings + cfg.CH - 1) / cfg.CH);
    cfg.nslots = (nside <= kOnePassNside || cfg.nchunks <= kSlotsPerPass) ? cfg.nchunks : kSlotsPerPass;
    return cfg;
}

LatConfig lat_config(int64_t nside, int64_t spin, bool mixed = false, int K = 1) {
    return make_config(nside, spin, mixed, false, K);
}

LatConfig syn_config(int64_t nside, int64_t spin, bool mixed = false, int K = 1) {
    return make_config(nside, spin, mixed, true, K);
}

// The launchers pair NW 1 with R 2, NW 2 with R 4 and NW 4 with the large R; a config that
// does not match its instantiation would run a kernel on the wrong number of threads.
int check_config(const char* name, const LatConfig& cfg, int large_r) {
    const int r = cfg.NW == 1 ? 2 : cfg.NW == 2 ? 4 : cfg.NW == 4 ? large_r : -1;
    if (r != cfg.R || cfg.CH != 64 * cfg.NW * cfg.R)
        return fail(GM_E_LAUNCH, "%s: internal launch configuration mismatch (R %d, NW %d, CH %d)", name, cfg.R, cfg.NW,
                    cfg.CH);
    return GM_OK;
}

int check_lat(int64_t nside, int64_t lmax, int64_t spin) {
    if (nside < 1 || (nside & (nside - 1))) return fail(GM_E_ARG, "nside must be a power of 2");
    if (nside > 8192) return fail(GM_E_UNSUPPORTED, "nside > 8192 not supported");
    if (lmax < 0 || lmax > 4 * nside * 6) return fail(GM_E_ARG, "bad lmax %lld", (long long)lmax);
    if (spin < 0 || spin > lmax) return fail(GM_E_ARG, "spin must satisfy 0 <= spin <= lmax");
    return GM_OK;
}

}  // namespace
}  // namespace gm

using namespace gm;

extern "C" int64_t gm_table_size(int64_t lmax) { return table_size(lmax + 1); }

extern "C" int64_t gm_lat_nchunks(int64_t nside, int64_t spin) { return lat_config(nside, spin).nslots; }

extern "C" int64_t gm_lat_nchunks_mixed(int64_t nside, int64_t spin, int64_t nmaps) {
    return lat_config(nside, spin, true, (int)nmaps).nslots;
}

extern "C" int64_t gm_lat_ring_chunks(int64_t nside, int64_t spin, int64_t mixed, int64_t nmaps) {
    return lat_config(nside, spin, mixed != 0, (int)nmaps).nchunks;
}

extern "C" int gm_lat_tables(int64_t lmax, int64_t spin, double* alpha, double* gamma, double* P, void* stream) {
    if (lmax < 0 || spin < 0 || spin > lmax) return fail(GM_E_ARG, "gm_lat_tables: bad lmax / spin");
    if (!alpha || !P || (spin && !gamma)) return fail(GM_E_ARG, "gm_lat_tables: null pointer");
    const int L = (int)lmax + 1;
    hipLaunchKernelGGL(k_lat_tables, dim3((L + 63) / 64), dim3(64), 0, (hipStream_t)stream, L, (int)spin, alpha,
                       gamma, P);
    return check_launch("gm_lat_tables");
}


namespace gm {
namespace {

template <typename F>
int launch_ana_s0(const char* name, const double* spec, int64_t nmaps, int64_t ld, int64_t smap, int64_t nside,
                  int64_t lmax, const double* alpha, const double* seed_mant, const int32_t* seed_exp,
                  const int32_t* nactive, double* partial, int64_t nchunks, int64_t chunk0, void* stream) {
    if (int e = check_lat(nside, lmax, 0)) return e;
    if (!spec || !alpha || !seed_mant || !seed_exp || !nactive || !partial) return fail(GM_E_ARG, "%s: null pointer", name);
    constexpr bool mixed = std::is_same<F, float>::value;
    if (!(nmaps == 1 || nmaps == 2 || (mixed && (nmaps == 4 || nmaps == 8))))
        return fail(GM_E_ARG, mixed ? "%s: nmaps must be 1, 2, 4 or 8" : "%s: nmaps must be 1 or 2", name);
    const int64_t L = lmax + 1;
    if (ld < L) return fail(GM_E_SHAPE, "%s: ld < lmax + 1", name);
    if (nmaps > 1 && smap < (4 * nside - 1) * ld) return fail(GM_E_SHAPE, "%s: map stride too small", name);
    const LatConfig cfg = lat_config(nside, 0, mixed, (int)nmaps);
    if (nchunks != cfg.nslots)
        return fail(GM_E_SHAPE, "%s: partial has %lld chunks, need %d", name, (long long)nchunks, cfg.nslots);
    if (chunk0 < 0 || chunk0 >= cfg.nchunks || chunk0 % cfg.nslots)
        return fail(GM_E_ARG, "%s: chunk0 %lld is not a pass start below %d", name, (long long)chunk0, cfg.nchunks);
    const int64_t ts = table_size(L);
    const int nblk = cfg.nchunks - (int)chunk0 < cfg.nslots ? cfg.nchunks - (int)chunk0 : cfg.nslots;
    dim3 grid(nblk, (unsigned)L), block(64 * cfg.NW);
    hipStream_t st = (hipStream_t)stream;
    const double2* sp = reinterpret_cast<const double2*>(spec);
    constexpr int RL = large_r(false, mixed, true);
    if (int e = check_config(name, cfg, large_r(false, mixed, true, (int)nmaps))) return e;
#define GM_ANA0(KK, RR, WW) hipLaunchKernelGGL((k_ana_s0<KK, RR, WW, F>), grid, block, 0, st, sp, ld, smap, (int)nside, (int)L, alpha, seed_mant, seed_exp, nactive, partial, ts, (int)chunk0)
    if (nmaps == 1) {
        if (cfg.NW == 1) GM_ANA0(1, 2, 1);
        else if (cfg.NW == 2) GM_ANA0(1, 4, 2);
        else GM_ANA0(1, RL, 4);
    } else if (nmaps == 2) {
        if (cfg.NW == 1) GM_ANA0(2, 2, 1);
        else if (cfg.NW == 2) GM_ANA0(2, 4, 2);
        else GM_ANA0(2, RL, 4);
    } else if constexpr (mixed) {
        if (nmaps == 4) {
            if (cfg.NW == 1) GM_ANA0(4, 2, 1);
            else if (cfg.NW == 2) GM_ANA0(4, 4, 2);
            else GM_ANA0(4, 3, 4);
        } else {
            if (cfg.NW == 1) GM_ANA0(8, 2, 1);
            else if (cfg.NW == 2) GM_ANA0(8, 4, 2);
            else GM_ANA0(8, 2, 4);
        }
    }
#undef GM_ANA0
    return check_launch(name);
}

// A single mixed spin-0 map (K = 1) synthesises on the matrix unit when this is 1, else on
// the vector ALU in FP32.
#ifndef GM_MIX_SYN0_SINGLE_MFMA
#define GM_MIX_SYN0_SINGLE_MFMA 1
#endif

template <typename F>
int launch_syn_s0(const char* name, const double* coef, int64_t nmaps, int64_t nside, int64_t lmax, const double* alpha,
                  const double* seed_mant, const int32_t* seed_exp, const int32_t* nactive, double* spec, int64_t ld,
                  int64_t smap, void* stream) {
    if (int e = check_lat(nside, lmax, 0)) return e;
    if (!coef || !alpha || !seed_mant || !seed_exp || !nactive || !spec) return fail(GM_E_ARG, "%s: null pointer", name);
    constexpr bool mixed = std::is_same<F, float>::value;
    if (!(nmaps == 1 || nmaps == 2 || (mixed && (nmaps == 4 || nmaps == 8))))
        return fail(GM_E_ARG, mixed ? "%s: nmaps must be 1, 2, 4 or 8" : "%s: nmaps must be 1 or 2", name);
    const int64_t L = lmax + 1;
    if (ld < L) return fail(GM_E_SHAPE, "%s: ld < lmax + 1", name);
    if (nmaps > 1 && smap < (4 * nside - 1) * ld) return fail(GM_E_SHAPE, "%s: map stride too small", name);
    const LatConfig cfg = syn_config(nside, 0, mixed, (int)nmaps);
    dim3 grid(cfg.nchunks, (unsigned)L), block(64 * cfg.NW);
    hipStream_t st = (hipStream_t)stream;
    const double2* cf = reinterpret_cast<const double2*>(coef);
    double2* sp = reinterpret_cast<double2*>(spec);
    constexpr int RL = large_r(true, mixed, true);
    if (int e = check_config(name, cfg, large_r(true, mixed, true, (int)nmaps))) return e;
#define GM_SYN0(KK, RR, WW)                                                                                       \
    do {                                                                                                          \
        if constexpr (mixed && (KK > 1 || GM_MIX_SYN0_SINGLE_MFMA))                                               \
            hipLaunchKernelGGL((k_syn_s0_mx<KK, RR, WW>), grid, block, 0, st, cf, (int)nside, (int)L, alpha,      \
                               seed_mant, seed_exp, nactive, sp, ld, smap);                                        \
        else                                                                                                      \
            hipLaunchKernelGGL((k_syn_s0<KK, RR, WW, F>), grid, block, 0, st, cf, (int)nside, (int)L, alpha,      \
                               seed_mant, seed_exp, nactive, sp, ld, smap);                                        \
    } while (0)
    if (nmaps == 1) {
        if (cfg.NW == 1) GM_SYN0(1, 2, 1);
        else if (cfg.NW == 2) GM_SYN0(1, 4, 2);
        else GM_SYN0(1, RL, 4);
    } else if (nmaps == 2) {
        if (cfg.NW == 1) GM_SYN0(2, 2, 1);
        else if (cfg.NW == 2) GM_SYN0(2, 4, 2);
        else GM_SYN0(2, RL, 4);
    } else if constexpr (mixed) {
        if (nmaps == 4) {
            if (cfg.NW == 1) GM_SYN0(4, 2, 1);
            else if (cfg.NW == 2) GM_SYN0(4, 4, 2);
            else GM_SYN0(4, 3, 4);
        } else {
            if (cfg.NW == 1) GM_SYN0(8, 2, 1);
            else if (cfg.NW == 2) GM_SYN0(8, 4, 2);
            else GM_SYN0(8, 2, 4);
        }
    }
#undef GM_SYN0
    return check_launch(name);
}

template <typename F>
int launch_ana_s(const char* name, const double* spec, int64_t ld, int64_t nside, int64_t lmax, int64_t spin,
                 const double* alpha, const double* gamma, const double* seed_mant, const int32_t* seed_exp,
                 const int32_t* nactive, double* partial, int64_t nchunks, int64_t chunk0, void* stream) {
    if (int e = check_lat(nside, lmax, spin)) return e;
    if (spin == 0) return fail(GM_E_ARG, "%s: use the spin-0 entry for spin 0", name);
    if (!spec || !alpha || !gamma || !seed_mant || !seed_exp || !nactive || !partial)
        return fail(GM_E_ARG, "%s: null pointer", name);
    const int64_t L = lmax + 1;
    if (ld < 2 * L - 1) return fail(GM_E_SHAPE, "%s: ld < 2 lmax + 1", name);
    const LatConfig cfg = lat_config(nside, spin, std::is_same<F, float>::value);
    if (nchunks != cfg.nslots)
        return fail(GM_E_SHAPE, "%s: partial has %lld chunks, need %d", name, (long long)nchunks, cfg.nslots);
    if (chunk0 < 0 || chunk0 >= cfg.nchunks || chunk0 % cfg.nslots)
        return fail(GM_E_ARG, "%s: chunk0 %lld is not a pass start below %d", name, (long long)chunk0, cfg.nchunks);
    const int nblk = cfg.nchunks - (int)chunk0 < cfg.nslots ? cfg.nchunks - (int)chunk0 : cfg.nslots;
    dim3 grid(nblk, (unsigned)L), block(64 * cfg.NW);
    hipStream_t st = (hipStream_t)stream;
    const double2* sp = reinterpret_cast<const double2*>(spec);
    const int64_t ts = table_size(L);
    constexpr int RL = large_r(false, std::is_same<F, float>::value, false);
    if (int e = check_config(name, cfg, RL)) return e;
#define GM_ANAS(RR, WW) hipLaunchKernelGGL((k_ana_s<RR, WW, F>), grid, block, 0, st, sp, ld, (int)nside, (int)L, (int)spin, alpha, gamma, seed_mant, seed_exp, nactive, partial, ts, (int)chunk0)
    if (cfg.NW == 1) GM_ANAS(2, 1);
    else if (cfg.NW == 2) GM_ANAS(4, 2);
    else GM_ANAS(RL, 4);
#undef GM_ANAS
    return check_launch(name);
}

template <typename F>
int launch_syn_s(const char* name, const double* coef, int64_t nside, int64_t lmax, int64_t spin, const double* alpha,
                 const double* gamma, const double* seed_mant, const int32_t* seed_exp, const int32_t* nactive,
                 double* spec, int64_t ld, void* stream) {
    if (int e = check_lat(nside, lmax, spin)) return e;
    if (spin == 0) return fail(GM_E_ARG, "%s: use the spin-0 entry for spin 0", name);
    if (!coef || !alpha || !gamma || !seed_mant || !seed_exp || !nactive || !spec)
        return fail(GM_E_ARG, "%s: null pointer", name);
    const int64_t L = lmax + 1;
    if (ld < 2 * L - 1) return fail(GM_E_SHAPE, "%s: ld < 2 lmax + 1", name);
    const LatConfig cfg = syn_config(nside, spin, std::is_same<F, float>::value);
    dim3 grid(cfg.nchunks, (unsigned)L), block(64 * cfg.NW);
    hipStream_t st = (hipStream_t)stream;
    const double2* cf = reinterpret_cast<const double2*>(coef);
    double2* sp = reinterpret_cast<double2*>(spec);
    constexpr int RL = large_r(true, std::is_same<F, float>::value, false);
    if (int e = check_config(name, cfg, RL)) return e;
#define GM_SYNS(RR, WW)                                                                                           \
    do {                                                                                                          \
        if constexpr (std::is_same<F, float>::value)                                                             \
            hipLaunchKernelGGL((k_syn_s_mx<RR, WW>), grid, block, 0, st, cf, (int)nside, (int)L, (int)spin,       \
                               alpha, gamma, seed_mant, seed_exp, nactive, sp, ld);                                \
        else                                                                                                      \
            hipLaunchKernelGGL((k_syn_s<RR, WW, F>), grid, block, 0, st, cf, (int)nside, (int)L, (int)spin,       \
                               alpha, gamma, seed_mant, seed_exp, nactive, sp, ld);                                \
    } while (0)
    if (cfg.NW == 1) GM_SYNS(2, 1);
    else if (cfg.NW == 2) GM_SYNS(4, 2);
    else GM_SYNS(RL, 4);
#undef GM_SYNS
    return check_launch(name);
}

}  // namespace
}  // namespace gm

extern "C" int gm_lat_ana_s0(const double* spec, int64_t nmaps, int64_t ld, int64_t smap, int64_t nside,
                             int64_t lmax, const double* alpha, const double* seed_mant, const int32_t* seed_exp,
                             const int32_t* nactive, double* partial, int64_t nchunks, int64_t chunk0,
                             void* stream) {
    return launch_ana_s0<double>("gm_lat_ana_s0", spec, nmaps, ld, smap, nside, lmax, alpha, seed_mant, seed_exp,
                                 nactive, partial, nchunks, chunk0, stream);
}

extern "C" int gm_lat_ana_s0_mixed(const double* spec, int64_t nmaps, int64_t ld, int64_t smap, int64_t nside,
                                   int64_t lmax, const double* alpha, const double* seed_mant, const int32_t* seed_exp,
                                   const int32_t* nactive, double* partial, int64_t nchunks, int64_t chunk0,
                                   void* stream) {
    return launch_ana_s0<float>("gm_lat_ana_s0_mixed", spec, nmaps, ld, smap, nside, lmax, alpha, seed_mant, seed_exp,
                                nactive, partial, nchunks, chunk0, stream);
}

namespace gm {
namespace {
int finish_s0(const char* name, bool mixed, const double* partial, int64_t nmaps, int64_t nside,
              const int32_t* nactive, const double* P, int64_t lmax, double* alm, int64_t alm_map_stride, int mode,
              int64_t chunk0, void* stream) {
    if (mode < 0 || mode > 2 || chunk0 < 0) return fail(GM_E_ARG, "%s: bad mode / chunk0", name);
    if (int e = check_lat(nside, lmax, 0)) return e;
    if (!partial || !nactive || !P || !alm) return fail(GM_E_ARG, "%s: null pointer", name);
    if (!(nmaps == 1 || nmaps == 2 || (mixed && (nmaps == 4 || nmaps == 8))))
        return fail(GM_E_ARG, mixed ? "%s: nmaps must be 1, 2, 4 or 8" : "%s: nmaps must be 1 or 2", name);
    const int64_t L = lmax + 1, nelem = L * (L + 1) / 2;
    if (nmaps > 1 && alm_map_stride < nelem) return fail(GM_E_SHAPE, "%s: alm stride too small", name);
    const LatConfig cfg = lat_config(nside, 0, mixed, (int)nmaps);
    hipLaunchKernelGGL(k_finish_s0, dim3((unsigned)((L + 255) / 256), (unsigned)L), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<const double2*>(partial), (int)nmaps, cfg.nslots, cfg.CH, nactive, P, (int)L,
                       table_size(L), reinterpret_cast<double2*>(alm), alm_map_stride, mode, (int)chunk0);
    return check_launch(name);
}
}  // namespace
}  // namespace gm

extern "C" int gm_lat_finish_s0(const double* partial, int64_t nmaps, int64_t nside, const int32_t* nactive,
                                const double* P, int64_t lmax, double* alm, int64_t alm_map_stride, int mode,
                                int64_t chunk0, void* stream) {
    return finish_s0("gm_lat_finish_s0", false, partial, nmaps, nside, nactive, P, lmax, alm, alm_map_stride, mode,
                     chunk0, stream);
}

extern "C" int gm_lat_finish_s0_mixed(const double* partial, int64_t nmaps, int64_t nside, const int32_t* nactive,
                                      const double* P, int64_t lmax, double* alm, int64_t alm_map_stride, int mode,
                                      int64_t chunk0, void* stream) {
    return finish_s0("gm_lat_finish_s0_mixed", true, partial, nmaps, nside, nactive, P, lmax, alm, alm_map_stride,
                     mode, chunk0, stream);
}

extern "C" int gm_lat_prep_s0(const double* alm, int64_t nmaps, int64_t alm_map_stride, const double* P,
                              int64_t lmax, double* coef, void* stream) {
    if (lmax < 0 || !alm || !P || !coef) return fail(GM_E_ARG, "gm_lat_prep_s0: bad arguments");
    if (nmaps != 1 && nmaps != 2 && nmaps != 4 && nmaps != 8)
        return fail(GM_E_ARG, "gm_lat_prep_s0: nmaps must be 1, 2, 4 or 8");
    const int64_t L = lmax + 1;
    if (nmaps > 1 && alm_map_stride < L * (L + 1) / 2) return fail(GM_E_SHAPE, "gm_lat_prep_s0: alm stride too small");
    hipLaunchKernelGGL(k_prep_s0, dim3((unsigned)((L + GM_TABLE_PAD + 255) / 256), (unsigned)L), dim3(256), 0,
                       (hipStream_t)stream, reinterpret_cast<const double2*>(alm), (int)nmaps, alm_map_stride, P,
                       (int)L, reinterpret_cast<double2*>(coef));
    return check_launch("gm_lat_prep_s0");
}

extern "C" int gm_lat_syn_s0(const double* coef, int64_t nmaps, int64_t nside, int64_t lmax, const double* alpha,
                             const double* seed_mant, const int32_t* seed_exp, const int32_t* nactive, double* spec,
                             int64_t ld, int64_t smap, void* stream) {
    return launch_syn_s0<double>("gm_lat_syn_s0", coef, nmaps, nside, lmax, alpha, seed_mant, seed_exp, nactive, spec,
                                 ld, smap, stream);
}

extern "C" int gm_lat_syn_s0_mixed(const double* coef, int64_t nmaps, int64_t nside, int64_t lmax, const double* alpha,
                                   const double* seed_mant, const int32_t* seed_exp, const int32_t* nactive,
                                   double* spec, int64_t ld, int64_t smap, void* stream) {
    return launch_syn_s0<float>("gm_lat_syn_s0_mixed", coef, nmaps, nside, lmax, alpha, seed_mant, seed_exp, nactive,
                                spec, ld, smap, stream);
}

extern "C" int gm_lat_ana_s(const double* spec, int64_t ld, int64_t nside, int64_t lmax, int64_t spin,
                            const double* alpha, const double* gamma, const double* seed_mant, const int32_t* seed_exp,
                            const int32_t* nactive, double* partial, int64_t nchunks, int64_t chunk0,
                            void* stream) {
    return launch_ana_s<double>("gm_lat_ana_s", spec, ld, nside, lmax, spin, alpha, gamma, seed_mant, seed_exp,
                                nactive, partial, nchunks, chunk0, stream);
}

extern "C" int gm_lat_ana_s_mixed(const double* spec, int64_t ld, int64_t nside, int64_t lmax, int64_t spin,
                                  const double* alpha, const double* gamma, const double* seed_mant,
                                  const int32_t* seed_exp, const int32_t* nactive, double* partial, int64_t nchunks,
                                  int64_t chunk0, void* stream) {
    return launch_ana_s<float>("gm_lat_ana_s_mixed", spec, ld, nside, lmax, spin, alpha, gamma, seed_mant, seed_exp,
                               nactive, partial, nchunks, chunk0, stream);
}

namespace gm {
namespace {
int finish_s(const char* name, bool mixed, const double* partial, int64_t nside, const int32_t* nactive,
             const double* P, int64_t lmax, int64_t spin, double* alm_e, double* alm_b, int mode, int64_t chunk0,
             void* stream) {
    if (mode < 0 || mode > 2 || chunk0 < 0) return fail(GM_E_ARG, "%s: bad mode / chunk0", name);
    if (int e = check_lat(nside, lmax, spin)) return e;
    if (!partial || !nactive || !P || !alm_e || !alm_b) return fail(GM_E_ARG, "%s: null pointer", name);
    const int64_t L = lmax + 1;
    const LatConfig cfg = lat_config(nside, spin, mixed);
    hipLaunchKernelGGL(k_finish_s, dim3((unsigned)((L + 255) / 256), (unsigned)L), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<const double2*>(partial), cfg.nslots, cfg.CH, nactive, P, (int)L, (int)spin,
                       table_size(L), reinterpret_cast<double2*>(alm_e), reinterpret_cast<double2*>(alm_b), mode,
                       (int)chunk0);
    return check_launch(name);
}
}  // namespace
}  // namespace gm

extern "C" int gm_lat_finish_s(const double* partial, int64_t nside, const int32_t* nactive, const double* P,
                               int64_t lmax, int64_t spin, double* alm_e, double* alm_b, int mode, int64_t chunk0,
                               void* stream) {
    return finish_s("gm_lat_finish_s", false, partial, nside, nactive, P, lmax, spin, alm_e, alm_b, mode, chunk0,
                    stream);
}

extern "C" int gm_lat_finish_s_mixed(const double* partial, int64_t nside, const int32_t* nactive, const double* P,
                                     int64_t lmax, int64_t spin, double* alm_e, double* alm_b, int mode,
                                     int64_t chunk0, void* stream) {
    return finish_s("gm_lat_finish_s_mixed", true, partial, nside, nactive, P, lmax, spin, alm_e, alm_b, mode, chunk0,
                    stream);
}

extern "C" int gm_lat_prep_s(const double* alm_e, const double* alm_b, const double* P, int64_t lmax, int64_t spin,
                             double* coef, void* stream) {
    if (lmax < 0 || spin < 0 || spin > lmax || !alm_e || !alm_b || !P || !coef)
        return fail(GM_E_ARG, "gm_lat_prep_s: bad arguments");
    const int64_t L = lmax + 1;
    hipLaunchKernelGGL(k_prep_s, dim3((unsigned)((L + GM_TABLE_PAD + 255) / 256), (unsigned)L), dim3(256), 0,
                       (hipStream_t)stream, reinterpret_cast<const double2*>(alm_e),
                       reinterpret_cast<const double2*>(alm_b), P, (int)L, (int)spin,
                       reinterpret_cast<double2*>(coef));
    return check_launch("gm_lat_prep_s");
}

extern "C" int gm_lat_syn_s(const double* coef, int64_t nside, int64_t lmax, int64_t spin, const double* alpha,
                            const double* gamma, const double* seed_mant, const int32_t* seed_exp,
                            const int32_t* nactive, double* spec, int64_t ld, void* stream) {
    return launch_syn_s<double>("gm_lat_syn_s", coef, nside, lmax, spin, alpha, gamma, seed_mant, seed_exp, nactive,
                                spec, ld, stream);
}

extern "C" int gm_lat_syn_s_mixed(const double* coef, int64_t nside, int64_t lmax, int64_t spin, const double* alpha,
                                  const double* gamma, const double* seed_mant, const int32_t* seed_exp,
                                  const int32_t* nactive, double* spec, int64_t ld, void* stream) {
    return launch_syn_s<float>("gm_lat_syn_s_mixed", coef, nside, lmax, spin, alpha, gamma, seed_mant, seed_exp,
                               nactive, spec, ld, stream);
}
