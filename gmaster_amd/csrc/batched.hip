// Batched latitudinal pass on gfx950: spin-0 maps and spin-2 fields sharing ONE spin-0
// recurrence lambda_lm per (m, ring), with the contraction on the matrix unit.
//
// The spin-2 rows are combinations of spin-0 rows (Gorski et al. 2005, the HEALPix F1/F2
// functions): with W = (lambda+_lm + (-1)^m lambda-_lm)/2, X = (lambda+_lm - (-1)^m lambda-_lm)/2,
// s^2 = sin^2(theta), and the recurrence x lambda_{l-1} = eps_l lambda_l + eps_{l-1} lambda_{l-2}
// removing every cos(theta) factor,
//     W_l = aW_l lambda_l / s^2 + bW_l lambda_l + cW_l lambda_{l-2} / s^2
//     X_l = dX_l lambda_{l+1} / s^2 + eX_l lambda_{l-1} / s^2.
// An analysis therefore needs, per order m and degree j = m..L, only the sums
//     S0_j(G) = sum_r lambda_j(theta_r) G_r     and     S1_j(G) = sum_r lambda_j(theta_r) G_r / s_r^2
// of the ring spectra G of Q and U (real maps: G_Q, G_U from the +-m columns of Q + iU),
// and a synthesis only sums of lambda_j against per-degree coefficients: its plain part
// folds into the 1/s^2 part through s^2 lambda_j = lambda_j - x^2 lambda_j (a combination of
// lambda_{j-2}, lambda_j, lambda_{j+2}), so Q(theta) = sum_j gamma_j lambda_j / s^2 with one
// coefficient per degree.  T, the mask and Q/U of one realisation are then 12 real analysis
// channels (8 synthesis channels) of one recurrence instead of three recurrences of 4.
//
// Analysis (k_bat_ana): lanes are folded ring pairs (north and south rings of one |z|,
// combined as E = N + S for even l + m and O = N - S for odd).  Every 32 degrees each lane
// writes its FP32-rounded mu values into a wave-private LDS tile, the wave reads them back
// transposed, and v_mfma_f32_16x16x4f32 contracts 4 rings x 16 channels x 16 degrees per
// instruction: A = channel values of 4 rings (constant over l, kept in registers), B = mu
// of those rings at 16 degrees of one parity.  The ring sum is the MFMA's K dimension, so
// no per-degree cross-lane reduction is needed; the waves of a workgroup combine their
// 16 x 32 tiles through LDS once per tile and write one partial per ring block.
//
// Polar rings: near the poles 1/s^2 is large and, for small m, the 1/s^2 terms of W cancel
// to O(s^2) of their size, so FP32 rounding there would be amplified by up to ~nside^2.
// Ring pairs with s^2 < 1/16 (within ~14.5 degrees of a pole) therefore run the same
// kernel in FP64 (v_mfma_f64_16x16x4f64, whose accumulator rows are interleaved
// differently from the f32 instruction's); elsewhere 1/s^2 <= 16 and FP32 keeps ~1e-7.
//
// Synthesis (k_bat_syn): as the spin-0 mixed synthesis, one v_mfma_f32_4x4x1f32 per
// (degree, ring set, 4 channels), lanes = ring pairs, E/O accumulators; the epilogue
// divides the spin-2 sums by s^2 and writes F+(+-m) = Q + iU, conj(Q) + i conj(U).  Polar
// ring pairs use an FP64 vector-ALU variant (k_bat_syn64).  Tables run to degree lmax + 2
// (the folded plain part reaches one degree past lmax + 1).
#include "common.h"
#include "recurrence.h"

namespace gm {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

template <typename F> struct Mfma16;
// Accumulator layouts (checked on the device): lane l, register v holds row
// 4 (l / 16) + v (f32) or 4 v + l / 16 (f64) of column l % 16.
template <> struct Mfma16<float> {
    typedef f4 acc_t;
    __device__ static constexpr int row(int group, int v) { return 4 * group + v; }
    __device__ static acc_t op(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};
template <> struct Mfma16<double> {
    typedef d4 acc_t;
    __device__ static constexpr int row(int group, int v) { return 4 * v + group; }
    __device__ static acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

// Launch shapes (tuning switches; the defaults were chosen by timing at Nside 1024 and 4096)
#ifndef GM_BAT_ANA_NW
#define GM_BAT_ANA_NW 8   // waves per FP32 analysis workgroup
#endif
#ifndef GM_BAT_ANA_R
#define GM_BAT_ANA_R 1    // ring sets per lane in the FP32 analysis
#endif
#ifndef GM_BAT_SUPER_H
#define GM_BAT_SUPER_H 4  // 32-degree tiles per cross-wave reduction in the FP16 analysis
#endif
#ifndef GM_BAT_SYN_NW
#define GM_BAT_SYN_NW 4
#endif
#ifndef GM_BAT_SYN_R
#define GM_BAT_SYN_R 2
#endif
#ifndef GM_BAT_SYN_WAVES
#define GM_BAT_SYN_WAVES 4  // minimum waves per SIMD the FP32 synthesis is compiled for
#endif

constexpr int kTile = 32;       // degrees per analysis tile (16 of each parity)
constexpr int kSuper = 2;       // analysis tiles per cross-wave reduction
// The same for the FP16 analysis (fewer cross-wave reductions and barriers; 2.5% faster at
// Nside 4096); spin-2-only batches keep 2, since they spill registers at 4.
template <int K0> constexpr int super_h() { return K0 > 0 ? GM_BAT_SUPER_H : 2; }
constexpr int kTabExtra = 2;    // table degrees beyond lmax + 1 (tables of lmax + 2)
constexpr int kAlphaTail = 128;  // zero doubles the caller appends to the alpha table
constexpr int kCoefTail = 3 * kTile;  // spare degrees (x the channel stride) after the coefficients
constexpr int kRow = 68;    // transpose row: 64 rings + 4 padding

// Channel counts.  Analysis: spin-0 map k at slots 2k, 2k+1 (re, im); spin-2 field f at
// 2 K0 + 8 f + (S0Q re, im, S0U re, im, S1Q re, im, S1U re, im).  Synthesis: spin-0 as
// in the analysis; spin-2 field f at syn_base2 + 4 f + (Q re, im, U re, im), aligned to 4
// so that the four components of one field sit in the four lanes of a 4x4x1 block.
template <int K0, int K2> struct Chan {
    static constexpr int nc = 2 * K0 + 8 * K2;          // analysis channels
    static constexpr int ana_stride = (nc + 3) / 4 * 4;  // partial-sum stride
    static constexpr int syn_base2 = (2 * K0 + 3) / 4 * 4;
    static constexpr int syn_stride = syn_base2 + 4 * K2;  // coefficient stride (multiple of 4)
    static_assert(nc <= 16, "at most 16 analysis channels");
};

__device__ inline double eps_lm(int j, int m) {
    if (j < m || j <= 0) return 0.0;
    const double dj = j, dm = m;
    return sqrt((dj * dj - dm * dm) / (4.0 * dj * dj - 1.0));
}

// The five per-(l, m) coefficients of W_l and X_l (zero below max(m, 2)).
__device__ inline void wx_coefs(int l, int m, double& aW, double& bW, double& cW, double& dX, double& eX) {
    aW = bW = cW = dX = eX = 0.0;
    if (l < 2 || l < m) return;
    const double dl = l, dm = m, m2 = dm * dm;
    const double c = 2.0 / sqrt((dl + 2.0) * (dl + 1.0) * dl * (dl - 1.0));
    const double k = sqrt((2.0 * dl + 1.0) * (dl * dl - m2) / (2.0 * dl - 1.0));
    auto eps = [&](int j) -> double {
        if (j < m || j <= 0) return 0.0;
        const double dj = j;
        return sqrt((dj * dj - m2) / (4.0 * dj * dj - 1.0));
    };
    aW = c * (m2 - dl + (dl * dl - m2) / (2.0 * dl - 1.0));
    bW = -c * dl * (dl - 1.0) * 0.5;
    cW = c * k * eps(l - 1);
    dX = -c * dm * (dl - 1.0) * eps(l + 1);
    eX = c * dm * (k - (dl - 1.0) * eps(l));
}

// E (p = 0) and O (p = 1) combinations of analysis channel ch at ring pair j, order m.
template <int K0, int K2>
__device__ inline void chan_eo(const double2* __restrict__ spec0, int64_t ld0, int64_t smap0,
                               const double2* __restrict__ spec2, int64_t ld2, int64_t smap2, int nside, int L,
                               int m, int j, int ch, double& e, double& o) {
    e = o = 0.0;
    const int r = 2 * nside - 1 - j, rp = 2 * nside - 1 + j;
    if (ch < 2 * K0) {
        const int k = ch >> 1, comp = ch & 1;
        const double vn = reinterpret_cast<const double*>(spec0 + k * smap0 + (int64_t)r * ld0 + m)[comp];
        if (j == 0) {
            e = vn;
            return;
        }
        const double vs = reinterpret_cast<const double*>(spec0 + k * smap0 + (int64_t)rp * ld0 + m)[comp];
        e = vn + vs;
        o = vn - vs;
        return;
    }
    const int c2 = ch - 2 * K0;
    if (c2 >= 8 * K2) return;
    const int f = c2 >> 3, sub = c2 & 7;
    const bool is_u = (sub & 2) != 0, s1 = (sub & 4) != 0;
    const int comp = sub & 1;
    // G_Q = (F(m) + conj F(-m)) / 2,  G_U = (F(m) - conj F(-m)) / (2i)
    auto gval = [&](int ring) -> double {
        const double2 fp = spec2[f * smap2 + (int64_t)ring * ld2 + (L - 1) + m];
        const double2 fm = spec2[f * smap2 + (int64_t)ring * ld2 + (L - 1) - m];
        if (!is_u) return comp ? 0.5 * (fp.y - fm.y) : 0.5 * (fp.x + fm.x);
        return comp ? 0.5 * (fm.x - fp.x) : 0.5 * (fp.y + fm.y);
    };
    double sc = 1.0;
    if (s1) {
        const Ring g = ring_info(nside, r);
        sc = 1.0 / (g.sin * g.sin);
    }
    const double vn = gval(r);
    if (j == 0) {
        e = vn * sc;
        return;
    }
    const double vs = gval(rp);
    e = (vn + vs) * sc;
    o = (vn - vs) * sc;
}

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename F> struct Vec4;
template <> struct Vec4<float> { typedef f4 type; };
template <> struct Vec4<double> { typedef d4 type; };

// ------------------------------------------------------------------ analysis
// Workgroup (blockIdx.x = ring block, blockIdx.y = m) over ring pairs
// [jbeg + blockIdx.x * CH, +CH) ∩ [jbeg, min(nact[m], jend)); degrees m..L (L = lmax + 1:
// S1_{lmax+1} enters X_lmax).  Writes partial[blockIdx.x][table entry of (m, l)][slot] for
// l <= L, tables of L + kTabExtra degrees (alpha valid for l <= L + 1).
template <int K0, int K2, int NW, int R, typename F>
__global__ __launch_bounds__(64 * NW, sizeof(F) == 4 ? 4 : 2) void k_bat_ana(const double2* __restrict__ spec0, int64_t ld0, int64_t smap0,
                                                     const double2* __restrict__ spec2, int64_t ld2, int64_t smap2,
                                                     int nside, int L, const double* __restrict__ alpha,
                                                     const double* __restrict__ smant, const int* __restrict__ sexp,
                                                     const int* __restrict__ nact, int jbeg, int jend,
                                                     F* __restrict__ partial, int64_t tsize) {
    typedef Chan<K0, K2> C;
    typedef typename Mfma16<F>::acc_t acc_t;
    typedef typename Vec4<F>::type v4;
    constexpr int T = 64 * NW;
    constexpr int CH = T * R;
    constexpr int SLOT = C::ana_stride;
    __shared__ __align__(32) F tb[NW][2 * 16 * kRow];
    const int Lt = L + kTabExtra;
    const int m = blockIdx.y;
    const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
    const int jhi = min(nact[m], jend);
    const int j0 = jbeg + blockIdx.x * CH;
    if (j0 >= jhi) return;  // uniform over the workgroup

    double x[R], mu0[R], mu1[R];
    int kx[R];
    bool inact[R], live[R];
    F av[R][2][16];  // A operands: [ring set][parity][ring quad q]
    const double sm = smant[m];
    const int se = sexp[m];
    const int kq = ln >> 4, ich = ln & 15;
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int js = j0 + (t * NW + w) * 64;
        // this lane runs ring 4 q + kq of the set with q = ln % 16, kq = ln / 16: its mu lands in
        // column ln of the transpose tile (consecutive lanes, consecutive banks), the column
        // the B operand of quad q reads for lane group kq
        const int j = js + 4 * (ln & 15) + (ln >> 4);
        live[t] = js < jhi;  // wave-uniform
        x[t] = 0.0;
        mu0[t] = 0.0;
        mu1[t] = 0.0;
        kx[t] = 0;
        if (j < jhi) {
            const Ring g = ring_info(nside, 2 * nside - 1 - j);
            x[t] = g.z;
            seed_lane<double>(sm, se, m, m, g.shalf, g.chalf, mu0[t], kx[t]);
        }
        inact[t] = __any(kx[t] < 0);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int jq = js + 4 * q + kq;
            double e = 0.0, o = 0.0;
            if (jq < jhi && ich < C::nc) chan_eo<K0, K2>(spec0, ld0, smap0, spec2, ld2, smap2, nside, L, m, jq, ich, e, o);
            av[t][0][q] = (F)e;
            av[t][1][q] = (F)o;
        }
    }

    // alpha_l at al[l]; the table array carries kAlphaTail zero entries past its end, so a
    // super-tile's reads up to degree l0 + 64 stay in bounds for every order (past the
    // table they are the next order's entries or zeros, and only feed degrees above L,
    // which are never stored)
    const double* al = alpha + table_base(m, Lt) - m;
    F* part = partial + ((int64_t)blockIdx.x * tsize + table_base(m, Lt) - m) * SLOT;
    F* mytb = tb[w];
    const int col = ln;  // transposed column of this lane's ring (4 (col % 16) + col / 16)
    for (int ls = m; ls <= L; ls += kSuper * kTile) {
        acc_t acc[kSuper][2];
#pragma unroll
        for (int h = 0; h < kSuper; ++h) {
            const int l0 = ls + h * kTile;
            double a[kTile];
#pragma unroll
            for (int i = 0; i < kTile; ++i) a[i] = al[l0 + i + 1];
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[h][0][v] = acc[h][1][v] = (F)0;
#pragma unroll
            for (int t = 0; t < R; ++t) {
                if (!live[t]) continue;
                // recurrence: mu of degrees l0 .. l0 + 31 into the tile (row = parity, degree / 2)
#pragma unroll
                for (int i4 = 0; i4 < kTile / 4; ++i4) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int i = 4 * i4 + u;
                        const double cur = (i & 1) ? mu1[t] : mu0[t];
                        mytb[((i & 1) * 16 + (i >> 1)) * kRow + col] = (F)cur;
                        if (i & 1) mu0[t] = fma(x[t] * a[i], mu1[t], -mu0[t]);
                        else mu1[t] = fma(x[t] * a[i], mu0[t], -mu1[t]);
                    }
                    if ((i4 & 1) && inact[t]) {  // every 8 degrees
                        rescale_lane(mu0[t], mu1[t], kx[t]);
                        inact[t] = __any(kx[t] < 0);
                    }
                }
                wave_sync();
                // lane (kq, n = ich) reads mu(degree 2n + p, ring 4q + kq) for q = 4g .. 4g + 3
#pragma unroll
                for (int g = 0; g < 4; ++g) {
#pragma unroll
                    for (int p = 0; p < 2; ++p) {
                        const v4 b = *reinterpret_cast<const v4*>(mytb + (p * 16 + ich) * kRow + kq * 16 + 4 * g);
#pragma unroll
                        for (int u = 0; u < 4; ++u) acc[h][p] = Mfma16<F>::op(av[t][p][4 * g + u], b[u], acc[h][p]);
                    }
                }
                wave_sync();
            }
        }
        // acc[h][p][v] at lane (kq, n): channel Mfma16<F>::row(kq, v), degree ls + 32 h + 2 n + p.
        // Each wave parks its accumulators in its (now free) transpose buffer in lane order
        // (region (h, p), lane, 4 values: consecutive lanes store consecutive chunks); thread t
        // of the workgroup sums region t / 64, lane t % 64 over the waves (reads in the same
        // order) and stores the 4 channels of its degree: FP32 rows 4 kq .. 4 kq + 3 are
        // consecutive slots (one 16-byte write), FP64 rows kq, kq + 4, kq + 8, kq + 12 are not.
        constexpr int NREG = kSuper * 2 * 64;  // (region, lane) chunks of 4 values
        static_assert(4 * NREG <= 2 * 16 * kRow, "super-tile must fit the transpose buffer");
        static_assert(SLOT % 4 == 0, "16-byte stores of 4 slots");
#pragma unroll
        for (int h = 0; h < kSuper; ++h)
#pragma unroll
            for (int p = 0; p < 2; ++p) *reinterpret_cast<acc_t*>(mytb + ((2 * h + p) * 64 + ln) * 4) = acc[h][p];
        __syncthreads();
        for (int t = threadIdx.x; t < NREG; t += T) {
            const int reg = t >> 6, lane = t & 63, n = lane & 15, g = lane >> 4;
            const int l = ls + 32 * (reg >> 1) + 2 * n + (reg & 1);
            if (l > L) continue;
            acc_t sum = {};
#pragma unroll
            for (int ww = 0; ww < NW; ++ww) sum += *reinterpret_cast<const acc_t*>(&tb[ww][4 * t]);
            if constexpr (sizeof(F) == 4) {
                if (4 * g < SLOT) *reinterpret_cast<acc_t*>(part + (int64_t)l * SLOT + 4 * g) = sum;
            } else {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int c = Mfma16<F>::row(g, v);
                    if (c < SLOT) part[(int64_t)l * SLOT + c] = sum[v];
                }
            }
        }
        __syncthreads();  // the next super-tile overwrites the buffers
    }
}

// ------------------------------------------------------------------ residual analysis
// The refinement passes analyse residual maps (map minus the synthesis of the current
// alms), a small fraction of the map, so their contraction can trade precision for rate:
// both operands are split into FP16 hi + lo parts (hi = the top 11 significand bits of
// the FP32 value, lo = the remainder rounded to FP16: together ~2^-22 relative) and
// v_mfma_f32_16x16x32_f16, at 16x the FP32 matrix rate, forms hi*hi + hi*lo + lo*hi in
// FP32 accumulators (each product to ~2^-21).  FP16's range is met by exact powers of
// two: mu is carried times 2^9 (|mu| < 2^5 on every ring outside the FP64 polar caps, so
// |2^9 mu| < 2^14), and each wave scales every channel so that its largest value over
// the wave's 64 ring pairs lies in [2^14, 2^15); the scales are divided out before the
// cross-wave sum.  Lane ln runs ring pair js + ln and writes its (hi, lo) pair as one
// word into column ln of the transpose tile (row = parity, degree / 2; two degrees per
// LDS write, as the FP32 kernel); MFMA lane (n, kg) reads the 8 words of ring pairs
// 32 b + 8 kg .. + 7 in degree row n (two 16-byte loads) and separates hi and lo with
// byte permutes.  Ring blocks, partials and their layout are those of k_bat_ana, R = 1.
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));

constexpr int kRowH = 64 + 4;      // words per transpose row: 64 ring pairs + 4 words of padding
constexpr double kMuScale = 512.0;  // 2^9

// v = hi + lo: hi keeps the top 11 significand bits (exact in FP16 for |v| >= 2^-14),
// lo = v - hi (exact in FP32) rounded to FP16; packed as (hi, lo).
__device__ inline h2 split16(float v) {
    const float hf = __uint_as_float(__float_as_uint(v) & 0xFFFFE000u);
    return h2{(_Float16)hf, (_Float16)(v - hf)};
}

// v = hi + lo as one packed word (hi in bits 0-15, lo in 16-31): hi = v rounded to FP16
// (v_cvt_f16_f32), lo = (v - hi) rounded to FP16 by v_fma_mixhi_f16 (FP16 x FP32 + FP32,
// one rounding, into the high half; v - hi is exact in FP32, so hi + lo carries v to
// ~2^-22 relative).  Two ALU operations; written out because the compiler expands the same
// C++ into an integer conversion sequence several times longer.
__device__ inline unsigned split16_packed(float v) {
    unsigned out;
    asm("v_cvt_f16_f32 %0, %1\n\tv_fma_mixhi_f16 %0, %0, %2, %1 op_sel_hi:[1,0,0]" : "=&v"(out) : "v"(v), "v"(-1.0f));
    return out;
}

template <int K0, int K2, int NW>
__global__ __launch_bounds__(64 * NW, 4) void k_bat_ana_h(const double2* __restrict__ spec0, int64_t ld0, int64_t smap0,
                                                         const double2* __restrict__ spec2, int64_t ld2, int64_t smap2,
                                                         int nside, int L, const double* __restrict__ alpha,
                                                         const double* __restrict__ smant, const int* __restrict__ sexp,
                                                         const int* __restrict__ nact, int jbeg, int jend,
                                                         float* __restrict__ partial, int64_t tsize) {
    typedef Chan<K0, K2> C;
    constexpr int kSuperH = super_h<K0>();
    constexpr int T = 64 * NW;
    constexpr int SLOT = C::ana_stride;
    constexpr int TILE = 2 * 16 * kRowH;  // words per wave: parity x 16 degree rows x row
    __shared__ __align__(32) unsigned tb[NW][TILE];
    const int Lt = L + kTabExtra;
    const int m = blockIdx.y;
    const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
    const int jhi = min(nact[m], jend);
    const int j0 = jbeg + blockIdx.x * T;
    if (j0 >= jhi) return;  // uniform over the workgroup
    const int js = j0 + w * 64;
    const bool live = js < jhi;  // wave-uniform
    const int kg = ln >> 4, ich = ln & 15;

    double x = 0.0, mu0 = 0.0, mu1 = 0.0;
    int kx = 0;
    if (js + ln < jhi) {
        const Ring g = ring_info(nside, 2 * nside - 1 - (js + ln));
        x = g.z;
        seed_lane<double>(smant[m], sexp[m], m, m, g.shalf, g.chalf, mu0, kx);
        mu0 *= kMuScale;
    }
    bool inact = __any(kx < 0);

    // A operands: channel ich of ring pairs js + 32 b + 8 kg + u (u = 0..7), parity p
    h8 ah[2][2], al[2][2];
    float inv[4];  // 1 / (2^9 x scale) of accumulator rows 4 kg + v
    {
        double ev[2][8], ov[2][8];
        double mx = 0.0;
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int jq = js + 32 * b + 8 * kg + u;
                double e = 0.0, o = 0.0;
                if (jq < jhi && ich < C::nc) chan_eo<K0, K2>(spec0, ld0, smap0, spec2, ld2, smap2, nside, L, m, jq, ich, e, o);
                ev[b][u] = e;
                ov[b][u] = o;
                mx = fmax(mx, fmax(fabs(e), fabs(o)));
            }
        mx = fmax(mx, __shfl_xor(mx, 16));
        mx = fmax(mx, __shfl_xor(mx, 32));
        const int ex = mx > 0.0 ? ilogb(mx) : 0;  // this channel's largest value lies in [2^ex, 2^(ex+1))
        const double sc = ldexp(1.0, 14 - ex);
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const h2 pe = split16((float)(ev[b][u] * sc)), po = split16((float)(ov[b][u] * sc));
                ah[0][b][u] = pe[0];
                al[0][b][u] = pe[1];
                ah[1][b][u] = po[0];
                al[1][b][u] = po[1];
            }
#pragma unroll
        for (int v = 0; v < 4; ++v) inv[v] = ldexpf(1.f, __shfl(ex, 4 * kg + v) - 14 - 9);
    }

    const double* al_ = alpha + table_base(m, Lt) - m;
    float* part = partial + ((int64_t)blockIdx.x * tsize + table_base(m, Lt) - m) * SLOT;
    unsigned* mytw = tb[w];
    for (int ls = m; ls <= L; ls += kSuperH * kTile) {
        f4 acc[kSuperH][2];
#pragma unroll
        for (int h = 0; h < kSuperH; ++h) {
            const int l0 = ls + h * kTile;
            double a[kTile];
#pragma unroll
            for (int i = 0; i < kTile; ++i) a[i] = al_[l0 + i + 1];
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[h][0][v] = acc[h][1][v] = 0.f;
            if (!live) continue;
#pragma unroll
            for (int i4 = 0; i4 < kTile / 4; ++i4) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int i = 4 * i4 + u;
                    const double cur = (i & 1) ? mu1 : mu0;
                    const unsigned pk = split16_packed((float)cur);
                    mytw[((i & 1) * 16 + (i >> 1)) * kRowH + ln] = pk;
                    if (i & 1) mu0 = fma(x * a[i], mu1, -mu0);
                    else mu1 = fma(x * a[i], mu0, -mu1);
                }
                if ((i4 & 1) && inact) {  // every 8 degrees
                    rescale_lane(mu0, mu1, kx);
                    inact = __any(kx < 0);
                }
            }
            wave_sync();
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const unsigned* src = mytw + (p * 16 + ich) * kRowH + 32 * b + 8 * kg;
                    const uint4 w0 = *reinterpret_cast<const uint4*>(src);
                    const uint4 w1 = *reinterpret_cast<const uint4*>(src + 4);
                    // words (hi_u | lo_u << 16) of ring pairs u = 0..7 -> hi halves, lo halves
                    const unsigned hw[4] = {__builtin_amdgcn_perm(w0.y, w0.x, 0x05040100u),
                                            __builtin_amdgcn_perm(w0.w, w0.z, 0x05040100u),
                                            __builtin_amdgcn_perm(w1.y, w1.x, 0x05040100u),
                                            __builtin_amdgcn_perm(w1.w, w1.z, 0x05040100u)};
                    const unsigned lw[4] = {__builtin_amdgcn_perm(w0.y, w0.x, 0x07060302u),
                                            __builtin_amdgcn_perm(w0.w, w0.z, 0x07060302u),
                                            __builtin_amdgcn_perm(w1.y, w1.x, 0x07060302u),
                                            __builtin_amdgcn_perm(w1.w, w1.z, 0x07060302u)};
                    const h8 bh = __builtin_bit_cast(h8, hw);
                    const h8 bl = __builtin_bit_cast(h8, lw);
                    acc[h][p] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[p][b], bh, acc[h][p], 0, 0, 0);
                    acc[h][p] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[p][b], bl, acc[h][p], 0, 0, 0);
                    acc[h][p] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[p][b], bh, acc[h][p], 0, 0, 0);
                }
            wave_sync();
        }
        // acc[h][p][v] at lane (kg, n): channel 4 kg + v, degree ls + 32 h + 2 n + p.  Each wave
        // parks its accumulators, unscaled, in its (now free) buffer in lane order (region
        // (h, p), lane, 4 values: consecutive lanes store consecutive 16 bytes); thread t of the
        // workgroup then sums region t / 64, lane t % 64 over the waves (reads in the same
        // order) and stores the 4 slots 4 kg .. 4 kg + 3 of its degree as one 16-byte write.
        constexpr int NREG = kSuperH * 2 * 64;  // (region, lane) chunks of 4 values
        static_assert(4 * NREG <= TILE, "super-tile must fit the buffer");
        static_assert(SLOT % 4 == 0, "16-byte stores of 4 slots");
        float* mytb = reinterpret_cast<float*>(tb[w]);
#pragma unroll
        for (int h = 0; h < kSuperH; ++h)
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                f4 o;
#pragma unroll
                for (int v = 0; v < 4; ++v) o[v] = acc[h][p][v] * inv[v];
                *reinterpret_cast<f4*>(mytb + ((2 * h + p) * 64 + ln) * 4) = o;
            }
        __syncthreads();
        for (int t = threadIdx.x; t < NREG; t += T) {
            const int reg = t >> 6, lane = t & 63, n = lane & 15, g = lane >> 4;
            const int l = ls + 32 * (reg >> 1) + 2 * n + (reg & 1);
            if (4 * g >= SLOT || l > L) continue;
            f4 sum = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ww = 0; ww < NW; ++ww) sum += *reinterpret_cast<const f4*>(reinterpret_cast<const float*>(tb[ww]) + 4 * t);
            *reinterpret_cast<f4*>(part + (int64_t)l * SLOT + 4 * g) = sum;
        }
        __syncthreads();  // the next super-tile overwrites the buffers
    }
}

// Sums of the analysis partials of (m, j) over the active ring blocks, times P_j.
struct BatBlocks {
    int CH32, jsplit, CH64, jend;
    int64_t tsize64;  // per-block stride of the FP64 partials (orders m < mpol only)
};

__device__ inline int blocks_used(int lo, int hi, int CH) { return hi > lo ? (hi - lo + CH - 1) / CH : 0; }

template <int K0, int K2>
__global__ __launch_bounds__(256) void k_bat_finish(const float* __restrict__ p32, const double* __restrict__ p64,
                                                    BatBlocks bb, const int* __restrict__ nact,
                                                    const double* __restrict__ P, int L, int64_t tsize,
                                                    double2* __restrict__ alm0, int64_t as0, double2* __restrict__ alm2,
                                                    int64_t as2, int mode) {
    typedef Chan<K0, K2> C;
    constexpr int SLOT = C::ana_stride;
    constexpr int NH = 256 + 4;  // degrees lb - 2 .. lb + 257
    __shared__ double S[NH][SLOT];
    const int Lt = L + kTabExtra;
    const int m = blockIdx.y;
    const int lb = m + blockIdx.x * 256;
    if (lb > L - 1) return;
    const int na = nact[m];
    const int n32 = blocks_used(0, min(na, bb.jsplit), bb.CH32);
    const int n64 = blocks_used(bb.jsplit, min(na, bb.jend), bb.CH64);
    const int64_t pb = table_base(m, Lt) - m;
    for (int h = threadIdx.x; h < NH; h += blockDim.x) {
        const int j = lb - 2 + h;
        double s[SLOT];
#pragma unroll
        for (int c = 0; c < SLOT; ++c) s[c] = 0.0;
        if (j >= m && j <= L) {
            for (int b = 0; b < n32; ++b) {
                const float* q = p32 + ((int64_t)b * tsize + pb + j) * SLOT;
#pragma unroll
                for (int c = 0; c < SLOT; ++c) s[c] += (double)q[c];
            }
            for (int b = 0; b < n64; ++b) {
                const double* q = p64 + ((int64_t)b * bb.tsize64 + pb + j) * SLOT;
#pragma unroll
                for (int c = 0; c < SLOT; ++c) s[c] += q[c];
            }
            const double pj = P[pb + j];
#pragma unroll
            for (int c = 0; c < SLOT; ++c) s[c] *= pj;
        }
#pragma unroll
        for (int c = 0; c < SLOT; ++c) S[h][c] = s[c];
    }
    __syncthreads();
    const int l = lb + threadIdx.x;
    if (l > L - 1) return;
    const int lmax = L - 1;
    const int64_t ia = alm_index(l, m, lmax);
    const int h = threadIdx.x + 2;  // S[h] holds degree l
    auto put = [&](double2* dst, double2 v) {
        *dst = mode == 1 ? csub(*dst, v) : mode == 2 ? cadd(*dst, v) : v;
    };
#pragma unroll
    for (int k = 0; k < K0; ++k) put(alm0 + k * as0 + ia, make_double2(S[h][2 * k], S[h][2 * k + 1]));
    double aW, bW, cW, dX, eX;
    wx_coefs(l, m, aW, bW, cW, dX, eX);
#pragma unroll
    for (int f = 0; f < K2; ++f) {
        const int b = 2 * K0 + 8 * f;
        auto cs = [&](int hh, int o) { return make_double2(S[hh][b + o], S[hh][b + o + 1]); };
        const double2 s0q = cs(h, 0), s0u = cs(h, 2);
        const double2 s1q = cs(h, 4), s1u = cs(h, 6);
        const double2 s1q_m2 = cs(h - 2, 4), s1u_m2 = cs(h - 2, 6);
        const double2 s1q_m1 = cs(h - 1, 4), s1u_m1 = cs(h - 1, 6);
        const double2 s1q_p1 = cs(h + 1, 4), s1u_p1 = cs(h + 1, 6);
        // sum_r W_l G = aW S1_l + bW S0_l + cW S1_{l-2};  sum_r X_l G = dX S1_{l+1} + eX S1_{l-1}
        const double2 wq = make_double2(aW * s1q.x + bW * s0q.x + cW * s1q_m2.x, aW * s1q.y + bW * s0q.y + cW * s1q_m2.y);
        const double2 wu = make_double2(aW * s1u.x + bW * s0u.x + cW * s1u_m2.x, aW * s1u.y + bW * s0u.y + cW * s1u_m2.y);
        const double2 xq = make_double2(dX * s1q_p1.x + eX * s1q_m1.x, dX * s1q_p1.y + eX * s1q_m1.y);
        const double2 xu = make_double2(dX * s1u_p1.x + eX * s1u_m1.x, dX * s1u_p1.y + eX * s1u_m1.y);
        // E = -(wq + i xu),  B = i xq - wu
        const double2 E = make_double2(-(wq.x - xu.y), -(wq.y + xu.x));
        const double2 B = make_double2(-xq.y - wu.x, xq.x - wu.y);
        put(alm2 + (2 * f) * as2 + ia, E);
        put(alm2 + (2 * f + 1) * as2 + ia, B);
    }
}

// ------------------------------------------------------------------ synthesis
// Coefficients per (m, degree j = m .. L + 1) and channel (synthesis layout), times P_j:
//   spin 0: a_jm;  spin 2: Q(theta) = (1/s^2) sum_j GQ_j lambda_j with
//   GQ_j = AQ_j + BQ_j (1 - eps_{j+1}^2 - eps_j^2) - BQ_{j-2} eps_{j-1} eps_j - BQ_{j+2} eps_{j+1} eps_{j+2},
//   AQ_j = -[aW_j E_j + cW_{j+2} E_{j+2} + i (dX_{j-1} B_{j-1} + eX_{j+1} B_{j+1})],  BQ_j = -bW_j E_j
//   (the plain part sum_j BQ_j lambda_j rewritten through s^2 lambda_j = lambda_j - x^2 lambda_j),
//   and GU_j likewise from AU_j = -[aW_j B_j + cW_{j+2} B_{j+2} - i (dX_{j-1} E_{j-1} + eX_{j+1} E_{j+1})],
//   BU_j = -bW_j B_j.
template <int K0, int K2>
__global__ __launch_bounds__(256) void k_bat_prep(const double2* __restrict__ alm0, int64_t as0,
                                                  const double2* __restrict__ alm2, int64_t as2,
                                                  const double* __restrict__ P, int L, double* __restrict__ coef) {
    typedef Chan<K0, K2> C;
    constexpr int NS = C::syn_stride;
    const int Lt = L + kTabExtra;
    const int m = blockIdx.y;
    const int lmax = L - 1;
    const int64_t pb = table_base(m, Lt) - m;
    const int l = m + blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= Lt + GM_TABLE_PAD) return;
    double* out = coef + (pb + l) * NS;
    const double pl = l <= L + 1 ? P[pb + l] : 0.0;
#pragma unroll
    for (int k = 0; k < K0; ++k) {
        double2 a = make_double2(0.0, 0.0);
        if (l < L) a = cscale(alm0[k * as0 + alm_index(l, m, lmax)], pl);
        out[2 * k] = a.x;
        out[2 * k + 1] = a.y;
    }
    for (int c = 2 * K0; c < C::syn_base2; ++c) out[c] = 0.0;
    const int lo = m > 2 ? m : 2;
#pragma unroll
    for (int f = 0; f < K2; ++f) {
        auto eb = [&](int comp, int j) -> double2 {
            if (j < lo || j > lmax) return make_double2(0.0, 0.0);
            return alm2[(2 * f + comp) * as2 + alm_index(j, m, lmax)];
        };
        // A part at degree j (1/s^2 coefficient before folding)
        auto apart = [&](int j, double2& aq, double2& au) {
            double aW, bW, cW, dX, eX, t1, t2, t3, t4, cW2, dXm, eXp;
            wx_coefs(j, m, aW, bW, cW, dX, eX);
            wx_coefs(j + 2, m, t1, t2, cW2, t3, t4);
            wx_coefs(j - 1, m, t1, t2, t3, dXm, t4);
            wx_coefs(j + 1, m, t1, t2, t3, t4, eXp);
            const double2 E0 = eb(0, j), B0 = eb(1, j), E2 = eb(0, j + 2), B2 = eb(1, j + 2);
            const double2 Em = eb(0, j - 1), Bm = eb(1, j - 1), Ep = eb(0, j + 1), Bp = eb(1, j + 1);
            const double wEx = aW * E0.x + cW2 * E2.x, wEy = aW * E0.y + cW2 * E2.y;
            const double wBx = aW * B0.x + cW2 * B2.x, wBy = aW * B0.y + cW2 * B2.y;
            const double xBx = dXm * Bm.x + eXp * Bp.x, xBy = dXm * Bm.y + eXp * Bp.y;
            const double xEx = dXm * Em.x + eXp * Ep.x, xEy = dXm * Em.y + eXp * Ep.y;
            aq = make_double2(-(wEx - xBy), -(wEy + xBx));  // -(wE + i xB)
            au = make_double2(-(wBx + xEy), -(wBy - xEx));  // -(wB - i xE)
        };
        // plain part at degree j: BQ_j = -bW_j E_j, BU_j = -bW_j B_j
        auto bpart = [&](int j, double2& bq, double2& bu) {
            double aW, bW, cW, dX, eX;
            wx_coefs(j, m, aW, bW, cW, dX, eX);
            const double2 E0 = eb(0, j), B0 = eb(1, j);
            bq = make_double2(-bW * E0.x, -bW * E0.y);
            bu = make_double2(-bW * B0.x, -bW * B0.y);
        };
        double2 gq = make_double2(0.0, 0.0), gu = gq;
        if (l <= L + 1) {
            double2 aq, au, bq, bu, bqm, bum, bqp, bup;
            apart(l, aq, au);
            bpart(l, bq, bu);
            bpart(l - 2, bqm, bum);
            bpart(l + 2, bqp, bup);
            const double e0 = eps_lm(l, m), e1 = eps_lm(l + 1, m), e2 = eps_lm(l + 2, m), em = eps_lm(l - 1, m);
            const double w0 = 1.0 - e1 * e1 - e0 * e0, wm = -em * e0, wp = -e1 * e2;
            gq = make_double2(aq.x + w0 * bq.x + wm * bqm.x + wp * bqp.x, aq.y + w0 * bq.y + wm * bqm.y + wp * bqp.y);
            gu = make_double2(au.x + w0 * bu.x + wm * bum.x + wp * bup.x, au.y + w0 * bu.y + wm * bum.y + wp * bup.y);
        }
        double* o = out + C::syn_base2 + 4 * f;
        o[0] = gq.x * pl;
        o[1] = gq.y * pl;
        o[2] = gu.x * pl;
        o[3] = gu.y * pl;
    }
}

__device__ inline f4 mfma4(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0); }

// Writes of ring pair j's outputs: spin-0 channel values (N = E + O, S = E - O) and the
// spin-2 components Q, U (1/s^2 part + plain part) as F+(m) = Q + iU, F+(-m) = conj Q + i conj U.
__device__ inline void put_spin2(double2* __restrict__ spec2, int64_t ld2, int L, int m, int ring, double qr, double qi,
                                 double ur, double ui) {
    double2* row = spec2 + (int64_t)ring * ld2 + (L - 1);
    row[m] = make_double2(qr - ui, qi + ur);
    if (m) row[-m] = make_double2(qr + ui, ur - qi);
}

template <int K0, int K2, int NW, int R>
__global__ __launch_bounds__(64 * NW, GM_BAT_SYN_WAVES) void k_bat_syn(const double* __restrict__ coef, int nside, int L,
                                                     const double* __restrict__ alpha, const double* __restrict__ smant,
                                                     const int* __restrict__ sexp, const int* __restrict__ nact,
                                                     int jbeg, int jend, double2* __restrict__ spec0, int64_t ld0,
                                                     int64_t smap0, double2* __restrict__ spec2, int64_t ld2,
                                                     int64_t smap2) {
    typedef Chan<K0, K2> C;
    constexpr int NS = C::syn_stride;
    constexpr int NG = NS / 4;
    constexpr int T = 64 * NW;
    constexpr int CH = T * R;
    constexpr int NV = kTile * NS / 64;  // coefficients each lane stages per block
    constexpr int SR = kTile + 4;        // staged row: one channel, kTile degrees (+ padding)
    static_assert(kTile * NS % 64 == 0, "a block of coefficients must split evenly over a wave");
    // per wave: the FP32 coefficients of kTile degrees, [channel][degree]
    __shared__ __align__(16) float cst[NW][NS * SR];
    const int Lt = L + kTabExtra;
    const int m = blockIdx.y;
    const int jhi = min(nact[m], jend);
    const int j0 = jbeg + blockIdx.x * CH;
    const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
    const int q4 = threadIdx.x & 3, qbase = threadIdx.x & ~3;

    double x[R], mu0[R], mu1[R];
    int kx[R];
    f4 acc[2][R][NG];
    const double sm = smant[m];
    const int se = sexp[m];
    bool anyact = false;
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int j = j0 + t * T + threadIdx.x;
        x[t] = mu0[t] = mu1[t] = 0.0;
        kx[t] = 0;
        if (j < jhi) {
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
        // alpha and the coefficients carry zero tails past the tables' ends (kAlphaTail,
        // kCoefTail), so whole blocks read in bounds; coefficients above degree L are zeroed
        // when staged, and alpha past L keeps mu finite for the discarded degrees
        const double* al = alpha + table_base(m, Lt) - m;
        const double* cf = coef + (table_base(m, Lt) - m) * NS;
        float* my = cst[w];
        double pf[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) pf[v] = cf[(int64_t)m * NS + v * 64 + ln];
        double an[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) an[i] = al[m + i + 1];
        for (int lb = m; lb <= L + 1; lb += kTile) {
            wave_sync();  // the previous block's reads are done
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const int e = v * 64 + ln, d = e / NS, ch = e % NS;
                my[ch * SR + d] = lb + d <= L + 1 ? (float)pf[v] : 0.f;
            }
            wave_sync();
#pragma unroll 1
            for (int i4 = 0; i4 < kTile / 4; ++i4) {
                const int l0 = lb + 4 * i4;
                if (i4 == kTile / 8) {  // the next block's coefficients, half a block ahead
#pragma unroll
                    for (int v = 0; v < NV; ++v) pf[v] = cf[(int64_t)(lb + kTile) * NS + v * 64 + ln];
                }
                double a[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    a[i] = an[i];
                    an[i] = al[l0 + 4 + i + 1];
                }
                f4 cb[NG];  // degrees l0 .. l0 + 3 of channel 4 q + q4
#pragma unroll
                for (int q = 0; q < NG; ++q) cb[q] = *reinterpret_cast<const f4*>(my + (4 * q + q4) * SR + 4 * i4);
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
                        for (int q = 0; q < NG; ++q) acc[i & 1][t][q] = mfma4(cur[t][i], cb[q][i], acc[i & 1][t][q]);
                __builtin_amdgcn_sched_barrier(0);
                if (i4 & 1) rescale<R, double>(mu0, mu1, kx, anyinact);
            }
        }
    }
    // acc[p][t][q][v] at lane (b, q4): ring pair qbase + v of set t, channel 4 q + q4
    const int NF = jend;
#pragma unroll
    for (int t = 0; t < R; ++t) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = j0 + t * T + qbase + v;
            const bool ok = j < NF;
            const int r = 2 * nside - 1 - j, rp = 2 * nside - 1 + j;
#pragma unroll
            for (int q = 0; q < C::syn_base2 / 4; ++q) {
                const int ch = 4 * q + q4;
                if (!ok || ch >= 2 * K0) continue;
                const int k = ch >> 1, comp = ch & 1;
                const double ev = acc[0][t][q][v], ov = acc[1][t][q][v];
                reinterpret_cast<double*>(spec0 + k * smap0 + (int64_t)r * ld0 + m)[comp] = ev + ov;
                if (j > 0) reinterpret_cast<double*>(spec0 + k * smap0 + (int64_t)rp * ld0 + m)[comp] = ev - ov;
            }
            double rs = 0.0;
            if (ok) {
                const Ring g = ring_info(nside, r);
                rs = 1.0 / (g.sin * g.sin);
            }
#pragma unroll
            for (int f = 0; f < K2; ++f) {
                const int qa = C::syn_base2 / 4 + f;
                // component q4 (Q re, Q im, U re, U im) of the north and south rings
                const double n = ((double)acc[0][t][qa][v] + (double)acc[1][t][qa][v]) * rs;
                const double s = ((double)acc[0][t][qa][v] - (double)acc[1][t][qa][v]) * rs;
                // lane 0 of the quad needs U im (lane 3), lane 1 U re (lane 2)
                const double n_x = __shfl_xor(n, 3), s_x = __shfl_xor(s, 3);
                if (!ok) continue;
                double2* row_n = spec2 + f * smap2 + (int64_t)r * ld2 + (L - 1);
                double2* row_s = spec2 + f * smap2 + (int64_t)rp * ld2 + (L - 1);
                if (q4 == 0) {  // re parts: F+(m) = Q re - U im, F+(-m) = Q re + U im
                    reinterpret_cast<double*>(row_n + m)[0] = n - n_x;
                    if (m) reinterpret_cast<double*>(row_n - m)[0] = n + n_x;
                    if (j > 0) {
                        reinterpret_cast<double*>(row_s + m)[0] = s - s_x;
                        if (m) reinterpret_cast<double*>(row_s - m)[0] = s + s_x;
                    }
                } else if (q4 == 1) {  // im parts: F+(m) = Q im + U re, F+(-m) = U re - Q im
                    reinterpret_cast<double*>(row_n + m)[1] = n + n_x;
                    if (m) reinterpret_cast<double*>(row_n - m)[1] = n_x - n;
                    if (j > 0) {
                        reinterpret_cast<double*>(row_s + m)[1] = s + s_x;
                        if (m) reinterpret_cast<double*>(row_s - m)[1] = s_x - s;
                    }
                }
            }
        }
    }
}

// Polar ring pairs in FP64 on the vector ALU: R ring pairs per lane, every channel in
// registers (no cross-lane exchange in the epilogue).  The coefficients of each block of
// kTile degrees are staged in a wave-private LDS buffer (one coalesced load, the next
// block's issued half a block ahead) and read back as broadcasts.
template <int K0, int K2, int NW, int R>
__global__ __launch_bounds__(64 * NW) void k_bat_syn64(const double* __restrict__ coef, int nside, int L,
                                                       const double* __restrict__ alpha,
                                                       const double* __restrict__ smant, const int* __restrict__ sexp,
                                                       const int* __restrict__ nact, int jbeg, int jend,
                                                       double2* __restrict__ spec0, int64_t ld0, int64_t smap0,
                                                       double2* __restrict__ spec2, int64_t ld2, int64_t smap2) {
    typedef Chan<K0, K2> C;
    constexpr int NS = C::syn_stride;
    constexpr int T = 64 * NW;
    constexpr int NV = kTile * NS / 64;  // coefficients each lane stages per block
    static_assert(kTile * NS % 64 == 0 && NS % 2 == 0, "blocks must split evenly over a wave");
    __shared__ __align__(16) double cst[NW][kTile * NS];
    const int Lt = L + kTabExtra;
    const int m = blockIdx.y;
    const int jhi = min(nact[m], jend);
    const int j0 = jbeg + blockIdx.x * T * R;
    double x[R], mu0[R], mu1[R];
    int kx[R];
    double acc[R][2][NS];
    bool any = false;
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int j = j0 + t * T + threadIdx.x;
        x[t] = mu0[t] = mu1[t] = 0.0;
        kx[t] = 0;
#pragma unroll
        for (int c = 0; c < NS; ++c) acc[t][0][c] = acc[t][1][c] = 0.0;
        if (j < jhi) {
            const Ring g = ring_info(nside, 2 * nside - 1 - j);
            x[t] = g.z;
            seed_lane<double>(smant[m], sexp[m], m, m, g.shalf, g.chalf, mu0[t], kx[t]);
            any = true;
        }
    }
    if (__any(any)) {
        bool anyinact = false;
#pragma unroll
        for (int t = 0; t < R; ++t) anyinact = anyinact || (kx[t] < 0);
        anyinact = __any(anyinact);
        const double* al = alpha + table_base(m, Lt) - m;
        const double* cf = coef + (table_base(m, Lt) - m) * NS;
        double* my = cst[threadIdx.x >> 6];
        const int ln = threadIdx.x & 63;
        double pf[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) pf[v] = cf[(int64_t)m * NS + v * 64 + ln];
        for (int lb = m; lb <= L + 1; lb += kTile) {
            wave_sync();  // the previous block's reads are done
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const int e = v * 64 + ln;  // = d * NS + c
                my[e] = lb + e / NS <= L + 1 ? pf[v] : 0.0;
            }
            wave_sync();
#pragma unroll 1
            for (int i4 = 0; i4 < kTile / 4; ++i4) {
                if (i4 == kTile / 8) {  // the next block's coefficients, half a block ahead
#pragma unroll
                    for (int v = 0; v < NV; ++v) pf[v] = cf[(int64_t)(lb + kTile) * NS + v * 64 + ln];
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int d = 4 * i4 + i, l = lb + d;
                    double cv[NS];
#pragma unroll
                    for (int c2 = 0; c2 < NS / 2; ++c2) {
                        const double2 v2 = *reinterpret_cast<const double2*>(my + d * NS + 2 * c2);
                        cv[2 * c2] = v2.x;
                        cv[2 * c2 + 1] = v2.y;
                    }
                    const double a = al[l + 1];
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        const double cur = (i & 1) ? mu1[t] : mu0[t];
#pragma unroll
                        for (int c = 0; c < NS; ++c) acc[t][i & 1][c] = fma(cur, cv[c], acc[t][i & 1][c]);
                        if (i & 1) mu0[t] = fma(x[t] * a, mu1[t], -mu0[t]);
                        else mu1[t] = fma(x[t] * a, mu0[t], -mu1[t]);
                    }
                }
                if (i4 & 1) rescale<R, double>(mu0, mu1, kx, anyinact);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int j = j0 + t * T + threadIdx.x;
        if (j >= jend) continue;
        const int r = 2 * nside - 1 - j, rp = 2 * nside - 1 + j;
#pragma unroll
        for (int c = 0; c < 2 * K0; ++c) {
            const int k = c >> 1, comp = c & 1;
            reinterpret_cast<double*>(spec0 + k * smap0 + (int64_t)r * ld0 + m)[comp] = acc[t][0][c] + acc[t][1][c];
            if (j > 0)
                reinterpret_cast<double*>(spec0 + k * smap0 + (int64_t)rp * ld0 + m)[comp] = acc[t][0][c] - acc[t][1][c];
        }
        const Ring g = ring_info(nside, r);
        const double rs = 1.0 / (g.sin * g.sin);
#pragma unroll
        for (int f = 0; f < K2; ++f) {
            const int b = C::syn_base2 + 4 * f;
            double nq[4], sq[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                nq[c] = (acc[t][0][b + c] + acc[t][1][b + c]) * rs;
                sq[c] = (acc[t][0][b + c] - acc[t][1][b + c]) * rs;
            }
            put_spin2(spec2 + f * smap2, ld2, L, m, r, nq[0], nq[1], nq[2], nq[3]);
            if (j > 0) put_spin2(spec2 + f * smap2, ld2, L, m, rp, sq[0], sq[1], sq[2], sq[3]);
        }
    }
}

// ------------------------------------------------------------------ launch configuration
constexpr int kAnaNW = GM_BAT_ANA_NW, kAnaR = GM_BAT_ANA_R;  // FP32 analysis: 512 ring pairs per workgroup
constexpr int kAna64NW = 4;             // FP64 analysis: 256 ring pairs per workgroup
constexpr int kSynNW = GM_BAT_SYN_NW, kSynR = GM_BAT_SYN_R;  // FP32 synthesis: 512 ring pairs per workgroup
constexpr int kSyn64NW = 2, kSyn64R = 1;  // FP64 synthesis: 128 ring pairs per workgroup

constexpr int kAnaCH32 = 64 * kAnaNW * kAnaR, kAnaCH64 = 64 * kAna64NW;

int check_bat(const char* name, int64_t k0, int64_t k2, int64_t nside, int64_t lmax, int64_t jsplit) {
    if (nside < 1 || (nside & (nside - 1)) || nside > 8192) return fail(GM_E_ARG, "%s: bad nside", name);
    if (lmax < 2 || lmax > 4 * nside) return fail(GM_E_ARG, "%s: lmax must be in [2, 4 nside]", name);
    if (jsplit < 0 || jsplit > 2 * nside) return fail(GM_E_ARG, "%s: jsplit outside [0, 2 nside]", name);
    const bool ok = (k0 == 2 && k2 == 1) || (k0 == 1 && k2 == 1) || (k0 == 0 && k2 == 1) || (k0 == 0 && k2 == 2) ||
                    (k0 == 2 && k2 == 0) || (k0 == 4 && k2 == 0) || (k0 == 8 && k2 == 0);
    if (!ok) return fail(GM_E_ARG, "%s: unsupported batch (%lld spin-0 maps, %lld spin-2 fields)", name, (long long)k0,
                         (long long)k2);
    return GM_OK;
}

#define GM_BAT_DISPATCH(K0V, K2V, CALL)                                   \
    if (k0 == 2 && k2 == 1) { constexpr int K0V = 2, K2V = 1; CALL; }       \
    else if (k0 == 1 && k2 == 1) { constexpr int K0V = 1, K2V = 1; CALL; }  \
    else if (k0 == 0 && k2 == 1) { constexpr int K0V = 0, K2V = 1; CALL; }  \
    else if (k0 == 0 && k2 == 2) { constexpr int K0V = 0, K2V = 2; CALL; }  \
    else if (k0 == 2 && k2 == 0) { constexpr int K0V = 2, K2V = 0; CALL; }  \
    else if (k0 == 4 && k2 == 0) { constexpr int K0V = 4, K2V = 0; CALL; }  \
    else { constexpr int K0V = 8, K2V = 0; CALL; }

int64_t stride_ana(int64_t k0, int64_t k2) { return (2 * k0 + 8 * k2 + 3) / 4 * 4; }
int64_t stride_syn(int64_t k0, int64_t k2) { return (2 * k0 + 3) / 4 * 4 + 8 * k2; }

}  // namespace
}  // namespace gm

using namespace gm;

extern "C" int64_t gm_bat_blocks(int64_t nside, int64_t jsplit, int64_t polar) {
    if (nside < 1 || jsplit < 0 || jsplit > 2 * nside) return -1;
    return polar ? (2 * nside - jsplit + kAnaCH64 - 1) / kAnaCH64 : (jsplit + kAnaCH32 - 1) / kAnaCH32;
}

extern "C" int64_t gm_bat_stride(int64_t k0, int64_t k2, int64_t synthesis) {
    return synthesis ? stride_syn(k0, k2) : stride_ana(k0, k2);
}

namespace {
int bat_ana(const char* name, bool residual, const double* spec0, int64_t k0, int64_t ld0, int64_t smap0,
            const double* spec2, int64_t k2, int64_t ld2, int64_t smap2, int64_t nside, int64_t lmax,
            const double* alpha, const double* seed_mant, const int32_t* seed_exp, const int32_t* nactive,
            int64_t jsplit, int64_t mpol, float* part32, double* part64, void* stream) {
    if (int e = check_bat(name, k0, k2, nside, lmax, jsplit)) return e;
    if (mpol < 0 || mpol > lmax + 1) return fail(GM_E_ARG, "%s: mpol outside [0, lmax + 1]", name);
    if ((k0 && !spec0) || (k2 && !spec2) || !alpha || !seed_mant || !seed_exp || !nactive)
        return fail(GM_E_ARG, "%s: null pointer", name);
    if ((jsplit > 0 && !part32) || (jsplit < 2 * nside && mpol > 0 && !part64))
        return fail(GM_E_ARG, "%s: null partial", name);
    const int64_t L = lmax + 1;
    if (k0 && ld0 < L) return fail(GM_E_SHAPE, "%s: spin-0 spectrum ld < lmax + 1", name);
    if (k2 && ld2 < 2 * L - 1) return fail(GM_E_SHAPE, "%s: spin-2 spectrum ld < 2 lmax + 1", name);
    const int64_t nring = 4 * nside - 1;
    if (k0 > 1 && smap0 < nring * ld0) return fail(GM_E_SHAPE, "%s: spin-0 map stride too small", name);
    if (k2 > 1 && smap2 < nring * ld2) return fail(GM_E_SHAPE, "%s: spin-2 field stride too small", name);
    const int64_t tsize = table_size(L + kTabExtra);
    hipStream_t st = (hipStream_t)stream;
    const auto* s0 = reinterpret_cast<const double2*>(spec0);
    const auto* s2 = reinterpret_cast<const double2*>(spec2);
    const int nb32 = (int)gm_bat_blocks(nside, jsplit, 0), nb64 = (int)gm_bat_blocks(nside, jsplit, 1);
    static_assert(kAnaCH32 % 64 == 0 && kAnaCH32 <= 1024, "residual kernel: one ring pair per lane, same blocks");
    GM_BAT_DISPATCH(K0, K2, {
        if (nb32 > 0 && residual)
            hipLaunchKernelGGL((k_bat_ana_h<K0, K2, kAnaCH32 / 64>), dim3(nb32, (unsigned)L), dim3(kAnaCH32), 0, st, s0,
                               ld0, smap0, s2, ld2, smap2, (int)nside, (int)L, alpha, seed_mant, seed_exp, nactive, 0,
                               (int)jsplit, part32, tsize);
        else if (nb32 > 0)
            hipLaunchKernelGGL((k_bat_ana<K0, K2, kAnaNW, kAnaR, float>), dim3(nb32, (unsigned)L), dim3(64 * kAnaNW), 0,
                               st, s0, ld0, smap0, s2, ld2, smap2, (int)nside, (int)L, alpha, seed_mant, seed_exp,
                               nactive, 0, (int)jsplit, part32, tsize);
        if (nb64 > 0 && mpol > 0)
            hipLaunchKernelGGL((k_bat_ana<K0, K2, kAna64NW, 1, double>), dim3(nb64, (unsigned)mpol),
                               dim3(64 * kAna64NW), 0, st, s0, ld0, smap0, s2, ld2, smap2, (int)nside, (int)L, alpha,
                               seed_mant, seed_exp, nactive, (int)jsplit, (int)(2 * nside), part64,
                               table_base(mpol, L + kTabExtra));
    })
    return check_launch(name);
}
}  // namespace

extern "C" int gm_bat_ana(const double* spec0, int64_t k0, int64_t ld0, int64_t smap0, const double* spec2, int64_t k2,
                          int64_t ld2, int64_t smap2, int64_t nside, int64_t lmax, const double* alpha,
                          const double* seed_mant, const int32_t* seed_exp, const int32_t* nactive, int64_t jsplit,
                          int64_t mpol, float* part32, double* part64, void* stream) {
    return bat_ana("gm_bat_ana", false, spec0, k0, ld0, smap0, spec2, k2, ld2, smap2, nside, lmax, alpha, seed_mant,
                   seed_exp, nactive, jsplit, mpol, part32, part64, stream);
}

extern "C" int gm_bat_ana_res(const double* spec0, int64_t k0, int64_t ld0, int64_t smap0, const double* spec2,
                              int64_t k2, int64_t ld2, int64_t smap2, int64_t nside, int64_t lmax, const double* alpha,
                              const double* seed_mant, const int32_t* seed_exp, const int32_t* nactive, int64_t jsplit,
                              int64_t mpol, float* part32, double* part64, void* stream) {
    return bat_ana("gm_bat_ana_res", true, spec0, k0, ld0, smap0, spec2, k2, ld2, smap2, nside, lmax, alpha, seed_mant,
                   seed_exp, nactive, jsplit, mpol, part32, part64, stream);
}

extern "C" int gm_bat_finish(const float* part32, const double* part64, int64_t k0, int64_t k2, int64_t nside,
                             int64_t lmax, const int32_t* nactive, int64_t jsplit, int64_t mpol, const double* P,
                             double* alm0, int64_t astride0, double* alm2, int64_t astride2, int mode, void* stream) {
    const char* name = "gm_bat_finish";
    if (int e = check_bat(name, k0, k2, nside, lmax, jsplit)) return e;
    if (mpol < 0 || mpol > lmax + 1) return fail(GM_E_ARG, "%s: mpol outside [0, lmax + 1]", name);
    if (mode < 0 || mode > 2) return fail(GM_E_ARG, "%s: bad mode", name);
    if (!nactive || !P || (k0 && !alm0) || (k2 && !alm2)) return fail(GM_E_ARG, "%s: null pointer", name);
    if ((jsplit > 0 && !part32) || (jsplit < 2 * nside && mpol > 0 && !part64))
        return fail(GM_E_ARG, "%s: null partial", name);
    const int64_t L = lmax + 1, nelem = L * (L + 1) / 2;
    if ((k0 > 1 && astride0 < nelem) || (k2 && astride2 < nelem)) return fail(GM_E_SHAPE, "%s: alm stride too small", name);
    // orders m >= mpol have no ring pair beyond jsplit (their nactive <= jsplit): no FP64 partials
    BatBlocks bb{kAnaCH32, (int)jsplit, kAnaCH64, (int)(2 * nside), table_base(mpol, L + kTabExtra)};
    const int64_t tsize = table_size(L + kTabExtra);
    hipStream_t st = (hipStream_t)stream;
    auto* a0 = reinterpret_cast<double2*>(alm0);
    auto* a2 = reinterpret_cast<double2*>(alm2);
    GM_BAT_DISPATCH(K0, K2, {
        hipLaunchKernelGGL((k_bat_finish<K0, K2>), dim3((unsigned)((L + 255) / 256), (unsigned)L), dim3(256), 0, st,
                           part32, part64, bb, nactive, P, (int)L, tsize, a0, astride0, a2, astride2, mode);
    })
    return check_launch(name);
}

extern "C" int gm_bat_prep(const double* alm0, int64_t k0, int64_t astride0, const double* alm2, int64_t k2,
                           int64_t astride2, const double* P, int64_t lmax, double* coef, void* stream) {
    const char* name = "gm_bat_prep";
    if (!P || !coef || (k0 && !alm0) || (k2 && !alm2)) return fail(GM_E_ARG, "%s: null pointer", name);
    if (lmax < 2) return fail(GM_E_ARG, "%s: lmax < 2", name);
    if (int e = check_bat(name, k0, k2, 8192, 2, 0)) return e;
    const int64_t L = lmax + 1, nelem = L * (L + 1) / 2;
    if ((k0 > 1 && astride0 < nelem) || (k2 && astride2 < nelem)) return fail(GM_E_SHAPE, "%s: alm stride too small", name);
    hipStream_t st = (hipStream_t)stream;
    const auto* a0 = reinterpret_cast<const double2*>(alm0);
    const auto* a2 = reinterpret_cast<const double2*>(alm2);
    GM_BAT_DISPATCH(K0, K2, {
        hipLaunchKernelGGL((k_bat_prep<K0, K2>), dim3((unsigned)((L + kTabExtra + GM_TABLE_PAD + 255) / 256), (unsigned)L),
                           dim3(256), 0, st, a0, astride0, a2, astride2, P, (int)L, coef);
    })
    return check_launch(name);
}

extern "C" int gm_bat_syn(const double* coef, int64_t k0, int64_t k2, int64_t nside, int64_t lmax, const double* alpha,
                          const double* seed_mant, const int32_t* seed_exp, const int32_t* nactive, int64_t jsplit,
                          double* spec0, int64_t ld0, int64_t smap0, double* spec2, int64_t ld2, int64_t smap2,
                          void* stream) {
    const char* name = "gm_bat_syn";
    if (int e = check_bat(name, k0, k2, nside, lmax, jsplit)) return e;
    if (!coef || !alpha || !seed_mant || !seed_exp || !nactive || (k0 && !spec0) || (k2 && !spec2))
        return fail(GM_E_ARG, "%s: null pointer", name);
    const int64_t L = lmax + 1;
    if (k0 && ld0 < L) return fail(GM_E_SHAPE, "%s: spin-0 spectrum ld < lmax + 1", name);
    if (k2 && ld2 < 2 * L - 1) return fail(GM_E_SHAPE, "%s: spin-2 spectrum ld < 2 lmax + 1", name);
    const int64_t nring = 4 * nside - 1;
    if (k0 > 1 && smap0 < nring * ld0) return fail(GM_E_SHAPE, "%s: spin-0 map stride too small", name);
    if (k2 > 1 && smap2 < nring * ld2) return fail(GM_E_SHAPE, "%s: spin-2 field stride too small", name);
    hipStream_t st = (hipStream_t)stream;
    auto* s0 = reinterpret_cast<double2*>(spec0);
    auto* s2 = reinterpret_cast<double2*>(spec2);
    constexpr int CH = 64 * kSynNW * kSynR, CH64 = 64 * kSyn64NW * kSyn64R;
    const int nb32 = (int)((jsplit + CH - 1) / CH), nb64 = (int)((2 * nside - jsplit + CH64 - 1) / CH64);
    GM_BAT_DISPATCH(K0, K2, {
        if (nb32 > 0)
            hipLaunchKernelGGL((k_bat_syn<K0, K2, kSynNW, kSynR>), dim3(nb32, (unsigned)L), dim3(64 * kSynNW), 0, st,
                               coef, (int)nside, (int)L, alpha, seed_mant, seed_exp, nactive, 0, (int)jsplit, s0, ld0,
                               smap0, s2, ld2, smap2);
        if (nb64 > 0)
            hipLaunchKernelGGL((k_bat_syn64<K0, K2, kSyn64NW, kSyn64R>), dim3(nb64, (unsigned)L), dim3(64 * kSyn64NW), 0, st,
                               coef, (int)nside, (int)L, alpha, seed_mant, seed_exp, nactive, (int)jsplit,
                               (int)(2 * nside), s0, ld0, smap0, s2, ld2, smap2);
    })
    return check_launch(name);
}
