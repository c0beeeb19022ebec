// Mode-coupling matrices on gfx950, FP64: the Wigner-3j rows (l1 l2 l3; 0 0 0) and
// (l1 l2 l3; 2 -2 0) are walked in l3 for each pair l1 <= l2 and all four MASTER blocks
// accumulate at once:
//     M00  = (2 l2 + 1) sum W'_l3 w0^2            M02  = (2 l2 + 1) sum W'_l3 w0 w2
//     M22+ = (2 l2 + 1) sum_{J even} W'_l3 w2^2    M22- = (2 l2 + 1) sum_{J odd} W'_l3 w2^2
// with W'_l = (2l + 1) W_l / (4 pi) (the caller passes W') and J = l1 + l2 + l3 (general
// coupling formula of gmaster/workspaces.py:1069-1077 with (s1, s2, n1, n2) = (0,0,0,0),
// (0,2,0,2), (2,2,2,2)).  The recursion itself (square-root free, Horner-form sums) is in
// coupling_rec.h.  The row sums are symmetric, so each pair writes (l1, l2) and (l2, l1).
//
// Full matrix (k_coupling): one wave per (l1, 64 consecutive l2).  Every lane of the wave
// has the same number of degrees (2 l1 + 1, or fewer where the window ends), so the loop is
// uniform; W' is staged in a wave-private LDS window of 128 entries per 64 degrees (the
// 64 lanes read 64 consecutive entries per degree, one ds_read2 per pair of degrees), zero
// past the window's end.  Waves take rows in descending l1, longest first.
#include "common.h"
#include "coupling_rec.h"

namespace gm {
namespace {

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// W' window of one wave in LDS (128 entries, see couple_wave_lane).
struct LdsWin {
    double* buf;
    const double* __restrict__ wp;
    int lmax_w, d0, lane;
    __device__ __forceinline__ void stage(int k0) {
        const int b = d0 + k0 + lane;
        const double v0 = b <= lmax_w ? wp[b] : 0.0;
        const double v1 = b + 64 <= lmax_w ? wp[b + 64] : 0.0;
        wave_sync();  // earlier reads of the window are done
        buf[lane] = v0;
        buf[lane + 64] = v1;
        wave_sync();
    }
    __device__ __forceinline__ double at(int t) const { return buf[t]; }
};

__device__ __forceinline__ void store_pair(int l1, int l2, int L, const double s[4], double* __restrict__ m00,
                                           double* __restrict__ m02, double* __restrict__ m22p,
                                           double* __restrict__ m22m) {
    const double c2 = 2.0 * l2 + 1.0, c1 = 2.0 * l1 + 1.0;
    const int64_t i12 = (int64_t)l1 * L + l2, i21 = (int64_t)l2 * L + l1;
    if (m00) { m00[i12] = c2 * s[0]; m00[i21] = c1 * s[0]; }
    if (m02) { m02[i12] = c2 * s[1]; m02[i21] = c1 * s[1]; }
    if (m22p) { m22p[i12] = c2 * s[2]; m22p[i21] = c1 * s[2]; }
    if (m22m) { m22m[i12] = c2 * s[3]; m22m[i21] = c1 * s[3]; }
}

__global__ __launch_bounds__(256) void k_coupling(const double* __restrict__ Wp, int lmax_w,
                                                  const double* __restrict__ g, int L, int64_t nitems,
                                                  double* __restrict__ m00, double* __restrict__ m02,
                                                  double* __restrict__ m22p, double* __restrict__ m22m) {
    __shared__ double win_buf[4][128];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t item = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(wave);
    if (item >= nitems) return;
    int l1, j;
    couple_decode(item, L, l1, j);
    l1 = __builtin_amdgcn_readfirstlane(l1);
    j = __builtin_amdgcn_readfirstlane(j);
    LdsWin win{win_buf[wave], Wp, lmax_w, 64 * j, lane};
    double s[4];
    couple_wave_lane(win, lmax_w, g, L, l1, l1 + 64 * j, lane, s);
    const int l2 = l1 + 64 * j + lane;
    if (l2 < L) store_pair(l1, l2, L, s, m00, m02, m22p, m22m);
}

// Listed row segments, one wave per item: items[3 k .. 3 k + 2] = (l1, l2_0, nlive), the
// pairs (l1, l2_0 + i) for i < nlive <= 64 (l1 <= l2_0, l2_0 + nlive <= L).
__global__ __launch_bounds__(256) void k_coupling_rows(const double* __restrict__ Wp, int lmax_w,
                                                       const double* __restrict__ g, int L,
                                                       const int32_t* __restrict__ items, int64_t nitems,
                                                       double* __restrict__ m00, double* __restrict__ m02,
                                                       double* __restrict__ m22p, double* __restrict__ m22m) {
    __shared__ double win_buf[4][128];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t item = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(wave);
    if (item >= nitems) return;
    const int l1 = __builtin_amdgcn_readfirstlane(items[3 * item]);
    const int l2_0 = __builtin_amdgcn_readfirstlane(items[3 * item + 1]);
    int nlive = __builtin_amdgcn_readfirstlane(items[3 * item + 2]);
    if (l1 < 0 || l2_0 < l1 || l2_0 >= L || nlive < 1) return;  // malformed item: nothing to do
    nlive = min(nlive, min(64, L - l2_0));
    LdsWin win{win_buf[wave], Wp, lmax_w, l2_0 - l1, lane};
    double s[4];
    couple_wave_lane(win, lmax_w, g, L, l1, l2_0, lane, s);
    if (lane < nlive) store_pair(l1, l2_0 + lane, L, s, m00, m02, m22p, m22m);
}

// The four row sums (without the column factor) of each listed pair, vals[k][0..3].
// (A malformed pair, outside 0 <= lo <= hi < L, gets NaN.)
__global__ void k_coupling_pair_values(const double* __restrict__ Wp, int lmax_w, const double* __restrict__ g,
                                       int L, const int32_t* __restrict__ lo, const int32_t* __restrict__ hi,
                                       int64_t npairs, double* __restrict__ vals) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npairs) return;
    const int l1 = lo[k], l2 = hi[k];
    double s[4];
    if (l1 < 0 || l2 < l1 || l2 >= L)
        s[0] = s[1] = s[2] = s[3] = __builtin_nan("");
    else
        couple_pair(Wp, lmax_w, g, l1, l2, s);
    for (int q = 0; q < 4; ++q) vals[k * 4 + q] = s[q];
}

// Only the listed pairs (lo[k] <= hi[k]); every other element of the outputs is left as it was.
__global__ void k_coupling_pairs(const double* __restrict__ Wp, int lmax_w, const double* __restrict__ g, int L,
                                 const int32_t* __restrict__ lo, const int32_t* __restrict__ hi, int64_t npairs,
                                 double* __restrict__ m00, double* __restrict__ m02, double* __restrict__ m22p,
                                 double* __restrict__ m22m) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npairs) return;
    const int l1 = lo[k], l2 = hi[k];
    if (l1 < 0 || l2 < l1 || l2 >= L) return;  // malformed pair: nothing to do
    double s[4];
    couple_pair(Wp, lmax_w, g, l1, l2, s);
    store_pair(l1, l2, L, s, m00, m02, m22p, m22m);
}

// Toeplitz approximation of one coupling block (Louis et al. 2020, NaMaster's fill order as
// in gmaster/workspaces.py:1141-1194), on the master coefficients S = M / (2 l' + 1).  The
// reference applies a sequence of overlapping fills to the lower triangle and then mirrors
// it; here each output element evaluates the last fill that covers it:
//   pure(i, j) = c_t[(i - j + lt) mod n] sqrt(D_i D_j),  c_t[k] = S[k][lt] / sqrt(D_k D_lt)
//   j <= le or i - j <= dl -> exact;  i, j >= lt -> pure;  i == j -> exact;
//   1 <= i - j <= dl with j <= lt -> exact;
//   le <= j <= lt, i >= n - 1 - lt + j -> c_e[le + i - j] sqrt(D_i D_j),  c_e[k] = S[k][le] / sqrt(D_k D_le)
// for i = max(row, col), j = min(row, col); band and diagonal read S from the upper
// triangle exactly as the reference's diag(S, offset) does.
//
// A workgroup fills a tile of kTR rows x kTC columns.  D_k, c_t and c_e depend on one index
// each and come from the diagonal and two columns of M (strided reads), so the tile stages
// the D of its rows and columns and c_t, c_e over its range of i - j in LDS once, instead
// of every element reading them from their own cache lines.
constexpr int kTR = 32, kTC = 256;

__global__ __launch_bounds__(kTC) void k_toeplitz(const double* __restrict__ M, int n, int lt, int le, int dl,
                                                 double* __restrict__ out) {
    __shared__ double Dc[kTC], Dr[kTR], ctv[kTC + kTR], cev[kTC + kTR];
    const int b0 = blockIdx.x * kTC, a0 = blockIdx.y * kTR;
    const int tid = threadIdx.x;
    auto S = [&](int r, int c) { return M[(int64_t)r * n + c] / (2.0 * c + 1.0); };
    auto D = [&](int k) { return fabs(S(k, k)); };
    const int amax = min(a0 + kTR, n) - 1, bmax = min(b0 + kTC, n) - 1;
    int dlo, dhi;  // range of |row - col| over the tile
    if (bmax < a0) {
        dlo = a0 - bmax;
        dhi = amax - b0;
    } else if (b0 > amax) {
        dlo = b0 - amax;
        dhi = bmax - a0;
    } else {
        dlo = 0;
        dhi = max(amax - b0, bmax - a0);
    }
    {
        const int b = b0 + tid;
        Dc[tid] = b < n ? D(b) : 0.0;
        if (tid < kTR) Dr[tid] = a0 + tid < n ? D(a0 + tid) : 0.0;
        const double Dlt = D(lt), Dle = D(le);
        for (int t = tid; t <= dhi - dlo; t += kTC) {
            const int dd = dlo + t;
            const int kt = (dd + lt) % n;
            const double dent = sqrt(D(kt) * Dlt);
            ctv[t] = dent > 0.0 ? S(kt, lt) / dent : 0.0;
            const int ke = le + dd;
            double ce = 0.0;
            if (ke < n) {
                const double Dk = D(ke);
                const double den = sqrt(Dk * Dle);
                ce = Dk > 0.0 ? S(ke, le) / den : 0.0;  // guarded by D_k only, as the reference
            }
            cev[t] = ce;
        }
    }
    __syncthreads();
    const int b = b0 + tid;
    if (b >= n) return;
    for (int r = 0; r < kTR && a0 + r < n; ++r) {
        const int a = a0 + r;
        const int i = a > b ? a : b, j = a > b ? b : a;
        const double Di = a > b ? Dr[r] : Dc[tid], Dj = a > b ? Dc[tid] : Dr[r];
        const int t = i - j - dlo;
        double v = ctv[t] * sqrt(Di * Dj);
        if (j <= le) v = S(i, j);
        if (i - j <= dl) v = (i == j) ? S(i, i) : S(j, i);
        if (i >= lt && j >= lt) v = ctv[t] * sqrt(Di * Dj);
        if (i == j) v = S(i, i);
        if (i - j >= 1 && i - j <= dl && j <= lt) v = S(j, i);
        if (j >= le && j <= lt && i >= n - 1 - lt + j) v = cev[t] * sqrt(Dj * Di);
        out[(int64_t)a * n + b] = v * (2.0 * b + 1.0);
    }
}

// Which fill of the Toeplitz approximation sets element (i >= j): 0 = an exact value,
// 1 = the pure Toeplitz extension, 2 = the corner extrapolation from column l_exact (the
// same precedence as k_toeplitz).
__device__ __forceinline__ int toeplitz_kind(int i, int j, int n, int lt, int le, int dl) {
    int kind = 1;
    if (j <= le) kind = 0;
    if (i - j <= dl) kind = 0;
    if (i >= lt && j >= lt) kind = 1;
    if (i == j) kind = 0;
    if (i - j >= 1 && i - j <= dl && j <= lt) kind = 0;
    if (j >= le && j <= lt && i >= n - 1 - lt + j) kind = 2;
    return kind;
}

// Learned correction of a Toeplitz-approximated block M (column factor included), in place:
// for every element not set exactly, with j = min(row, col) and d = |row - col|,
//     M += e(j, d) sqrt(D_j D_{j+d}) (2 col + 1),   D_k = |M_kk| / (2k + 1),
// e bilinear in (log j, log d) over the grid lg[0..nl) x dg[0..nd) (increasing integers)
// and zero outside it.
__global__ void k_toeplitz_correct(double* __restrict__ M, int n, int lt, int le, int dl,
                                   const int32_t* __restrict__ lg, int nl, const int32_t* __restrict__ dg, int nd,
                                   const float* __restrict__ e) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    if (b >= n) return;
    const int i = a > b ? a : b, j = a > b ? b : a, d = i - j;
    if (j < lg[0] || j > lg[nl - 1] || d < dg[0] || d > dg[nd - 1]) return;
    if (toeplitz_kind(i, j, n, lt, le, dl) == 0) return;
    int p = 0, q = 0;
    while (p + 2 < nl && lg[p + 1] <= j) ++p;
    while (q + 2 < nd && dg[q + 1] <= d) ++q;
    const double tl = nl > 1 ? (log((double)j) - log((double)lg[p])) / (log((double)lg[p + 1]) - log((double)lg[p])) : 0.0;
    const double td = nd > 1 ? (log((double)d) - log((double)dg[q])) / (log((double)dg[q + 1]) - log((double)dg[q])) : 0.0;
    const int p1 = nl > 1 ? p + 1 : p, q1 = nd > 1 ? q + 1 : q;
    const double v = (1.0 - tl) * ((1.0 - td) * e[p * nd + q] + td * e[p * nd + q1]) +
                     tl * ((1.0 - td) * e[p1 * nd + q] + td * e[p1 * nd + q1]);
    const double Di = fabs(M[(int64_t)i * n + i]) / (2.0 * i + 1.0);
    const double Dj = fabs(M[(int64_t)j * n + j]) / (2.0 * j + 1.0);
    M[(int64_t)a * n + b] += v * sqrt(Di * Dj) * (2.0 * b + 1.0);
}

}  // namespace
}  // namespace gm

using namespace gm;

extern "C" int gm_coupling_pair_values(const double* window, int64_t lmax_window, const double* g, int64_t lmax,
                                       const int32_t* lo, const int32_t* hi, int64_t npairs, double* vals,
                                       void* stream) {
    if (!window || !g || !vals || (npairs > 0 && (!lo || !hi)))
        return fail(GM_E_ARG, "gm_coupling_pair_values: null pointer");
    if (lmax < 0 || lmax > 65535) return fail(GM_E_ARG, "gm_coupling_pair_values: bad lmax %lld", (long long)lmax);
    if (lmax_window < 0 || lmax_window > 2 * lmax)
        return fail(GM_E_ARG, "gm_coupling_pair_values: need 0 <= lmax_window <= 2 lmax");
    if (npairs < 0) return fail(GM_E_ARG, "gm_coupling_pair_values: negative pair count");
    if (npairs == 0) return GM_OK;
    hipLaunchKernelGGL(k_coupling_pair_values, dim3((unsigned)((npairs + 127) / 128)), dim3(128), 0,
                       (hipStream_t)stream, window, (int)lmax_window, g, (int)lmax + 1, lo, hi, npairs, vals);
    return check_launch("gm_coupling_pair_values");
}

extern "C" int gm_toeplitz_correct(double* M, int64_t lmax, int64_t l_toeplitz, int64_t l_exact, int64_t dl_band,
                                   const int32_t* l_grid, int64_t n_l, const int32_t* d_grid, int64_t n_d,
                                   const float* e, void* stream) {
    if (!M || !l_grid || !d_grid || !e) return fail(GM_E_ARG, "gm_toeplitz_correct: null pointer");
    if (lmax < 1 || l_toeplitz <= 0 || l_exact <= 0 || dl_band < 0 || l_exact > l_toeplitz || l_toeplitz >= lmax ||
        dl_band >= lmax)
        return fail(GM_E_ARG, "gm_toeplitz_correct: need 0 < l_exact <= l_toeplitz < lmax and 0 <= dl_band < lmax");
    if (n_l < 1 || n_d < 1 || n_l > 4096 || n_d > 4096) return fail(GM_E_ARG, "gm_toeplitz_correct: bad grid sizes");
    const int n = (int)lmax + 1;
    hipLaunchKernelGGL(k_toeplitz_correct, dim3((unsigned)((n + 255) / 256), (unsigned)n), dim3(256), 0,
                       (hipStream_t)stream, M, n, (int)l_toeplitz, (int)l_exact, (int)dl_band, l_grid, (int)n_l,
                       d_grid, (int)n_d, e);
    return check_launch("gm_toeplitz_correct");
}

extern "C" int gm_toeplitz(const double* M, int64_t lmax, int64_t l_toeplitz, int64_t l_exact, int64_t dl_band,
                           double* out, void* stream) {
    if (!M || !out) return fail(GM_E_ARG, "gm_toeplitz: null pointer");
    if (M == out) return fail(GM_E_ARG, "gm_toeplitz: out must not alias M");
    if (lmax < 1 || l_toeplitz <= 0 || l_exact <= 0 || dl_band < 0 || l_exact > l_toeplitz ||
        l_toeplitz >= lmax || dl_band >= lmax)
        return fail(GM_E_ARG, "gm_toeplitz: need 0 < l_exact <= l_toeplitz < lmax and 0 <= dl_band < lmax");
    const int n = (int)lmax + 1;
    hipLaunchKernelGGL(k_toeplitz, dim3((unsigned)((n + kTC - 1) / kTC), (unsigned)((n + kTR - 1) / kTR)), dim3(kTC), 0,
                       (hipStream_t)stream,
                       M, n, (int)l_toeplitz, (int)l_exact, (int)dl_band, out);
    return check_launch("gm_toeplitz");
}

extern "C" int gm_coupling(const double* window, int64_t lmax_window, const double* g, int64_t lmax, double* m00,
                           double* m02, double* m22p, double* m22m, void* stream) {
    if (!window || !g) return fail(GM_E_ARG, "gm_coupling: null window / g table");
    if (lmax < 0 || lmax > 65535) return fail(GM_E_ARG, "gm_coupling: bad lmax %lld", (long long)lmax);
    if (!m00 && !m02 && !m22p && !m22m) return fail(GM_E_ARG, "gm_coupling: no output requested");
    if (lmax_window < 0 || lmax_window > 2 * lmax) return fail(GM_E_ARG, "gm_coupling: need 0 <= lmax_window <= 2 lmax");
    const int L = (int)lmax + 1;
    const int64_t items = couple_items(L);
    hipLaunchKernelGGL(k_coupling, dim3((unsigned)((items + 3) / 4)), dim3(256), 0, (hipStream_t)stream, window,
                       (int)lmax_window, g, L, items, m00, m02, m22p, m22m);
    return check_launch("gm_coupling");
}

extern "C" int gm_coupling_rows(const double* window, int64_t lmax_window, const double* g, int64_t lmax,
                                const int32_t* items, int64_t nitems, double* m00, double* m02, double* m22p,
                                double* m22m, void* stream) {
    if (!window || !g || (nitems > 0 && !items)) return fail(GM_E_ARG, "gm_coupling_rows: null pointer");
    if (lmax < 0 || lmax > 65535) return fail(GM_E_ARG, "gm_coupling_rows: bad lmax %lld", (long long)lmax);
    if (!m00 && !m02 && !m22p && !m22m) return fail(GM_E_ARG, "gm_coupling_rows: no output requested");
    if (lmax_window < 0 || lmax_window > 2 * lmax)
        return fail(GM_E_ARG, "gm_coupling_rows: need 0 <= lmax_window <= 2 lmax");
    if (nitems < 0) return fail(GM_E_ARG, "gm_coupling_rows: negative item count");
    if (nitems == 0) return GM_OK;
    hipLaunchKernelGGL(k_coupling_rows, dim3((unsigned)((nitems + 3) / 4)), dim3(256), 0, (hipStream_t)stream, window,
                       (int)lmax_window, g, (int)lmax + 1, items, nitems, m00, m02, m22p, m22m);
    return check_launch("gm_coupling_rows");
}

extern "C" int gm_coupling_pairs(const double* window, int64_t lmax_window, const double* g, int64_t lmax,
                                 const int32_t* lo, const int32_t* hi, int64_t npairs, double* m00, double* m02,
                                 double* m22p, double* m22m, void* stream) {
    if (!window || !g || (npairs > 0 && (!lo || !hi))) return fail(GM_E_ARG, "gm_coupling_pairs: null pointer");
    if (lmax < 0 || lmax > 65535) return fail(GM_E_ARG, "gm_coupling_pairs: bad lmax %lld", (long long)lmax);
    if (!m00 && !m02 && !m22p && !m22m) return fail(GM_E_ARG, "gm_coupling_pairs: no output requested");
    if (lmax_window < 0 || lmax_window > 2 * lmax)
        return fail(GM_E_ARG, "gm_coupling_pairs: need 0 <= lmax_window <= 2 lmax");
    if (npairs < 0) return fail(GM_E_ARG, "gm_coupling_pairs: negative pair count");
    if (npairs == 0) return GM_OK;
    const int L = (int)lmax + 1;
    const int bx = 128;
    hipLaunchKernelGGL(k_coupling_pairs, dim3((unsigned)((npairs + bx - 1) / bx)), dim3(bx), 0, (hipStream_t)stream,
                       window, (int)lmax_window, g, L, lo, hi, npairs, m00, m02, m22p, m22m);
    return check_launch("gm_coupling_pairs");
}
