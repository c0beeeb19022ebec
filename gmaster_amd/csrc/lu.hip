// Dense FP64 LU of the binned mode-coupling matrix and its triangular solves (the
// decoupling step: jnp.linalg.solve(mcm_binned, binned) in gmaster/workspaces.py:2106-2147).
//
// The binned coupling matrix is close to the identity for the masks MASTER is used with, so
// partial pivoting almost never exchanges rows.  gm_lu_factor therefore factors without
// pivoting (A = L U, L unit lower, in place) and raises a device flag when a pivot is zero
// or some |L_ij| > 1: that is exactly when partial pivoting would have chosen another row
// (at step j it takes the largest |A(j)_ij|, i >= j, the first one on ties, and
// L_ij = A(j)_ij / A(j)_jj), so without the flag the factors are the partially pivoted
// ones with identity permutation.  The caller uses a pivoting LU when the flag is set.
//
// Right-looking blocked algorithm, panels of kNB = 64 columns, three launches per panel:
//   k_lu_diag    the 64 x 64 diagonal block, factored by one workgroup (registers, pivot
//                row and column exchanged through LDS);
//   k_lu_panel   L21 = A21 U11^-1 (one thread per row) and U12 = L11^-1 A12 (one thread
//                per column), the 64 values of the row / column in registers;
//   k_lu_update  A22 -= L21 U12 in 64 x 64 tiles on v_mfma_f64_16x16x4f64 (operands
//                staged in LDS; each wave a 32 x 32 quarter, 2 x 2 MFMA blocks).
// gm_lu_solve: B <- U^-1 L^-1 B for row-major B (n x nrhs), one workgroup per group of
// kSolveR right-hand sides, 64-row blocks: the block's diagonal system in LDS, then the
// rows below (above) updated in parallel from the transposed factors (coalesced).
#include "common.h"

namespace gm {
namespace {

constexpr int kNB = 64;
constexpr int kSolveR = 4;
typedef double d4 __attribute__((ext_vector_type(4)));

__global__ void k_lu_clear(int* __restrict__ flag) {
    if (threadIdx.x == 0) *flag = 0;
}

// The diagonal block: thread (ti, tc) keeps elements (ti + 16 u, tc + 16 v) in registers.
// One barrier per step: step j reads the pivot, the raw column j and row j from buffer j & 1
// (published at the end of step j - 1), every thread forms l = a / pivot for its rows (the
// column's owner keeps it as L: the same quotient), updates, and the owners of step j + 1's
// pivot, column and row publish them into the other buffer.
// Step s's pivot, column (below the pivot) and row sit in register column / row V = s / 16 of
// the threads with tc = s % 16 / ti = s % 16; V is a template argument so that t stays in
// registers (a run-time index would move it to private memory).
template <int V>
__device__ __forceinline__ void lu_publish_v(const double (&t)[4][4], int s, int b, int ti, int tc, double* pv,
                                             double* col, double* row) {
    const int rs = s & 15;
    if (tc == rs) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = ti + 16 * u;
            if (i == s) *pv = t[u][V];
            else if (i > s && i < b) col[i] = t[u][V];
        }
    }
    if (ti == rs) {
#pragma unroll
        for (int v = 0; v < 4; ++v) row[tc + 16 * v] = t[V][v];
    }
}

// Steps j = 16 JB .. 16 JB + 15 (JB a template argument, so every register index is a
// constant); returns at j = b (uniform over the workgroup).
template <int JB>
__device__ __forceinline__ void lu_steps(double (&t)[4][4], bool& bad, int b, int ti, int tc, double (*colj)[kNB],
                                         double (*rowj)[kNB], double* pv) {
    for (int r = 0; r < 16; ++r) {
        const int j = 16 * JB + r;
        if (j >= b) return;
        const int q = j & 1;
        const double piv = pv[q];
        if (piv == 0.0) bad = true;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = ti + 16 * u;
            if (i <= j || i >= b) continue;
            const double lij = colj[q][i] / piv;  // column j: l = a / pivot below the diagonal
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = tc + 16 * v;
                if (c == j) {
                    t[u][v] = lij;
                    if (fabs(lij) > 1.0) bad = true;
                } else if (c > j && c < b) {
                    t[u][v] = fma(-lij, rowj[q][c], t[u][v]);
                }
            }
        }
        if (j + 1 < b) {
            if (r < 15) lu_publish_v<JB>(t, j + 1, b, ti, tc, &pv[q ^ 1], colj[q ^ 1], rowj[q ^ 1]);
            else if constexpr (JB < 3) lu_publish_v<JB + 1>(t, j + 1, b, ti, tc, &pv[q ^ 1], colj[q ^ 1], rowj[q ^ 1]);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_lu_diag(double* __restrict__ A, int n, int k0, int* __restrict__ flag) {
    __shared__ double colj[2][kNB], rowj[2][kNB], pv[2];
    const int b = min(kNB, n - k0);
    const int tid = threadIdx.x;
    const int ti = tid >> 4, tc = tid & 15;
    double t[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = ti + 16 * u, c = tc + 16 * v;
            t[u][v] = (i < b && c < b) ? A[(int64_t)(k0 + i) * n + k0 + c] : 0.0;
        }
    bool bad = false;
    lu_publish_v<0>(t, 0, b, ti, tc, &pv[0], colj[0], rowj[0]);
    __syncthreads();
    lu_steps<0>(t, bad, b, ti, tc, colj, rowj, pv);
    lu_steps<1>(t, bad, b, ti, tc, colj, rowj, pv);
    lu_steps<2>(t, bad, b, ti, tc, colj, rowj, pv);
    lu_steps<3>(t, bad, b, ti, tc, colj, rowj, pv);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = ti + 16 * u, c = tc + 16 * v;
            if (i < b && c < b) A[(int64_t)(k0 + i) * n + k0 + c] = t[u][v];
        }
    if (bad) atomicOr(flag, 1);
}

__global__ __launch_bounds__(256) void k_lu_panel(double* __restrict__ A, int n, int k0, int nrb,
                                                  int* __restrict__ flag) {
    __shared__ double T[kNB][kNB + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < kNB * kNB; e += 256) T[e / kNB][e % kNB] = A[(int64_t)(k0 + e / kNB) * n + k0 + e % kNB];
    __syncthreads();
    const int first = k0 + kNB;
    if ((int)blockIdx.x < nrb) {
        // row i of L21: x U11 = a, right-looking over the 64 columns
        const int i = first + blockIdx.x * 256 + tid;
        if (i >= n) return;
        double* row = A + (int64_t)i * n + k0;
        double x[kNB];
#pragma unroll
        for (int c = 0; c < kNB; ++c) x[c] = row[c];
        bool bad = false;
#pragma unroll
        for (int j = 0; j < kNB; ++j) {
            x[j] = x[j] / T[j][j];
            if (fabs(x[j]) > 1.0) bad = true;
#pragma unroll
            for (int c = j + 1; c < kNB; ++c) x[c] = fma(-x[j], T[j][c], x[c]);
        }
#pragma unroll
        for (int c = 0; c < kNB; ++c) row[c] = x[c];
        if (bad) atomicOr(flag, 1);
    } else {
        // column c of U12: L11 y = a (unit diagonal)
        const int c = first + (blockIdx.x - nrb) * 256 + tid;
        if (c >= n) return;
        double y[kNB];
#pragma unroll
        for (int r = 0; r < kNB; ++r) y[r] = A[(int64_t)(k0 + r) * n + c];
#pragma unroll
        for (int r = 0; r < kNB; ++r)
#pragma unroll
            for (int s = r + 1; s < kNB; ++s) y[s] = fma(-T[s][r], y[r], y[s]);
#pragma unroll
        for (int r = 0; r < kNB; ++r) A[(int64_t)(k0 + r) * n + c] = y[r];
    }
}

// A22 -= L21 U12 for the 64 x 64 tile (blockIdx.y, blockIdx.x) of the trailing matrix.
// MFMA 16x16x4 f64: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator
// register v holds D[4 v + l / 16][l % 16].
__global__ __launch_bounds__(256) void k_lu_update(double* __restrict__ A, int n, int k0) {
    __shared__ double As[kNB][kNB + 1];  // -L21 rows of the tile, columns k0 .. k0 + 63
    __shared__ double Bs[kNB][kNB + 1];  // U12 rows k0 .. k0 + 63, columns of the tile
    const int first = k0 + kNB;
    const int r0 = first + blockIdx.y * kNB, c0 = first + blockIdx.x * kNB;
    const int tid = threadIdx.x;
    for (int e = tid; e < kNB * kNB; e += 256) {
        const int r = e / kNB, c = e % kNB;
        As[r][c] = r0 + r < n ? -A[(int64_t)(r0 + r) * n + k0 + c] : 0.0;
        Bs[r][c] = c0 + c < n ? A[(int64_t)(k0 + r) * n + c0 + c] : 0.0;
    }
    __syncthreads();
    const int w = tid >> 6, l = tid & 63;
    const int wr = 32 * (w >> 1), wc = 32 * (w & 1);
    const int lr = l & 15, lk = l >> 4;
    d4 acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = r0 + wr + 16 * bi + 4 * v + lk, c = c0 + wc + 16 * bj + lr;
                acc[bi][bj][v] = (r < n && c < n) ? A[(int64_t)r * n + c] : 0.0;
            }
#pragma unroll 4
    for (int kk = 0; kk < kNB; kk += 4) {
        double a[2], bb[2];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) a[bi] = As[wr + 16 * bi + lr][kk + lk];
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) bb[bj] = Bs[kk + lk][wc + 16 * bj + lr];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
                acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], bb[bj], acc[bi][bj], 0, 0, 0);
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = r0 + wr + 16 * bi + 4 * v + lk, c = c0 + wc + 16 * bj + lr;
                if (r < n && c < n) A[(int64_t)r * n + c] = acc[bi][bj][v];
            }
}

// B <- L^-1 B then U^-1 B for the right-hand sides kSolveR * blockIdx.x .. + kSolveR - 1.
// Per 64-row block: wave q solves the block's triangular system for right-hand side q with
// lane s holding row s (one shuffle per row), then all threads update the remaining rows.
__global__ __launch_bounds__(1024) void k_lu_solve(const double* __restrict__ LU, const double* __restrict__ LUT, int n,
                                                   double* __restrict__ B, int nrhs) {
    __shared__ double T[kNB][kNB + 1];
    __shared__ double X[kNB][kSolveR];
    const int tid = threadIdx.x;
    const int w = tid >> 6, s = tid & 63;
    const int c0 = blockIdx.x * kSolveR;
    const int nr = min(kSolveR, nrhs - c0);
    const int nblk = (n + kNB - 1) / kNB;
    auto stage = [&](int j0, int b) {
        for (int e = tid; e < kNB * kNB; e += 1024) {
            const int r = e >> 6, c = e & 63;
            T[r][c] = (r < b && c < b) ? LU[(int64_t)(j0 + r) * n + j0 + c] : 0.0;
        }
    };
    // B[i] -= LU[i][j0 .. j0 + b) X for lo <= i < hi, reading the transposed factors so that
    // consecutive threads (rows) read consecutive words
    auto update = [&](int lo, int hi, int j0, int b) {
        for (int i = lo + tid; i < hi; i += 1024) {
            const double* col = LUT + (int64_t)j0 * n + i;
            double acc[kSolveR];
#pragma unroll
            for (int q = 0; q < kSolveR; ++q) acc[q] = 0.0;
            for (int c = 0; c < b; ++c) {
                const double v = col[(int64_t)c * n];
#pragma unroll
                for (int q = 0; q < kSolveR; ++q) acc[q] = fma(v, X[c][q], acc[q]);
            }
            for (int q = 0; q < nr; ++q) B[(int64_t)i * nrhs + c0 + q] -= acc[q];
        }
    };
    // forward: L y = b (unit lower)
    for (int kb = 0; kb < nblk; ++kb) {
        const int j0 = kb * kNB, b = min(kNB, n - j0);
        stage(j0, b);
        __syncthreads();
        if (w < nr) {
            double x = s < b ? B[(int64_t)(j0 + s) * nrhs + c0 + w] : 0.0;
            for (int r = 0; r < b; ++r) {
                const double xr = __shfl(x, r);
                if (s > r) x = fma(-T[s][r], xr, x);
            }
            X[s][w] = x;
            if (s < b) B[(int64_t)(j0 + s) * nrhs + c0 + w] = x;
        } else if (w < kSolveR) {
            X[s][w] = 0.0;
        }
        __syncthreads();
        update(j0 + b, n, j0, b);
        __syncthreads();
    }
    // backward: U x = y
    for (int kb = nblk - 1; kb >= 0; --kb) {
        const int j0 = kb * kNB, b = min(kNB, n - j0);
        stage(j0, b);
        __syncthreads();
        if (w < nr) {
            double x = s < b ? B[(int64_t)(j0 + s) * nrhs + c0 + w] : 0.0;
            for (int r = b - 1; r >= 0; --r) {
                if (s == r) x = x / T[r][r];
                const double xr = __shfl(x, r);
                if (s < r) x = fma(-T[s][r], xr, x);
            }
            X[s][w] = x;
            if (s < b) B[(int64_t)(j0 + s) * nrhs + c0 + w] = x;
        } else if (w < kSolveR) {
            X[s][w] = 0.0;
        }
        __syncthreads();
        update(0, j0, j0, b);
        __syncthreads();
    }
}

}  // namespace
}  // namespace gm

using namespace gm;

extern "C" int gm_lu_factor(double* A, int64_t n, int32_t* flag, void* stream) {
    if (!A || !flag) return fail(GM_E_ARG, "gm_lu_factor: null pointer");
    if (n < 1 || n > (1 << 20)) return fail(GM_E_ARG, "gm_lu_factor: bad size %lld", (long long)n);
    hipStream_t st = (hipStream_t)stream;
    const int nn = (int)n;
    hipLaunchKernelGGL(k_lu_clear, dim3(1), dim3(64), 0, st, flag);
    for (int k0 = 0; k0 < nn; k0 += kNB) {
        hipLaunchKernelGGL(k_lu_diag, dim3(1), dim3(256), 0, st, A, nn, k0, flag);
        const int m = nn - k0 - kNB;
        if (m <= 0) break;
        const int nrb = (m + 255) / 256;
        hipLaunchKernelGGL(k_lu_panel, dim3(2 * nrb), dim3(256), 0, st, A, nn, k0, nrb, flag);
        const int nt = (m + kNB - 1) / kNB;
        hipLaunchKernelGGL(k_lu_update, dim3(nt, nt), dim3(256), 0, st, A, nn, k0);
    }
    return check_launch("gm_lu_factor");
}

extern "C" int gm_lu_solve(const double* LU, const double* LUT, int64_t n, double* B, int64_t nrhs, void* stream) {
    if (!LU || !LUT || !B) return fail(GM_E_ARG, "gm_lu_solve: null pointer");
    if (n < 1 || n > (1 << 20) || nrhs < 1 || nrhs > (1 << 26))
        return fail(GM_E_ARG, "gm_lu_solve: bad sizes n=%lld nrhs=%lld", (long long)n, (long long)nrhs);
    hipLaunchKernelGGL(k_lu_solve, dim3((unsigned)((nrhs + kSolveR - 1) / kSolveR)), dim3(1024), 0,
                       (hipStream_t)stream, LU, LUT, (int)n, B, (int)nrhs);
    return check_launch("gm_lu_solve");
}
