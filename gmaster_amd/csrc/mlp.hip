// Small dense network (the learned coupling surrogate) on gfx950: bf16 operands on the
// matrix cores with FP32 accumulation, FP32 master weights and AdamW.
//
// One tiled GEMM serves every product of training and inference,
//     C[M x N] = A[M x K] B[K x N],  A(i, k) = A[i*a_i + k*a_k],  B(k, j) = B[k*b_k + j*b_j],
// so forward (X W^T), weight gradient (dZ^T X) and input gradient (dZ W) differ only in
// strides.  A 256-thread workgroup computes a 64 x 64 tile; each wave owns 32 x 32 as 2 x 2
// v_mfma_f32_16x16x16_bf16 tiles.  Operand tiles are staged in LDS with k contiguous, so a
// lane's fragment (4 consecutive k) is one 64-bit LDS read.  Fragment layout of the
// 16x16x16 instruction, lane = 0..63:
//     A[i][k]: i = lane % 16, k = 4 (lane / 16) + e     (e = element of the 4-vector)
//     B[k][j]: j = lane % 16, k = 4 (lane / 16) + e
//     D[i][j]: j = lane % 16, i = 4 (lane / 16) + e
// The epilogue adds the bias and applies the activation (0 = identity, 1 = SiLU), storing
// the pre-activation Z (FP32), the output Y (FP32) and / or its bf16 copy.
#include "common.h"

namespace gm {
namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 64, TK = 32, PADK = TK + 4;

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// round to nearest even; NaN stays NaN
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float act_fwd(float z, int act) {
    if (act == 1) return z / (1.0f + expf(-z));
    return z;
}

__device__ __forceinline__ float act_grad(float z, int act) {
    if (act == 1) {
        const float s = 1.0f / (1.0f + expf(-z));
        return s * (1.0f + z * (1.0f - s));
    }
    return 1.0f;
}

__global__ __launch_bounds__(256) void k_gemm_bf16(const uint16_t* __restrict__ A, int64_t a_i, int64_t a_k,
                                                   const uint16_t* __restrict__ B, int64_t b_k, int64_t b_j, int M,
                                                   int N, int K, const float* __restrict__ bias, int act,
                                                   float* __restrict__ Z, float* __restrict__ Y,
                                                   uint16_t* __restrict__ Ybf, int64_t ldc) {
    __shared__ __attribute__((aligned(16))) uint16_t As[TM][PADK];  // [i][k]
    __shared__ __attribute__((aligned(16))) uint16_t Bs[TN][PADK];  // [j][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.y * TM, j0 = blockIdx.x * TN;
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
    float4_t acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = float4_t{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < K; k0 += TK) {
        // stage: 64 x 32 elements of each operand, 8 per thread, zero outside the matrix
        for (int e = tid; e < TM * TK; e += 256) {
            const int r = e / TK, k = e % TK;
            const int gi = i0 + r, gk = k0 + k;
            As[r][k] = (gi < M && gk < K) ? A[(int64_t)gi * a_i + (int64_t)gk * a_k] : (uint16_t)0;
            const int gj = j0 + r;
            Bs[r][k] = (gj < N && gk < K) ? B[(int64_t)gk * b_k + (int64_t)gj * b_j] : (uint16_t)0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; kk += 16) {
            const int kq = kk + 4 * (lane >> 4);
            short4_t fa[2], fb[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) fa[a] = *reinterpret_cast<const short4_t*>(&As[wi + 16 * a + (lane & 15)][kq]);
#pragma unroll
            for (int b = 0; b < 2; ++b) fb[b] = *reinterpret_cast<const short4_t*>(&Bs[wj + 16 * b + (lane & 15)][kq]);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[a], fb[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }

#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int j = j0 + wj + 16 * b + (lane & 15);
            if (j >= N) continue;
            const float bj = bias ? bias[j] : 0.0f;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = i0 + wi + 16 * a + 4 * (lane >> 4) + e;
                if (i >= M) continue;
                const float z = acc[a][b][e] + bj;
                const int64_t o = (int64_t)i * ldc + j;
                if (Z) Z[o] = z;
                const float y = act_fwd(z, act);
                if (Y) Y[o] = y;
                if (Ybf) Ybf[o] = f32_to_bf16(y);
            }
        }
}

// dZ = dY * act'(Z) in FP32 and bf16
__global__ void k_act_back(const float* __restrict__ dy, const float* __restrict__ z, int act, int64_t n,
                           float* __restrict__ dz, uint16_t* __restrict__ dzbf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = dy[i] * (z ? act_grad(z[i], act) : 1.0f);
    dz[i] = g;
    dzbf[i] = f32_to_bf16(g);
}

// db[o] = sum_b dZ[b][o]: one thread per column, rows in order
__global__ void k_colsum(const float* __restrict__ dz, int rows, int cols, float* __restrict__ db) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= cols) return;
    float s = 0.0f;
    for (int b = 0; b < rows; ++b) s += dz[(int64_t)b * cols + o];
    db[o] = s;
}

__global__ void k_to_bf16(const float* __restrict__ x, int64_t n, uint16_t* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f32_to_bf16(x[i]);
}

// AdamW on FP32 master weights, with the bias-corrected step folded into lr_t
__global__ void k_adamw(float* __restrict__ w, uint16_t* __restrict__ wbf, const float* __restrict__ g,
                        float* __restrict__ m, float* __restrict__ v, int64_t n, float lr_t, float b1, float b2,
                        float eps_t, float decay) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] * (1.0f - decay) - lr_t * mi / (sqrtf(vi) + eps_t);
    w[i] = wi;
    if (wbf) wbf[i] = f32_to_bf16(wi);
}

// weighted squared error: dy = 2 wt (y - t) / n_eff, sq = wt (y - t)^2
__global__ void k_mse_grad(const float* __restrict__ y, const float* __restrict__ t, const float* __restrict__ wt,
                           int64_t n, float inv_n, float* __restrict__ dy, float* __restrict__ sq) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float w = wt ? wt[i] : 1.0f;
    const float d = y[i] - t[i];
    dy[i] = 2.0f * w * d * inv_n;
    if (sq) sq[i] = w * d * d;
}

int gemm(const uint16_t* A, int64_t a_i, int64_t a_k, const uint16_t* B, int64_t b_k, int64_t b_j, int64_t M,
         int64_t N, int64_t K, const float* bias, int act, float* Z, float* Y, uint16_t* Ybf, hipStream_t st) {
    dim3 grid((unsigned)((N + TN - 1) / TN), (unsigned)((M + TM - 1) / TM));
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, A, a_i, a_k, B, b_k, b_j, (int)M, (int)N, (int)K, bias,
                       act, Z, Y, Ybf, N);
    return GM_OK;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace gm

using namespace gm;

extern "C" int gm_mlp_fwd(const uint16_t* x, const uint16_t* w, const float* b, int64_t batch, int64_t n_in,
                          int64_t n_out, int act, float* z, float* y, uint16_t* y_bf16, void* stream) {
    if (!x || !w) return fail(GM_E_ARG, "gm_mlp_fwd: null input / weight");
    if (!z && !y && !y_bf16) return fail(GM_E_ARG, "gm_mlp_fwd: no output requested");
    if (batch < 1 || n_in < 1 || n_out < 1 || batch > (1 << 24) || n_in > (1 << 20) || n_out > (1 << 20))
        return fail(GM_E_ARG, "gm_mlp_fwd: bad sizes");
    if (act != 0 && act != 1) return fail(GM_E_ARG, "gm_mlp_fwd: act must be 0 (identity) or 1 (SiLU)");
    // Y[b][o] = sum_c X[b][c] W[o][c] + bias[o]
    gemm(x, n_in, 1, w, 1, n_in, batch, n_out, n_in, b, act, z, y, y_bf16, (hipStream_t)stream);
    return check_launch("gm_mlp_fwd");
}

extern "C" int gm_mlp_bwd(const float* dy, const float* z, int act, const uint16_t* x, const uint16_t* w,
                          int64_t batch, int64_t n_in, int64_t n_out, float* dz, uint16_t* dz_bf16, float* dw,
                          float* db, float* dx, uint16_t* dx_bf16, void* stream) {
    if (!dy || !x || !w || !dz || !dz_bf16 || !dw || !db) return fail(GM_E_ARG, "gm_mlp_bwd: null pointer");
    if (batch < 1 || n_in < 1 || n_out < 1 || batch > (1 << 24) || n_in > (1 << 20) || n_out > (1 << 20))
        return fail(GM_E_ARG, "gm_mlp_bwd: bad sizes");
    if (act != 0 && act != 1) return fail(GM_E_ARG, "gm_mlp_bwd: act must be 0 (identity) or 1 (SiLU)");
    if (act == 1 && !z) return fail(GM_E_ARG, "gm_mlp_bwd: SiLU needs the pre-activation z");
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = batch * n_out;
    hipLaunchKernelGGL(k_act_back, dim3(blocks_for(n)), dim3(256), 0, st, dy, act ? z : nullptr, act, n, dz, dz_bf16);
    hipLaunchKernelGGL(k_colsum, dim3(blocks_for(n_out)), dim3(256), 0, st, dz, (int)batch, (int)n_out, db);
    // dW[o][c] = sum_b dZ[b][o] X[b][c]
    gemm(dz_bf16, 1, n_out, x, n_in, 1, n_out, n_in, batch, nullptr, 0, nullptr, dw, nullptr, st);
    // dX[b][c] = sum_o dZ[b][o] W[o][c]
    if (dx || dx_bf16) gemm(dz_bf16, n_out, 1, w, n_in, 1, batch, n_in, n_out, nullptr, 0, nullptr, dx, dx_bf16, st);
    return check_launch("gm_mlp_bwd");
}

extern "C" int gm_to_bf16(const float* x, int64_t n, uint16_t* y, void* stream) {
    if (!x || !y || n < 0) return fail(GM_E_ARG, "gm_to_bf16: bad arguments");
    if (n == 0) return GM_OK;
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, x, n, y);
    return check_launch("gm_to_bf16");
}

extern "C" int gm_adamw(float* w, uint16_t* w_bf16, const float* g, float* m, float* v, int64_t n, double lr,
                        double beta1, double beta2, double eps, double weight_decay, int64_t step, void* stream) {
    if (!w || !g || !m || !v || n < 0) return fail(GM_E_ARG, "gm_adamw: bad arguments");
    if (step < 1) return fail(GM_E_ARG, "gm_adamw: step counts from 1");
    if (!(beta1 >= 0 && beta1 < 1 && beta2 >= 0 && beta2 < 1 && eps > 0 && lr >= 0))
        return fail(GM_E_ARG, "gm_adamw: bad hyper-parameters");
    if (n == 0) return GM_OK;
    const double c1 = 1.0 - pow(beta1, (double)step), c2 = 1.0 - pow(beta2, (double)step);
    // m_hat / (sqrt(v_hat) + eps) = (m / c1) / (sqrt(v) / sqrt(c2) + eps)
    const float lr_t = (float)(lr * sqrt(c2) / c1), eps_t = (float)(eps * sqrt(c2));
    hipLaunchKernelGGL(k_adamw, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, w, w_bf16, g, m, v, n, lr_t,
                       (float)beta1, (float)beta2, eps_t, (float)(lr * weight_decay));
    return check_launch("gm_adamw");
}

extern "C" int gm_mse_grad(const float* y, const float* t, const float* wt, int64_t n, double n_eff, float* dy,
                           float* sq, void* stream) {
    if (!y || !t || !dy || n < 0 || !(n_eff > 0)) return fail(GM_E_ARG, "gm_mse_grad: bad arguments");
    if (n == 0) return GM_OK;
    hipLaunchKernelGGL(k_mse_grad, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, y, t, wt, n,
                       (float)(1.0 / n_eff), dy, sq);
    return check_launch("gm_mse_grad");
}
