// Power spectra and binning on gfx950, FP64: coupled pseudo-C_ell of healpy-packed alms
// (HBM-bound segmented reduction over m), binned spectra and binned coupling blocks.
#include <mutex>

#include "common.h"

namespace gm {
namespace {

constexpr int kClMChunk = 32;  // orders per partial in the coupled-C_ell reduction

// partial[mc][pair][l] = sum_{m in chunk mc, m <= l} c_m Re(a_lm conj b_lm)
__global__ void k_cl_partial(const double2* __restrict__ a, int n1, const double2* __restrict__ b, int n2,
                             int64_t stride, int L, double* __restrict__ part) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int mc = blockIdx.y;
    const int m0 = mc * kClMChunk;
    if (l >= L || m0 > l) return;
    const int mend = min(m0 + kClMChunk, l + 1);
    const int lmax = L - 1;
    const int np = n1 * n2;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int m = m0; m < mend; ++m) {
        const int64_t ia = alm_index(l, m, lmax);
        const double w = m ? 2.0 : 1.0;
        for (int i = 0; i < n1; ++i) {
            const double2 x = a[i * stride + ia];
            for (int j = 0; j < n2; ++j) {
                const double2 y = b[j * stride + ia];
                acc[i * n2 + j] += w * (x.x * y.x + x.y * y.y);
            }
        }
    }
    for (int p = 0; p < np; ++p) part[((int64_t)mc * np + p) * L + l] = acc[p];
}

__global__ void k_cl_final(const double* __restrict__ part, int np, int L, double* __restrict__ cl) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (l >= L) return;
    const int nmc = l / kClMChunk + 1;
    double s = 0.0;
    for (int mc = 0; mc < nmc; ++mc) s += part[((int64_t)mc * np + p) * L + l];
    cl[(int64_t)p * L + l] = s / (2.0 * l + 1.0);
}

__global__ void k_bin_cell(const double* __restrict__ cl, int64_t ld, const int* __restrict__ ells,
                           const int* __restrict__ bstart, int nb, const double* __restrict__ w,
                           double* __restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (b >= nb) return;
    double s = 0.0;
    for (int e = bstart[b]; e < bstart[b + 1]; ++e) s += w[e] * cl[c * ld + ells[e]];
    out[(int64_t)c * nb + b] = s;
}

// out[b1][b2] = sum_{e1 in b1, e2 in b2} wout[e1] M[ells[e1]][ells[e2]] tin[e2]
__global__ void k_bin_block(const double* __restrict__ M, int L, const int* __restrict__ ells,
                            const int* __restrict__ bstart, int nb, const double* __restrict__ wout,
                            const double* __restrict__ tin, double* __restrict__ out) {
    const int b2 = blockIdx.x * blockDim.x + threadIdx.x;
    const int b1 = blockIdx.y;
    if (b2 >= nb) return;
    double s = 0.0;
    for (int e1 = bstart[b1]; e1 < bstart[b1 + 1]; ++e1) {
        const double* row = M + (int64_t)ells[e1] * L;
        double r = 0.0;
        for (int e2 = bstart[b2]; e2 < bstart[b2 + 1]; ++e2) r += row[ells[e2]] * tin[e2];
        s += wout[e1] * r;
    }
    out[(int64_t)b1 * nb + b2] = s;
}

}  // namespace
}  // namespace gm

using namespace gm;

extern "C" int64_t gm_cl_work_size(int64_t lmax, int64_t npairs) {
    const int64_t L = lmax + 1;
    return ((L + kClMChunk - 1) / kClMChunk) * npairs * L;
}

extern "C" int gm_coupled_cell(const double* alm1, int64_t n1, const double* alm2, int64_t n2, int64_t alm_stride,
                               int64_t lmax, double* cl, double* work, void* stream) {
    if (!alm1 || !alm2 || !cl || !work) return fail(GM_E_ARG, "gm_coupled_cell: null pointer");
    if (lmax < 0) return fail(GM_E_ARG, "gm_coupled_cell: lmax < 0");
    if (n1 < 1 || n1 > 2 || n2 < 1 || n2 > 2) return fail(GM_E_ARG, "gm_coupled_cell: 1 or 2 components per field");
    const int64_t L = lmax + 1, nelem = L * (L + 1) / 2;
    if ((n1 > 1 || n2 > 1) && alm_stride < nelem) return fail(GM_E_SHAPE, "gm_coupled_cell: alm stride too small");
    const int nmc = (int)((L + kClMChunk - 1) / kClMChunk);
    const int np = (int)(n1 * n2);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_cl_partial, dim3((unsigned)((L + 255) / 256), nmc), dim3(256), 0, st,
                       reinterpret_cast<const double2*>(alm1), (int)n1, reinterpret_cast<const double2*>(alm2),
                       (int)n2, alm_stride, (int)L, work);
    hipLaunchKernelGGL(k_cl_final, dim3((unsigned)((L + 255) / 256), np), dim3(256), 0, st, work, np, (int)L, cl);
    return check_launch("gm_coupled_cell");
}

// ------------------------------------------------------------------ no-allocation variant
// The partial sums of gm_coupled_cell live in a file-scope device array instead of a
// caller's workspace.  Each GPU holds its own copy for the life of the process; host code
// uses the per-device address returned by hipGetSymbolAddress, never the symbol itself.
namespace gm {
namespace {
constexpr int64_t kClScratchDoubles = int64_t(1) << 23;  // 64 MiB: lmax <= 8191 for 4 pairs
__device__ __align__(256) double scratch[kClScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Address of the current device's scratch array (resolved once per device).
int scratch_for_current_device(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(GM_E_LAUNCH, "gm_coupled_cell_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(GM_E_ARG, "gm_coupled_cell_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(GM_E_LAUNCH, "gm_coupled_cell_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return GM_OK;
}
}  // namespace
}  // namespace gm

extern "C" int64_t gm_coupled_cell_noalloc_max_work(void) { return kClScratchDoubles; }

extern "C" int gm_coupled_cell_noalloc(const double* alm1, int64_t n1, const double* alm2, int64_t n2,
                                       int64_t alm_stride, int64_t lmax, double* cl, void* stream) {
    if (!alm1 || !alm2 || !cl) return fail(GM_E_ARG, "gm_coupled_cell_noalloc: null pointer");
    if (lmax < 0) return fail(GM_E_ARG, "gm_coupled_cell_noalloc: lmax < 0");
    if (n1 < 1 || n1 > 2 || n2 < 1 || n2 > 2)
        return fail(GM_E_ARG, "gm_coupled_cell_noalloc: 1 or 2 components per field");
    const int64_t need = gm_cl_work_size(lmax, n1 * n2);
    if (need > kClScratchDoubles)
        return fail(GM_E_SHAPE,
                    "gm_coupled_cell_noalloc: lmax %lld with %lld pairs needs %lld scratch doubles, the maximum is "
                    "%lld (use gm_coupled_cell with a workspace)",
                    (long long)lmax, (long long)(n1 * n2), (long long)need, (long long)kClScratchDoubles);
    double* work = nullptr;
    const int rc = scratch_for_current_device(&work);
    if (rc != GM_OK) return rc;
    // every partial the final pass reads is written by the partial pass: nothing to clear
    return gm_coupled_cell(alm1, n1, alm2, n2, alm_stride, lmax, cl, work, stream);
}

extern "C" int gm_bin_cell(const double* cl, int64_t ncl, int64_t ld, const int32_t* ells, const int32_t* band_start,
                           int64_t nbands, const double* w, double* out, void* stream) {
    if (!cl || !ells || !band_start || !w || !out) return fail(GM_E_ARG, "gm_bin_cell: null pointer");
    if (ncl < 1 || nbands < 1) return fail(GM_E_ARG, "gm_bin_cell: empty input");
    hipLaunchKernelGGL(k_bin_cell, dim3((unsigned)((nbands + 127) / 128), (unsigned)ncl), dim3(128), 0,
                       (hipStream_t)stream, cl, ld, ells, band_start, (int)nbands, w, out);
    return check_launch("gm_bin_cell");
}

extern "C" int gm_bin_block(const double* M, int64_t lmax, const int32_t* ells, const int32_t* band_start,
                            int64_t nbands, const double* wout, const double* tin, double* out, void* stream) {
    if (!M || !ells || !band_start || !wout || !tin || !out) return fail(GM_E_ARG, "gm_bin_block: null pointer");
    if (lmax < 0 || nbands < 1) return fail(GM_E_ARG, "gm_bin_block: bad sizes");
    hipLaunchKernelGGL(k_bin_block, dim3((unsigned)((nbands + 63) / 64), (unsigned)nbands), dim3(64), 0,
                       (hipStream_t)stream, M, (int)(lmax + 1), ells, band_start, (int)nbands, wout, tin, out);
    return check_launch("gm_bin_block");
}

// ------------------------------------------------------------------ content fingerprint
namespace gm {
namespace {

// splitmix64 finaliser (Steele, Lea & Flood 2014): a bijective 64-bit mixer.
__device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ void k_fingerprint_clear(unsigned long long* out) {
    if (threadIdx.x < 2) out[threadIdx.x] = 0ull;
}

// Two order-independent 64-bit sums of position-keyed word hashes; any change of one word
// changes both sums unless two independent 64-bit hashes collide.
__global__ void k_fingerprint(const uint64_t* __restrict__ w, int64_t nwords, const unsigned char* __restrict__ tail,
                              int ntail, unsigned long long* out) {
    uint64_t s1 = 0, s2 = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += stride) {
        const uint64_t h = mix64(w[i] ^ mix64((uint64_t)i + 0x9e3779b97f4a7c15ull));
        s1 += h;
        s2 += mix64(h ^ 0x632be59bd9b4e019ull);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && ntail > 0) {
        uint64_t t = 0;
        for (int k = 0; k < ntail; ++k) t |= (uint64_t)tail[k] << (8 * k);
        const uint64_t h = mix64(t ^ mix64((uint64_t)nwords + 0x9e3779b97f4a7c15ull) ^ 0xff51afd7ed558ccdull);
        s1 += h;
        s2 += mix64(h ^ 0x632be59bd9b4e019ull);
    }
    for (int off = 32; off > 0; off >>= 1) {
        s1 += __shfl_down(s1, off);
        s2 += __shfl_down(s2, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&out[0], (unsigned long long)s1);
        atomicAdd(&out[1], (unsigned long long)s2);
    }
}

}  // namespace
}  // namespace gm

extern "C" int gm_fingerprint(const void* data, int64_t nbytes, uint64_t* out, void* stream) {
    if (!out || (nbytes > 0 && !data)) return fail(GM_E_ARG, "gm_fingerprint: null pointer");
    if (nbytes < 0) return fail(GM_E_ARG, "gm_fingerprint: nbytes < 0");
    if (reinterpret_cast<uintptr_t>(data) & 7) return fail(GM_E_ARG, "gm_fingerprint: data not 8-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    auto* o = reinterpret_cast<unsigned long long*>(out);
    hipLaunchKernelGGL(k_fingerprint_clear, dim3(1), dim3(64), 0, st, o);
    const int64_t nwords = nbytes / 8;
    const int ntail = (int)(nbytes - nwords * 8);
    int64_t blocks = (nwords + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fingerprint, dim3((unsigned)blocks), dim3(256), 0, st,
                       reinterpret_cast<const uint64_t*>(data), nwords,
                       reinterpret_cast<const unsigned char*>(data) + nwords * 8, ntail, o);
    return check_launch("gm_fingerprint");
}
