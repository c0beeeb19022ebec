/*
 * gmaster_amd C-ABI: the drop-in boundary between the pymaster-compatible Python layer
 * (gmaster_amd/*.py) and the hand-written gfx950 kernels (gmaster_amd/csrc/*.hip).
 *
 * Conventions shared by every entry point
 *   - All pointers are device pointers (owned by the caller, e.g. torch tensors); no
 *     entry point allocates device memory unless stated.  Complex arrays are interleaved
 *     (re, im) doubles; "ld" arguments count complex elements per row.
 *   - `stream` is a hipStream_t (NULL = default stream).  Calls are asynchronous on it.
 *   - Return 0 on success or a GM_E_* code; gm_last_error() gives the message (per thread).
 *   - Every entry validates its sizes before launching, so the error paths are testable
 *     without a GPU.
 *
 * Array layouts
 *   ring spectrum, spin 0:  spec[map][ring][m],     m in [0, L),        ld >= L
 *   ring spectrum, spin s:  spec[ring][m + L - 1],  m in (-L, L),       ld >= 2L - 1
 *       "phased" spectra: spec[r][m] = (4 pi / npix) e^{-i m phi0_r} DFT_r[m mod nphi_r]
 *   alm:                    healpy m-major packing, nelem = (lmax+1)(lmax+2)/2, complex
 *   per-order tables:       order m owns l = m .. L-1+GM_TABLE_PAD at
 *                           base(m) = m (L + GM_TABLE_PAD) - m (m - 1) / 2
 *
 * The reference (licongxu/GMaster) crosses this boundary through XLA FFI custom calls;
 * each declaration cites the reference interface it replaces.
 */
#ifndef GMASTER_AMD_H
#define GMASTER_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GM_TABLE_PAD 8

enum {
    GM_OK = 0,
    GM_E_ARG = 1,         /* null pointer or invalid scalar argument */
    GM_E_SHAPE = 2,       /* inconsistent sizes / leading dimensions */
    GM_E_UNSUPPORTED = 3, /* valid request outside what this build supports */
    GM_E_LAUNCH = 4       /* HIP launch / runtime error */
};

/* Message of the last failing call on this thread ("" if none). */
const char* gm_last_error(void);
/* Version of this ABI (bumped on any signature change). */
int gm_abi_version(void);

/* Size in elements of one per-order table for band limit L = lmax + 1. */
int64_t gm_table_size(int64_t lmax);

/* ---------------------------------------------------------------- ring (azimuthal) stage */

/* Forward ring FFT of nmaps real HEALPix maps (RING order) into phased spin-0 spectra,
 * m in [0, lmax].  Replaces gmaster/_sht/rings.py:251-299 (_forward_ring_fft_positive)
 * together with the weight / phase multiply of gmaster/_sht/healpix.py:72-94. */
int gm_ring_fft_fwd_s0(const double* maps, int64_t nmaps, int64_t map_stride, int64_t nside,
                       int64_t lmax, double* spec, int64_t ld, int64_t spec_map_stride, void* stream);

/* Forward ring FFT of a spin field (Q, U) into the phased spectrum of Q + iU,
 * m in (-L, L).  Replaces gmaster/_sht/rings.py:628-685 (_forward_ring_fft_full). */
int gm_ring_fft_fwd_s(const double* q, const double* u, int64_t nside, int64_t lmax, double* spec,
                      int64_t ld, void* stream);

/* Inverse: maps[p] = sum_m F[r][m] e^{i m phi_p} (spin 0, Hermitian in m).
 * Replaces gmaster/_sht/rings.py:477-524. */
int gm_ring_fft_inv_s0(const double* spec, int64_t nmaps, int64_t ld, int64_t spec_map_stride,
                       int64_t nside, int64_t lmax, double* maps, int64_t map_stride, void* stream);

/* Inverse for spin fields: Q + iU = sum_m F+[r][m] e^{i m phi_p}.
 * Replaces gmaster/_sht/rings.py:702-771. */
int gm_ring_fft_inv_s(const double* spec, int64_t ld, int64_t nside, int64_t lmax, double* q,
                      double* u, void* stream);

/* Jacobi residual in ring-Fourier space, no FFT:
 *   R[r][k] = w nphi sum_{m = k mod nphi, |m| < L} e^{i (m-k) phi0} F[r][m] - G[r][k]
 * (spin 0: F[-m] = conj F[m]).  In-place allowed (R == G).
 * Replaces gm_ring_fold (gmaster/_native/cuda/march_v2.cu:426-470, sht/march_v2.py:614-625). */
int gm_ring_fold_residual(const double* F, const double* G, double* R, int64_t nmaps, int64_t ld,
                          int64_t map_stride, int64_t nside, int64_t lmax, void* stream);

/* Spin-s version on (-L, L) spectra.  Replaces gm_ring_fold_c (march_v2.cu:474-493). */
int gm_ring_fold_residual_c(const double* F, const double* G, double* R, int64_t ld,
                            int64_t nside, int64_t lmax, void* stream);

/* --------------------------------------------------------------- latitudinal (Legendre) */

/* Recurrence tables for spin s (0 <= s <= lmax): mu-rescaled coefficients alpha, gamma
 * (gamma may be NULL for s = 0) and output scale P, each gm_table_size(lmax) doubles;
 * seed_mant / seed_exp (lmax+1 entries, host-computed) are passed to the kernels below.
 * Replaces the window tables of gmaster/_sht/march_v2.py:388-450. */
int gm_lat_tables(int64_t lmax, int64_t spin, double* alpha, double* gamma, double* P, void* stream);

/* Spin-0 analysis of nmaps (1 or 2) phased spectra sharing one recurrence, one pass over
 * the ring chunks chunk0 .. chunk0 + nchunks - 1:
 * partial[c - chunk0][base(m) + l - m][map] = sum_{rings of chunk c} mu_lm(theta) (G_N +- G_S)
 * nchunks = gm_lat_nchunks(nside, 0) slots; chunk0 = 0, nchunks, 2 nchunks, ... below
 * gm_lat_ring_chunks (one pass up to Nside 4096; passes bound the partial buffer above).
 * Replaces gm_march_ana_s0 / gm_march_ana_s0_pair (march_v2.cu:155-285, :551-552). */
int gm_lat_ana_s0(const double* spec, int64_t nmaps, int64_t ld, int64_t spec_map_stride,
                  int64_t nside, int64_t lmax, const double* alpha, const double* seed_mant,
                  const int32_t* seed_exp, const int32_t* nactive, double* partial,
                  int64_t nchunks, int64_t chunk0, void* stream);

/* Sum the partial slots of the analysis pass that started at chunk0, scale by P and write
 * alm (mode 0: alm = result; 1: alm -= result, the Jacobi update; 2: alm += result, the
 * later passes of a multi-pass analysis). */
int gm_lat_finish_s0(const double* partial, int64_t nmaps, int64_t nside, const int32_t* nactive,
                     const double* P, int64_t lmax, double* alm, int64_t alm_map_stride, int mode,
                     int64_t chunk0, void* stream);

/* Synthesis coefficients: coef[base(m)+l-m][map] = alm[map][idx(l,m)] P_lm (zero padded);
 * nmaps 1, 2, 4 or 8 (4 and 8 feed the batched mixed synthesis). */
int gm_lat_prep_s0(const double* alm, int64_t nmaps, int64_t alm_map_stride, const double* P,
                   int64_t lmax, double* coef, void* stream);

/* Spin-0 synthesis: spec[map][r][m] = sum_l coef * mu_lm(theta_r).
 * Replaces gm_march_syn_s0 / _pair (march_v2.cu:291-407, :554-555). */
int gm_lat_syn_s0(const double* coef, int64_t nmaps, int64_t nside, int64_t lmax, const double* alpha,
                  const double* seed_mant, const int32_t* seed_exp, const int32_t* nactive,
                  double* spec, int64_t ld, int64_t spec_map_stride, void* stream);

/* Spin-s analysis over all rings (direct + mirror channel per order m >= 0), one pass over
 * ring chunks chunk0 .. chunk0 + nchunks - 1 as for spin 0.
 * Replaces gm_march_ana_s2 (march_v2.cu:553). */
int gm_lat_ana_s(const double* spec, int64_t ld, int64_t nside, int64_t lmax, int64_t spin,
                 const double* alpha, const double* gamma, const double* seed_mant,
                 const int32_t* seed_exp, const int32_t* nactive, double* partial, int64_t nchunks,
                 int64_t chunk0, void* stream);

/* Finish spin s: sum the pass's partials, scale, combine +-m into E/B (healpy packing,
 * 2 x nelem); modes as gm_lat_finish_s0.
 * Replaces gmaster/_sht/healpix.py:117-143 and :812-814. */
int gm_lat_finish_s(const double* partial, int64_t nside, const int32_t* nactive, const double* P,
                    int64_t lmax, int64_t spin, double* alm_e, double* alm_b, int mode, int64_t chunk0,
                    void* stream);

/* Synthesis coefficients for spin s from E/B (gmaster/_sht/healpix.py:599-611). */
int gm_lat_prep_s(const double* alm_e, const double* alm_b, const double* P, int64_t lmax,
                  int64_t spin, double* coef, void* stream);

/* Spin-s synthesis into spec[r][m + L - 1].  Replaces gm_march_syn_s2 (march_v2.cu:556). */
int gm_lat_syn_s(const double* coef, int64_t nside, int64_t lmax, int64_t spin, const double* alpha,
                 const double* gamma, const double* seed_mant, const int32_t* seed_exp,
                 const int32_t* nactive, double* spec, int64_t ld, void* stream);

/* Partial slots of one analysis pass (FP64 kernels), and the number of ring chunks the
 * analysis kernels split the rings of one order into (mixed != 0: the mixed kernels with
 * nmaps maps per recurrence); passes start at chunk0 = 0, slots, 2 slots, ... */
int64_t gm_lat_nchunks(int64_t nside, int64_t spin);
int64_t gm_lat_ring_chunks(int64_t nside, int64_t spin, int64_t mixed, int64_t nmaps);

/* ---------------------------------------------------------------- power spectra, MCM */

/* Coupled pseudo-C_ell of every (a, b) pair: cl[a*n2+b][l] = sum_m c_m Re(a conj b)/(2l+1).
 * work: >= gm_cl_work_size(lmax, n1*n2) doubles.  Replaces gmaster/workspaces.py:86-95. */
int64_t gm_cl_work_size(int64_t lmax, int64_t npairs);
int gm_coupled_cell(const double* alm1, int64_t n1, const double* alm2, int64_t n2, int64_t alm_stride,
                    int64_t lmax, double* cl, double* work, void* stream);

/* ---- batched latitudinal pass: K0 real spin-0 maps and K2 spin-2 fields sharing one
 * spin-0 recurrence lambda_lm per (m, ring pair), contraction on the matrix unit.
 * Supported (K0, K2): (2, 1), (1, 1), (0, 1), (0, 2), (2, 0), (4, 0), (8, 0).  Tables
 * (alpha, P, seeds) are the spin-0 ones of lmax + 2 (gm_lat_tables(lmax + 2, 0, ...)): the
 * spin-2 rows need lambda up to degree lmax + 1, the folded synthesis up to lmax + 2.  nactive[m] counts ring pairs (equator
 * first); pairs j >= jsplit (near the poles) contract in FP64, the others in FP32.
 * The kernels read whole blocks of degrees: alpha must be followed by 128 zero doubles and
 * the synthesis coefficients by 96 spare degrees (96 x gm_bat_stride(.., 1) doubles).
 * Replaces, for this batched form, the march analysis/synthesis FFI handlers
 * gm_march_ana_s0_pair / gm_march_ana_s2 / gm_march_syn_s0_pair / gm_march_syn_s2
 * (gmaster/_native/cuda/march_v2.cu:496-594), which run one recurrence per spin. */

/* Ring blocks of the FP32 (polar = 0) or FP64 (polar = 1) analysis launch; -1 if invalid. */
int64_t gm_bat_blocks(int64_t nside, int64_t jsplit, int64_t polar);
/* Channel stride of the analysis partial sums (synthesis = 0) or synthesis coefficients (1). */
int64_t gm_bat_stride(int64_t k0, int64_t k2, int64_t synthesis);
/* Analysis partial sums: spec0 (k0, nring, ld0 >= lmax+1) phased spectra of real maps,
 * spec2 (k2, nring, ld2 >= 2 lmax + 1) phased spectra F+ of Q + iU (column lmax + m);
 * part32: gm_bat_blocks(.., 0) x gm_table_size(lmax + 2) x stride floats; part64, the
 * polar blocks, gm_bat_blocks(.., 1) x T x stride doubles with T the table entries of the
 * orders m < mpol (every order with nactive[m] > jsplit). */
int gm_bat_ana(const double* spec0, int64_t k0, int64_t ld0, int64_t smap0, const double* spec2, int64_t k2,
               int64_t ld2, int64_t smap2, int64_t nside, int64_t lmax, const double* alpha, const double* seed_mant,
               const int32_t* seed_exp, const int32_t* nactive, int64_t jsplit, int64_t mpol, float* part32,
               double* part64, void* stream);
/* gm_bat_ana for residual maps (the Jacobi refinement passes): the FP32 ring blocks contract
 * with FP16 hi + lo operands on the matrix unit (each product to ~2^-21 relative, ~4x the
 * per-term error of gm_bat_ana, on an input that is a small fraction of the map); same
 * arguments, partials and polar FP64 blocks.  Replaces, in the refinement loop of the
 * batched form, the same march analysis handlers as gm_bat_ana. */
int gm_bat_ana_res(const double* spec0, int64_t k0, int64_t ld0, int64_t smap0, const double* spec2, int64_t k2,
                   int64_t ld2, int64_t smap2, int64_t nside, int64_t lmax, const double* alpha,
                   const double* seed_mant, const int32_t* seed_exp, const int32_t* nactive, int64_t jsplit,
                   int64_t mpol, float* part32, double* part64, void* stream);
/* Sums the partials, applies P_l and forms spin-0 alms (alm0 + k astride0) and E, B
 * (alm2 + (2 f + c) astride2, c = 0 for E, 1 for B); mode 0 writes, 1 subtracts, 2 adds. */
int gm_bat_finish(const float* part32, const double* part64, int64_t k0, int64_t k2, int64_t nside, int64_t lmax,
                  const int32_t* nactive, int64_t jsplit, int64_t mpol, const double* P, double* alm0,
                  int64_t astride0, double* alm2, int64_t astride2, int mode, void* stream);
/* Synthesis coefficients (gm_table_size(lmax + 2) x gm_bat_stride(.., 1) doubles). */
int gm_bat_prep(const double* alm0, int64_t k0, int64_t astride0, const double* alm2, int64_t k2, int64_t astride2,
                const double* P, int64_t lmax, double* coef, void* stream);
/* Synthesis of every ring pair into spec0 / spec2 (layouts as for gm_bat_ana). */
int gm_bat_syn(const double* coef, int64_t k0, int64_t k2, int64_t nside, int64_t lmax, const double* alpha,
               const double* seed_mant, const int32_t* seed_exp, const int32_t* nactive, int64_t jsplit,
               double* spec0, int64_t ld0, int64_t smap0, double* spec2, int64_t ld2, int64_t smap2, void* stream);

/* gm_coupled_cell without a caller workspace and without any allocation: the partial sums
 * use a 64 MiB device array of the library (gm_coupled_cell_noalloc_max_work() doubles per
 * GPU, held until the process ends), so lmax <= 8191 with 2 x 2 components (lmax <= 16383
 * with 1 x 1); larger requests fail with GM_E_SHAPE.  No counterpart in the reference: for
 * callers that cannot allocate device memory (graph capture, tight serving loops).  All
 * calls on one GPU share its array: calls that overlap on one GPU (other streams, other
 * threads) race and must be serialised by the caller. */
int64_t gm_coupled_cell_noalloc_max_work(void);
int gm_coupled_cell_noalloc(const double* alm1, int64_t n1, const double* alm2, int64_t n2, int64_t alm_stride,
                            int64_t lmax, double* cl, void* stream);

/* Mode-coupling blocks 0-0, 0-2, 2-2 even / odd in one pass over the Wigner-3j rows:
 *   M[l1][l2] = (2 l2 + 1) sum_l3 W'_l3 (l1 l2 l3; ...)(l1 l2 l3; ...)
 * window: the weighted mask pseudo-C_ell W'_l = (2l+1) W_l / (4 pi), lmax_window+1 doubles
 * (lmax_window <= 2 lmax; terms above it are zero); g: g(p) = Gamma(p+1/2)/(sqrt(pi) Gamma(p+1))
 * for p <= lmax.  Any of m00, m02, m22p, m22m may be NULL.  Replaces gm_coupling_tt
 * (gmaster/_native/cuda/coupling_tt.cu:18-81) and the quadrature of gmaster/workspaces.py:682-732. */
int gm_coupling(const double* window, int64_t lmax_window, const double* g, int64_t lmax, double* m00,
                double* m02, double* m22p, double* m22m, void* stream);

/* gm_coupling restricted to the pairs (lo[k], hi[k]), lo[k] <= hi[k] <= lmax (int32); writes
 * both (lo, hi) and (hi, lo) of each requested block and leaves every other element alone.
 * Replaces the exact-entry evaluation of gmaster/workspaces.py:542-590
 * (_coupling_matrix_tt_toeplitz over _toeplitz_exact_pairs, :520-539). */
int gm_coupling_pairs(const double* window, int64_t lmax_window, const double* g, int64_t lmax,
                      const int32_t* lo, const int32_t* hi, int64_t npairs, double* m00, double* m02,
                      double* m22p, double* m22m, void* stream);

/* gm_coupling for listed row segments: items (int32 triples, nitems of them) give (l1, l2_0,
 * nlive), the pairs (l1, l2_0 + i) for i < nlive <= 64, with 0 <= l1 <= l2_0 and
 * l2_0 + nlive <= lmax + 1 (not checked on the device); one wavefront per item, the same
 * recursion and window staging as gm_coupling.  Writes both (l1, l2) and (l2, l1) and
 * leaves every other element alone.  Replaces the exact-entry evaluation of
 * gmaster/workspaces.py:542-590 when the entries come in rows, as the Toeplitz fill's do. */
int gm_coupling_rows(const double* window, int64_t lmax_window, const double* g, int64_t lmax,
                     const int32_t* items, int64_t nitems, double* m00, double* m02, double* m22p,
                     double* m22m, void* stream);

/* Toeplitz approximation of one coupling block M ((lmax+1)^2, column factor included)
 * into out (not aliasing M): NaMaster's fill with l_toeplitz, l_exact, dl_band
 * (0 < l_exact <= l_toeplitz < lmax).  Replaces gmaster/workspaces.py:1141-1194
 * (_apply_toeplitz) as used at :1834-1840. */
int gm_toeplitz(const double* M, int64_t lmax, int64_t l_toeplitz, int64_t l_exact, int64_t dl_band,
                double* out, void* stream);

/* In-place LU of the row-major n x n matrix A (unit lower L below the diagonal, U on and
 * above) without row exchanges; *flag (device int32) is cleared and then set nonzero when a
 * pivot is zero or some |L_ij| > 1, i.e. when partial pivoting would have exchanged rows
 * (the factors are then unusable and a pivoting LU is needed).  Replaces the factorisation
 * inside jnp.linalg.solve(mcm_binned, ...) of gmaster/workspaces.py:2146 (decouple_cell). */
int gm_lu_factor(double* A, int64_t n, int32_t* flag, void* stream);

/* B <- U^-1 L^-1 B for the factors of gm_lu_factor and row-major B (n x nrhs), in place;
 * LUT is the transpose of LU (n x n, row-major), which the row updates read.  Replaces the
 * triangular solves of the same jnp.linalg.solve. */
int gm_lu_solve(const double* LU, const double* LUT, int64_t n, double* B, int64_t nrhs, void* stream);

/* Binned block: out[b1][b2] = sum_{l1 in b1, l2 in b2} wout[l1] M[l1][l2] tin[l2]
 * (CSR bands: ells sorted by band, band_start has nbands+1 entries).
 * Replaces gmaster/workspaces.py:292-310 (_BlockMCM.banded). */
int gm_bin_block(const double* M, int64_t lmax, const int32_t* ells, const int32_t* band_start,
                 int64_t nbands, const double* wout, const double* tin, double* out, void* stream);

/* Binned spectra: out[c][b] = sum_{l in b} w[l] cl[c][l].  Replaces gmaster/bins.py:15-20. */
int gm_bin_cell(const double* cl, int64_t ncl, int64_t ld, const int32_t* ells,
                const int32_t* band_start, int64_t nbands, const double* w, double* out, void* stream);

/* 128-bit content fingerprint of a device buffer (two 64-bit sums of position-keyed
 * splitmix64 word hashes, written to out[0..1]; data 8-byte aligned).  Keys the mask-alms
 * cache on the mask's bytes, so a new mask that lands at a freed mask's address never
 * reuses its alms.  No reference counterpart: the reference recomputes mask alms per field
 * (gmaster/field.py:273-289, 348-376). */
int gm_fingerprint(const void* data, int64_t nbytes, uint64_t* out, void* stream);

/* ---------------------------------------------------------------- contaminant deprojection */

/* Weighted inner products of flattened maps (each vector n doubles, e.g. nmaps x npix):
 *   out[i*nb + j] = w sum_p a_i[p] b_j[p] v[p mod period]   (v = weight_map; NULL for 1)
 * with a_i at a + i*lda, b_j at b + j*ldb.  work: >= gm_map_dots_work_size(na, nb, n)
 * doubles.  The sum is taken in a fixed order (deterministic).  Replaces the template
 * covariance and template-map products of gmaster/field.py:231-241 and dot_map
 * (gmaster/utils.py:256-263, weight 4 pi / npix). */
int64_t gm_map_dots_work_size(int64_t na, int64_t nb, int64_t n);
int gm_map_dots(const double* a, int64_t na, int64_t lda, const double* b, int64_t nb, int64_t ldb, int64_t n,
                const double* weight_map, int64_t period, double w, double* out, double* work, void* stream);

/* maps[p] -= sum_i alphas[i] templates[i*ldt + p] for p < n (alphas: ntemp device doubles).
 * Replaces gmaster/field.py:243-247. */
int gm_map_deproject(double* maps, int64_t n, const double* templates, int64_t ntemp, int64_t ldt,
                     const double* alphas, void* stream);

/* Product of two-component maps, out = conj(w) p per pixel:
 *   out[0] = w[0] p[0] + w[1] p[1],  out[1] = w[0] p[1] - w[1] p[0]
 * (w, p, out: 2 x npix; out may alias p).  The mask-derivative products of the E/B
 * purification, gmaster/field.py:458-462 and :473-476. */
int gm_spin_mul(const double* w, const double* p, int64_t npix, double* out, void* stream);

/* out[o] = sum_k cl[k][o][l] alm[k] per healpy-packed element (l, m): cl is nin x nout x
 * (lmax+1) doubles, alm nin and out nout complex arrays with the given strides (complex
 * elements); out must not alias alm.  Replaces _filter_alms (gmaster/workspaces.py:2258-2265). */
int gm_alm_filter(const double* alm, int64_t nin, int64_t alm_stride, const double* cl, int64_t nout,
                  int64_t lmax, double* out, int64_t out_stride, void* stream);

/* ---- pure-E/B coupling blocks ------------------------------------------------------------
 * The purified spin-2 coupling by Gauss-Legendre quadrature over Wigner d-functions
 * (gmaster/workspaces.py:928-1052).  Tables are degree-major: row l holds every node. */
/* Ascending nodes x and weights w of the order-point Gauss-Legendre rule on [-1, 1] (Newton
 * on P_order from the standard initial guesses).  Replaces _gauss_legendre
 * (gmaster/workspaces.py:593-606), which reads them from ducc0. */
int gm_gauss_legendre(int64_t order, double* x, double* w, void* stream);
/* out[l][k] = lscale[l] d^l_{mn}(beta_k), cos(beta_k) = x[k], l = 0..lmax (zero below
 * max(|m|, |n|)); lscale may be NULL (1); |m|, |n| <= 16.  Replaces _wigner_d_table
 * (gmaster/workspaces.py:609-679), same recurrence, sign and normalisation. */
int gm_wigner_d_table(const double* x, int64_t nnodes, int64_t m, int64_t n, int64_t lmax,
                      const double* lscale, double* out, void* stream);
/* c[k] = sum_l coef[l] d^l_{mn}(beta_k) and, when s is not NULL, s[k] = sum_l (-1)^l coef[l]
 * d^l_{mn}(beta_k), l <= lmax, without a table (|m|, |n| <= 16): the mask correlations of
 * gmaster/workspaces.py:973-985 and of the general coupling (:682-732, :1069-1122). */
int gm_wigner_d_corr(const double* x, int64_t nnodes, int64_t m, int64_t n, int64_t lmax,
                     const double* coef, double* c, double* s, void* stream);
/* out[l][k] = node_weight[k] sum_t coef[t] node_factors[t][k] degree_factors[t][l] tables[t][l][k]
 * for 1..8 terms (host arrays of device pointers; a NULL factor array or entry is 1, a NULL
 * node_weight is 1): the "left" tables of gmaster/workspaces.py:999-1037 with the quadrature
 * weight folded in. */
int gm_table_combine(int64_t nterms, const double* const* tables, const double* const* node_factors,
                     const double* const* degree_factors, const double* coef, const double* node_weight,
                     int64_t L, int64_t nnodes, double* out, void* stream);
/* even = (total + (-1)^(l1+l2) sgnd) / 2, odd = (total - (-1)^(l1+l2) sgnd) / 2 over L x L
 * (gmaster/workspaces.py:1038-1042). */
int gm_parity_split(const double* total, const double* sgnd, int64_t L, double* even, double* odd, void* stream);

/* ---- input side: mask apodization and Gaussian alms ----------------------------------------
 * Bytes of device workspace gm_mask_apodize needs at this nside (0 for an invalid nside). */
int64_t gm_apodize_work_size(int64_t nside);
/* C1 (kind 1) or C2 (kind 2) apodization of a RING-ordered mask over aposize_rad radians:
 * unmasked pixels within the scale of the nearest masked pixel (mask <= 0) are multiplied by
 * x - sin(2 pi x)/(2 pi) or (1 - cos(pi x))/2, x = sqrt((1 - cos d) / (1 - cos aposize));
 * kind 3 zeroes every pixel (masked ones included) whose centre lies closer than aposize_rad
 * to a masked pixel's centre, the other pixels keep their value.  out may alias mask.  Replaces the cKDTree search of mask_apodization
 * (gmaster/utils.py:61-147) with a ring search on the device. */
int gm_mask_apodize(const double* mask, int64_t nside, double aposize_rad, int64_t kind, double* out, void* work,
                    void* stream);
/* out (ncomp x nelem complex, healpy packing) = per-l covariance root (lmax+1 x ncomp x ncomp,
 * row-major) times complex Gaussian draws with unit variance per real part, scaled by
 * 1/sqrt(2) and with a zero imaginary part at m = 0 (ncomp <= 8).  The draws come from
 * Philox4x32-10 keyed by seed, one counter per (element, component).  Replaces
 * _gaussian_alms (gmaster/utils.py:806-820), whose JAX generator is not reproduced. */
int gm_gaussian_alms(const double* root, int64_t lmax, int64_t ncomp, uint64_t seed, double* out, void* stream);
/* The raw normal pairs of the same generator: out[2 (e ndraw + d) + {0, 1}] for elements
 * e < nelem and draws d < ndraw (nelem * ndraw <= 2^31). */
int gm_gauss_draws(uint64_t seed, int64_t nelem, int64_t ndraw, double* out, void* stream);

/* ---- mixed precision: FP32 contraction ---------------------------------------------------
 * Same arguments as the FP64 entries above.  The recurrence, its tables and seeds stay
 * FP64; each step's value is rounded to FP32 once and the contraction over rings and the
 * block reductions run in FP32 (spectra read and partial sums / spectra written in FP64),
 * ~1e-7 relative.  No counterpart in the reference, whose march is FP32 throughout
 * (gmaster/_native/cuda/march_v2.cu:155-407). */
/* Chunk counts and finishes of the mixed analyses (their ring chunks differ from FP64's,
 * and for spin 0 depend on the number of maps sharing the recurrence: 1, 2, 4 or 8). */
int64_t gm_lat_nchunks_mixed(int64_t nside, int64_t spin, int64_t nmaps);
int gm_lat_finish_s0_mixed(const double* partial, int64_t nmaps, int64_t nside, const int32_t* nactive,
                           const double* P, int64_t lmax, double* alm, int64_t alm_map_stride, int mode,
                           int64_t chunk0, void* stream);
int gm_lat_finish_s_mixed(const double* partial, int64_t nside, const int32_t* nactive, const double* P, int64_t lmax,
                          int64_t spin, double* alm_e, double* alm_b, int mode, int64_t chunk0, void* stream);
int gm_lat_ana_s0_mixed(const double* spec, int64_t nmaps, int64_t ld, int64_t smap, int64_t nside, int64_t lmax,
                        const double* alpha, const double* seed_mant, const int32_t* seed_exp, const int32_t* nactive,
                        double* partial, int64_t nchunks, int64_t chunk0, void* stream);
int gm_lat_syn_s0_mixed(const double* coef, int64_t nmaps, int64_t nside, int64_t lmax, const double* alpha,
                        const double* seed_mant, const int32_t* seed_exp, const int32_t* nactive, double* spec,
                        int64_t ld, int64_t smap, void* stream);
int gm_lat_ana_s_mixed(const double* spec, int64_t ld, int64_t nside, int64_t lmax, int64_t spin, const double* alpha,
                       const double* gamma, const double* seed_mant, const int32_t* seed_exp, const int32_t* nactive,
                       double* partial, int64_t nchunks, int64_t chunk0, void* stream);
int gm_lat_syn_s_mixed(const double* coef, int64_t nside, int64_t lmax, int64_t spin, const double* alpha,
                       const double* gamma, const double* seed_mant, const int32_t* seed_exp, const int32_t* nactive,
                       double* spec, int64_t ld, void* stream);

/* ---- learned coupling surrogate: dense network kernels --------------------------------
 * No counterpart in the reference, whose coupling is exact or Toeplitz only
 * (gmaster/workspaces.py:542-590, :1141-1194); these train and run the network that predicts
 * the Toeplitz residual.  bf16 = uint16 bit patterns, round-to-nearest-even. */

/* y = act(x w^T + b): x (batch, n_in) bf16, w (n_out, n_in) bf16, b (n_out) FP32 or NULL;
 * act 0 = identity, 1 = SiLU.  Writes any of z (pre-activation), y (FP32), y_bf16. */
int gm_mlp_fwd(const uint16_t* x, const uint16_t* w, const float* b, int64_t batch, int64_t n_in, int64_t n_out,
               int act, float* z, float* y, uint16_t* y_bf16, void* stream);

/* Backward of gm_mlp_fwd: dz = dy * act'(z) (FP32 and bf16 scratch, batch x n_out),
 * dw = dz^T x (n_out x n_in), db = column sums of dz, dx = dz w (FP32 and / or bf16, may be NULL). */
int gm_mlp_bwd(const float* dy, const float* z, int act, const uint16_t* x, const uint16_t* w, int64_t batch,
               int64_t n_in, int64_t n_out, float* dz, uint16_t* dz_bf16, float* dw, float* db, float* dx,
               uint16_t* dx_bf16, void* stream);

/* The four row sums of gm_coupling (00, 02, 22+, 22-; without the (2 l2 + 1) column factor)
 * of each pair (lo[k] <= hi[k] <= lmax) into vals[k][4].  Spot checks of an approximate
 * coupling matrix; no counterpart in the reference. */
int gm_coupling_pair_values(const double* window, int64_t lmax_window, const double* g, int64_t lmax,
                            const int32_t* lo, const int32_t* hi, int64_t npairs, double* vals, void* stream);

/* Learned correction of a Toeplitz-approximated block M (column factor included), in place:
 * every element not set exactly by the fill gets e(j, d) sqrt(D_j D_{j+d}) (2 col + 1), with
 * j = min(row, col), d = |row - col|, D_k = |M_kk| / (2k + 1), and e (n_l x n_d, FP32)
 * interpolated bilinearly in (log j, log d) on the strictly increasing int32 grids
 * l_grid[n_l], d_grid[n_d] (first entries >= 1), zero outside them. */
int gm_toeplitz_correct(double* M, int64_t lmax, int64_t l_toeplitz, int64_t l_exact, int64_t dl_band,
                        const int32_t* l_grid, int64_t n_l, const int32_t* d_grid, int64_t n_d, const float* e,
                        void* stream);

/* FP32 -> bf16 (round to nearest even). */
int gm_to_bf16(const float* x, int64_t n, uint16_t* y, void* stream);

/* AdamW step `step` (from 1) on FP32 master weights w, refreshing the bf16 copy if non-NULL. */
int gm_adamw(float* w, uint16_t* w_bf16, const float* g, float* m, float* v, int64_t n, double lr, double beta1,
             double beta2, double eps, double weight_decay, int64_t step, void* stream);

/* Weighted squared error: dy = 2 wt (y - t) / n_eff and sq = wt (y - t)^2 (sq may be NULL,
 * wt may be NULL for weight 1). */
int gm_mse_grad(const float* y, const float* t, const float* wt, int64_t n, double n_eff, float* dy, float* sq,
                void* stream);

#ifdef __cplusplus
}
#endif

#endif
