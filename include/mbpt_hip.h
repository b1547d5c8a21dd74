/*
 * mbpt_hip.h -- C ABI of libmbpt_hip.so, the gfx950 (MI355X) kernels of the
 * ISDF BSE sigma-vector and the space-time polarizability chi0(i tau).
 *
 * Conventions
 *   - Matrices are row-major, FP64 at the boundary.  The library converts to
 *     its working precision (MBPT_FP64, MBPT_FP32 or MBPT_BF16X3) into buffers
 *     a handle owns.
 *   - Pointers named *_dev are device pointers (HBM), *_host host pointers.
 *   - `stream` is a hipStream_t passed as void* (NULL: the null stream).
 *     Every call is stream-ordered and asynchronous unless stated.
 *   - Pair layout z[k][i][a]: occupied index slow, virtual fast.
 *   - Every function returns MBPT_OK or an error code; mbpt_last_error()
 *     gives the message of the calling thread's last failure.
 *   - A handle is not re-entrant: one handle per stream.
 */
#ifndef MBPT_HIP_H
#define MBPT_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* precision of the bulk GEMMs */
#define MBPT_FP64 0   /* FP64 MFMA (v_mfma_f64_16x16x4f64)                     */
#define MBPT_FP32 1   /* FP32 MFMA (v_mfma_f32_16x16x4f32), FP32 accumulation */
#define MBPT_BF16X3 2 /* split BF16 hi+lo, 3 BF16 MFMAs per product, FP32 acc   */

/* status codes */
#define MBPT_OK 0
#define MBPT_ERR_VALUE 1       /* bad shape / argument: the reference raises ValueError */
#define MBPT_ERR_HIP 2         /* a HIP runtime call failed                             */
#define MBPT_ERR_UNSUPPORTED 3 /* a valid request this build does not implement         */

const char *mbpt_last_error(void);
int mbpt_abi_version(void);

/* ------------------------------------------------------------------------
 * Sigma-vector of the Tamm-Dancoff ISDF BSE/Casida operator
 *
 *   A z = d o z + kappa X_o^T (u o X_v) - (X_o^T (W o (X_o z X_v^T))) X_v,
 *   u = V p,  p_mu = sum_j X_o[mu, j] (z X_v^T)[j, mu],  d_ia = eps_a - eps_i
 *
 * on the interpolation grid (W = D W_aux D^T, V = D D^T of the reference's
 * factors).  Replaces the closure `apply_AB` returned by
 * `isdf_block_action` (src/SingleReference/LinearResponse/block_actions.py:
 * 365, the serial tda=True branch :563-661).
 * ------------------------------------------------------------------------ */
typedef struct mbpt_sigma mbpt_sigma;

/* Build a handle: the action's set-up, block_actions.py:365-543.
 *   M, nocc, nvir  grid points, occupied and virtual orbitals (nocc <= 128)
 *   precision      MBPT_FP64 | MBPT_FP32 | MBPT_BF16X3
 *   kappa          Hartree prefactor, KAPPA[spin] (2 singlet, 0 triplet)
 *   X_dev          (M, nocc + nvir) collocation, occupied columns first
 *   W_dev          (M, M) screened interaction on the grid, symmetric
 *   V_dev          (M, M) bare Coulomb on the grid (unread when kappa == 0)
 *   eps_host       (nocc + nvir) orbital energies
 * The inputs are copied; the caller may free them after the call returns
 * and the stream has reached it. */
int mbpt_sigma_create(mbpt_sigma **out, int M, int nocc, int nvir,
                      int precision, double kappa, const double *X_dev,
                      const double *W_dev, const double *V_dev,
                      const double *eps_host, void *stream);

/* Az = A z for k trial vectors, z_dev and Az_dev (k, nocc, nvir) FP64,
 * distinct buffers: `apply_AB(z)[0]`, block_actions.py:563-661. */
int mbpt_sigma_apply(mbpt_sigma *h, int k, const double *z_dev,
                     double *Az_dev, void *stream);

/* (Az, Bz) = (A z, B z) of the full Casida action for k trial vectors, all
 * (k, nocc, nvir) FP64 and distinct: `apply_AB(z)` with tda=False,
 * block_actions.py:563-661 (the B block, kappa (ia|jb) - (ib|W|ja), from
 * U = (W o P) X_o and U^T X_v, :632-652), in the handle's precision. */
int mbpt_sigma_apply_ab(mbpt_sigma *h, int k, const double *z_dev, double *Az_dev,
                        double *Bz_dev, void *stream);

/* out = d - (ii|W|aa), (nocc, nvir) FP64: `screened_diagonal()`,
 * block_actions.py:663-689 (grid gauge: (X_o o X_o)^T W (X_v o X_v)). */
int mbpt_sigma_screened_diag(mbpt_sigma *h, double *out_dev, void *stream);

/* Device bytes the handle holds (`apply_AB.held_bytes`, block_actions.py:705). */
size_t mbpt_sigma_held_bytes(const mbpt_sigma *h);

/* Per-apply workspace in bytes that an apply of k vectors needs (with_b != 0:
 * mbpt_sigma_apply_ab); the workspace grows to it on first use and is kept.
 * The counterpart of the reference's tile budget, the S/T/Tb/U row tiles
 * sized by `tile_memory_gb` (block_actions.py:526-543): a caller bounds the
 * workspace by applying at most the k this allows per call.  Host only. */
size_t mbpt_sigma_workspace_bytes(const mbpt_sigma *h, int k, int with_b);

/* Per-stage device timing, the counterpart of the action's `piece_clock`
 * (block_actions.py:56-70).  enable != 0 starts (and zeroes) accumulation:
 * every later apply records HIP events around its four launches on its
 * stream without waiting for them; mbpt_sigma_profile_read waits for the
 * last timed apply and adds up every one since.  ms_out[4] receives the accumulated
 * milliseconds of head, hartree, exchange and tail; launches the number of
 * applies timed; exchange_flops the algorithmic flops of the exchange
 * launches, sum of k (4 nocc M^2 + M^2). */
int mbpt_sigma_profile(mbpt_sigma *h, int enable);
int mbpt_sigma_profile_read(mbpt_sigma *h, double *ms_out, long *launches,
                            double *exchange_flops);

int mbpt_sigma_destroy(mbpt_sigma *h);

/* ------------------------------------------------------------------------
 * Space-time polarizability on the grid
 *
 *   out = [accumulate ? out : 0] + sum_t weight[t] * (-2) Go(t) o Gv(t),
 *   Go(t) = X_o e^{(eps_o - mu) tau_t} X_o^T,  Gv(t) = X_v e^{-(eps_v - mu) tau_t} X_v^T
 *
 * weight = e_t gives proj(tau_t) (without the D projection, which is the
 * identity in the grid gauge); weight = cosft_wt[k, :] gives chi0(i w_k).
 * Replaces `polarizability_projected_tau` (src/SingleReference/
 * LinearResponse/space_time.py:478) and the tau sums of
 * `chi0_tau_terms` / `static_screening_matrix` (:634, :1003).
 * ------------------------------------------------------------------------ */
typedef struct mbpt_chi0 mbpt_chi0;

#define MBPT_CHI0_MAX_TAU 64

/* X_dev (M, nocc + nvir) device FP64, eps_host (nocc + nvir), mu the shift
 * of `split_branches` (space_time.py:448) -- any mu inside the gap. */
int mbpt_chi0_create(mbpt_chi0 **out, int M, int nocc, int nvir,
                     int precision, const double *X_dev,
                     const double *eps_host, double mu, void *stream);

/* ntau <= MBPT_CHI0_MAX_TAU time points tau_host and weights weight_host;
 * out_dev (M, M) FP64, symmetric. */
int mbpt_chi0_build(mbpt_chi0 *h, int ntau, const double *tau_host,
                    const double *weight_host, double *out_dev, int accumulate,
                    void *stream);

/* The per-time-point terms of a frequency sweep, kept in the handle:
 *
 *   P(t) = -2 Go(t) o Gv(t),  t < ntau  (proj(tau_t) in the grid gauge)
 *
 * each built once, so that mbpt_chi0_transform can form chi0 at any number
 * of frequencies from them: the reference's `chi0_tau_terms` and
 * `polarizability_projected_sweep` (space_time.py:533-560, :634-651), whose
 * terms its frequency blocks accumulate as they are built (:854-905).  Held
 * until the next call: ntau M^2 floats (FP32 handles, the tiles padded to
 * 128) or doubles (FP64).  ntau <= MBPT_CHI0_MAX_TAU. */
int mbpt_chi0_terms(mbpt_chi0 *h, int ntau, const double *tau_host, void *stream);

/* out_dev[f] = [accumulate ? out_dev[f] : 0] + sum_t weight_host[f][t] P(t)
 * for f < nfreq: weight_host (nfreq, ntau) row-major (cosft_wt of a
 * TimeFrequencyGrid, or a unit row for one term), out_dev (nfreq, M, M) FP64,
 * symmetric, summed in FP64: chi0(i w_f) for all frequencies in one pass
 * over the terms, the counterpart of `chi0_imaginary_frequency`
 * (space_time.py:607-631).  Needs mbpt_chi0_terms first. */
int mbpt_chi0_transform(mbpt_chi0 *h, int nfreq, const double *weight_host,
                        double *out_dev, int accumulate, void *stream);

/* Device milliseconds and calls of mbpt_chi0_transform since
 * mbpt_chi0_profile (HIP events, waited for). */
int mbpt_chi0_transform_profile_read(const mbpt_chi0 *h, double *ms, long *calls);

/* Device timing of the fused chi0 launches (HIP events on the build's
 * stream, waited for): accumulated milliseconds, launches, and their
 * algorithmic flops, sum over time points of M(M+1)(nocc + nvir + 1)
 * for the upper triangle (with diagonal) each launch computes.  Term builds
 * (mbpt_chi0_terms) count here too, one launch per call. */
int mbpt_chi0_profile(mbpt_chi0 *h, int enable);
int mbpt_chi0_profile_read(const mbpt_chi0 *h, double *ms, long *launches,
                           double *flops);

/* The space-time GW self-energy's branch sums on the grid (FP32 handles):
 *
 *   out[0][t][j] = + sum_PQ X[P][p] (Zt(t) o Go(t))_PQ X[Q][p]
 *   out[1][t][j] = - sum_PQ X[P][p] (Zt(t) o Gv(t))_PQ X[Q][p],  p = states[j]
 *
 * with Go, Gv the handle's Green's functions at tau_host[t] (as in
 * mbpt_chi0_build) and Zt_dev (ntau, M, M) FP32 row-major, the correlation
 * part of the screened interaction on the grid at each time point.
 * out_dev (2, ntau, nstates) FP64.  Replaces `self_energy_branch_sums` /
 * `SigmaPairs.add` (src/SingleReference/GW/imaginary_time.py:564-578,
 * :75-92) in the grid gauge, where Zt = D Wt D^T is formed once per time. */
#define MBPT_GW_MAX_STATES 4096
int mbpt_chi0_sigma(mbpt_chi0 *h, int ntau, const double *tau_host, const float *Zt_dev,
                    int nstates, const int *states_host, double *out_dev, void *stream);

/* Device timing of mbpt_chi0_sigma (enabled and reset by mbpt_chi0_profile):
 * milliseconds, calls and algorithmic flops, ntau M(M+1)(nocc + nvir + 1 +
 * 2 nstates) per call (the triangle's Green's function products, the
 * Hadamard product and the two branches' state contractions). */
int mbpt_chi0_sigma_profile_read(const mbpt_chi0 *h, double *ms, long *launches,
                                 double *flops);

int mbpt_chi0_destroy(mbpt_chi0 *h);

/* ------------------------------------------------------------------------
 * Davidson subspace reductions (FP64)
 *
 * C (a, b) = A^T B for A (n, a) and B (n, b) with row strides lda, ldb (so
 * column slices of a wider trial-space buffer pass directly), C row stride
 * ldc.  The products of the Rayleigh-Ritz step and the orthogonalisation of
 * the trial space, V^T (A V), (A V)^T (A V), V^T T (the reference runs them
 * inside pyscf's lib.davidson1, called at src/Solvers/davidson.py:277).
 * work: mbpt_gemm_tn_f64_work(n, a, b) doubles of device scratch.  n > 0.
 * ------------------------------------------------------------------------ */
long mbpt_gemm_tn_f64_work(int n, int a, int b);
int mbpt_gemm_tn_f64(int n, int a, int b, const double *A, int lda,
                     const double *B, int ldb, double *C, int ldc,
                     double *work, long work_elems, void *stream);

/* mbpt_gemm_tn_f64 without the work argument and without allocating: no
 * counterpart in the reference.  It is for callers that cannot allocate
 * device memory (code under hipGraph capture, tight serving loops, hosts
 * that forbid hipMalloc).  The chunk partials go to a fixed device array of
 * mbpt_gemm_tn_f64_noalloc_max() doubles (32 MiB) that each GPU holds from
 * the library's first use to the process's end; the kernels and the result
 * are those of mbpt_gemm_tn_f64.  Products needing more partials than that
 * (one 64-wide chunk with a * b above it) return MBPT_ERR_VALUE.  Every call
 * on one GPU shares that GPU's array, so calls that overlap on one GPU (two
 * streams, two threads) race: order them on one stream. */
long mbpt_gemm_tn_f64_noalloc_max(void);
int mbpt_gemm_tn_f64_noalloc(int n, int a, int b, const double *A, int lda,
                             const double *B, int ldb, double *C, int ldc,
                             void *stream);

/* ------------------------------------------------------------------------
 * Products of the single-instance Davidson iteration (FP64), the work inside
 * pyscf's lib.davidson1 (called at src/Solvers/davidson.py:277) and its
 * Jacobi-Davidson preconditioner (Solvers/davidson.py:43-77), on a trial
 * space kept as column blocks of row-major (n, ld) buffers, n the pair count.
 *
 * An operand is up to two column ranges of one buffer: logical column i is
 * column off1 + i for i < n1 and off2 + (i - n1) after, so [V | AV] or
 * [V | T] pass as one operand.
 *
 * TN: C (a, b), row stride ldc, = A^T B for A (n, a) and B (n, b) so given;
 *     work: mbpt_gemm_tn_seg_f64_work(n, a, b) doubles.
 * NN: C[:, c_off : c_off + c] = beta C + alpha A Y, A (n, k) so given, Y
 *     (k, c) row-major with ldy, c <= 64 (beta 0: C is not read).
 * Ritz: X = V Y and R = AV Y - X diag(theta), V and AV the m columns at
 *     voff / avoff, Y (m, k), k <= 64: the Ritz vectors and their residuals.
 * Block scale: before the operator is applied to a correction block T (n,
 *     p <= 64, row stride ldt) projected by the two Gram-Schmidt passes of
 *     the orthonormalisation, pr ((3m + 2p + 2) x p, row-major) holding
 *     S1 = V^T T, G0 = T^T T, AV^T T, S2 = V^T T1, G1 = T1^T T1 in its rows:
 *     column j's norm a_j = sqrt(G1_jj - |S2_:j|^2) goes to row 3m + 2p,
 *     its keep flag (a_j > drop |T_:j| before) to the last row, and T's
 *     columns are scaled to unit norm (kept) or zeroed (dropped).
 * JD: T = K^-1 R - alpha K^-1 U per column, K = diag - theta floored at
 *     `floor` in magnitude, alpha = U.K^-1 R / U.K^-1 U; work:
 *     mbpt_jd_correction_f64_work(n, k) doubles.
 * ------------------------------------------------------------------------ */
long mbpt_gemm_tn_seg_f64_work(int n, int a, int b);
int mbpt_gemm_tn_seg_f64(int n, int a, int b, const double *A, int lda, int a1,
                         int aoff1, int aoff2, const double *B, int ldb, int b1,
                         int boff1, int boff2, double *C, int ldc, double *work,
                         long work_elems, void *stream);
int mbpt_gemm_nn_seg_f64(int n, int k, int c, const double *A, int lda, int k1,
                         int aoff1, int aoff2, const double *Y, int ldy, double *C,
                         int ldc, int c_off, double alpha, double beta, void *stream);
int mbpt_ritz_residual_f64(int n, int m, int k, const double *V, int ldv, int voff,
                           const double *AV, int ldav, int avoff, const double *Y,
                           int ldy, const double *theta, double *X, int ldx,
                           double *R, int ldr, void *stream);
int mbpt_block_scale_f64(int n, int m, int p, double drop, double *pr, double *T,
                         int ldt, void *stream);
long mbpt_jd_correction_f64_work(int n, int k);
int mbpt_jd_correction_f64(int n, int k, const double *U, int ldu, const double *R,
                           int ldr, const double *diag, const double *theta,
                           double floor, double *T, int ldt, double *work,
                           long work_elems, void *stream);

/* ------------------------------------------------------------------------
 * Device Rayleigh-Ritz step of the single-instance Davidson (FP64): the m x m
 * eigh of pyscf's lib.davidson1 (called at src/Solvers/davidson.py:277) for
 * the lowest k Ritz pairs, kept on the device.
 *
 * H (m x m, row stride ldh) is the Rayleigh matrix V^T A V, G (row stride
 * ldg) = (AV)^T AV.  The symmetric part of H is reduced to tridiagonal form
 * (Householder), its lowest k eigenvalues found by Sturm-count bisection and
 * their eigenvectors by inverse iteration (orthogonalised within clusters
 * closer than 1e-6 ||T||) and taken back: Y (m x k, row stride ldy) holds
 * them as columns, wq (3k) the eigenvalues ascending, then y^T G y, then
 * y^T H y per column (H as given).  1 <= m <= mbpt_ritz_max_m() (192),
 * 1 <= k <= min(m, 64); work: mbpt_ritz_eigh_f64_work(m, k) doubles. */
int mbpt_ritz_max_m(void);
long mbpt_ritz_eigh_f64_work(int m, int k);
int mbpt_ritz_eigh_f64(int m, int k, const double *H, int ldh, const double *G, int ldg,
                       double *Y, int ldy, double *wq, double *work, long work_elems,
                       void *stream);

/* ------------------------------------------------------------------------
 * Batched small symmetric eigensolver (FP64): the Rayleigh-Ritz step of many
 * Davidson runs at once, the batched counterpart of the m x m eigh inside
 * pyscf's lib.davidson1 (called at src/Solvers/davidson.py:277).
 *
 * For b < nbatch: the symmetric part of A_b = A_dev + b strideA (m x m, row
 * stride lda) is diagonalised; w_dev + b strideW receives the eigenvalues
 * ascending, V_dev + b strideV (row stride ldv) the eigenvectors as columns.
 * 1 <= m <= 64; cyclic Jacobi, at most max_sweeps sweeps (10-20 suffice),
 * stopping when every |a_pq| <= 1e-15 max(sqrt|a_pp a_qq|, scale): scale is
 * the magnitude below which entries count as zero (the spectrum's size).
 * ------------------------------------------------------------------------ */
int mbpt_syev_batched_f64(int nbatch, int m, const double *A_dev, int lda,
                          long strideA, double *w_dev, long strideW,
                          double *V_dev, int ldv, long strideV, int max_sweeps,
                          double scale, void *stream);

/* ------------------------------------------------------------------------
 * Batched subspace products (FP64) of many Davidson runs whose vectors are
 * padded to P pairs with rows_dev[b] (int32, device) real rows each: the
 * V^T W, V y and T - V C steps of the Rayleigh-Ritz and orthogonalisation
 * inside pyscf's lib.davidson1 (called at src/Solvers/davidson.py:277), for
 * a batch.  Operand b starts at ptr + b * stride; rows are row-major with
 * the given leading dimensions.  Rows at and past rows[b] are not read.
 *
 * NN: Y_b = A_b S_b (subtract = 0; rows past rows[b] of Y_b set to zero) or
 *     Y_b -= A_b S_b (subtract = 1; rows past rows[b] untouched), A_b (P x m),
 *     S_b (m x c), m <= 64.
 * TN: C_b = A_b^T B_b, A_b (P x a), B_b (P x c), C_b (a x c), a, c <= 64;
 *     work of mbpt_btn_f64_work(B, P, a, c) doubles.
 * ------------------------------------------------------------------------ */
int mbpt_bnn_f64(int B, int P, int m, int c, const int *rows_dev,
                 const double *A, long sA, int lda, const double *S, long sS,
                 int lds, double *Y, long sY, int ldy, int subtract, void *stream);
long mbpt_btn_f64_work(int B, int P, int a, int c);
int mbpt_btn_f64(int B, int P, int a, int c, const int *rows_dev,
                 const double *A, long sA, int lda, const double *Bm, long sB,
                 int ldb, double *C, long sC, int ldc, double *work,
                 long work_elems, void *stream);

/* ------------------------------------------------------------------------
 * Batched sigma-vectors of n independent ISDF instances of uneven size, the
 * TDA `apply_AB` (block_actions.py:563-661) of every instance in one pass.
 * The reference has no batched path: it would run n separate closures.
 *
 * Each stage (head, Hartree, exchange per 16-wide occupied bucket, tail) is
 * one launch over all instances driven by a device work table; the launch
 * sequence is fixed per handle, so it can be captured in a hipGraph.
 * ------------------------------------------------------------------------ */
typedef struct mbpt_batch mbpt_batch;

/* n instances: M[b], nocc[b] (<= 64), nvir[b]; X_dev[b] (M, nocc + nvir),
 * W_dev[b] and V_dev[b] (M, M) device FP64 (V unread when kappa == 0);
 * eps_host the n orbital-energy vectors concatenated; precision MBPT_FP32 or
 * MBPT_FP64; kmax trial vectors per instance per apply.  The operands are
 * copied.  A default plan is made (see mbpt_batch_plan). */
int mbpt_batch_create(mbpt_batch **out, int n, const int *M, const int *nocc,
                      const int *nvir, int precision, double kappa,
                      const double *const *X_dev, const double *const *W_dev,
                      const double *const *V_dev, const double *eps_host,
                      int kmax, void *stream);

/* Work decomposition and tile shape (the tuning knobs; results do not
 * depend on them beyond rounding): grid rows per exchange work item
 * (>= 64), K chunk of the tail product T' X_v and of the Hartree product
 * p V (>= 16), the persistent grid size of every launch (0: one workgroup
 * per work item), and the exchange's bucket width (16 or 32): each
 * instance's n_o is padded to a multiple of it, which sets the K depth of
 * the exchange's P product and the instances sharing one launch.
 * Synchronises the stream; reallocates the workspace, so a graph captured
 * before must be captured again. */
int mbpt_batch_plan(mbpt_batch *h, int rows_per_split, int tail_kchunk,
                    int hartree_kchunk, int grid_cap, int bucket_width, void *stream);

/* out_dev (n, npair_max) FP64: row b, first nocc[b] nvir[b] entries, gets
 * instance b's screened diagonal d - (ii|W|aa) (`screened_diagonal()`,
 * block_actions.py:663-689); the padding is not written.  Synchronises. */
int mbpt_batch_screened_diag(mbpt_batch *h, double *out_dev, void *stream);

/* Largest nocc * nvir over the instances: the pair stride of the vectors. */
int mbpt_batch_npair_max(const mbpt_batch *h);
size_t mbpt_batch_held_bytes(const mbpt_batch *h);

/* Az = A z for every instance: z_dev and Az_dev (n, kmax, npair_max) FP64,
 * instance b's (nocc, nvir) pairs first in each row; kact_dev (n) int32 on
 * the device (NULL: all kmax) gives the live vectors of each instance, the
 * leading kact[b] rows -- rows after them are neither read nor written, nor
 * is the pair padding. */
int mbpt_batch_apply(mbpt_batch *h, const double *z_dev, double *Az_dev,
                     const int *kact_dev, void *stream);
int mbpt_batch_destroy(mbpt_batch *h);

/* ------------------------------------------------------------------------
 * Learned Davidson preconditioner: a 3 -> 256 -> 256 -> 1 SiLU network, BF16
 * weights and activations, FP32 accumulation.  No counterpart in the
 * reference: it stands in for the diagonal correction r / (d - theta)
 * (LR/davidson_shared.py:157-163, Solvers/davidson.py:43-77).
 *
 * Parameters (FP32, host order): W1 (3, 256), b1 (256), W2 (256, 256),
 * b2 (256), w3 (256), b3 (1); mbpt_mlp_nparams() of them.
 * ------------------------------------------------------------------------ */
typedef struct mbpt_mlp mbpt_mlp;

int mbpt_mlp_nparams(void);
int mbpt_mlp_create(mbpt_mlp **out, const float *params_host, void *stream);
int mbpt_mlp_get_params(mbpt_mlp *h, float *params_host, void *stream);
/* y (N) = net(X (N, 3)) */
int mbpt_mlp_forward(mbpt_mlp *h, long N, const float *X_dev, float *y_dev,
                     void *stream);
/* mean-squared-error loss against T (N) and its gradient (nparams, device;
 * NULL to skip the copy), loss on the host (NULL to skip) */
int mbpt_mlp_gradient(mbpt_mlp *h, long N, const float *X_dev,
                      const float *T_dev, float *grad_dev, float *loss_host,
                      void *stream);
/* one Adam step (beta 0.9 / 0.999) on the mean-squared error */
int mbpt_mlp_train_step(mbpt_mlp *h, long N, const float *X_dev,
                        const float *T_dev, float lr, float *loss_host,
                        void *stream);
/* Davidson correction for k residual columns: R (n, k) and T (n, k) FP64
 * with row strides ldr, ldt; d (n) pair energies, theta (k) Ritz values.
 * T[i][j] = rms_j * net(R[i][j] / rms_j, d_i - theta_j, theta_j), rms_j the
 * root-mean-square of column j of R. */
int mbpt_mlp_correction(mbpt_mlp *h, int n, int k, const double *R_dev,
                        int ldr, const double *d_dev, const double *theta_dev,
                        double *T_dev, int ldt, void *stream);
/* The same for a batch of Davidson runs: nbatch instances, each a block of P
 * rows of R (row stride ldr, k columns), d and T (row stride ldt), of which
 * the first npair[b] are real; off_dev (nbatch + 1) int32 on the device is
 * the prefix sum of npair, ntot = off[nbatch].  theta_dev (nbatch, k); rms
 * per instance and column over its real rows.  Rows past npair[b] of T are
 * not written. */
int mbpt_mlp_correction_batched(mbpt_mlp *h, int nbatch, int P, int k,
                                const int *off_dev, long ntot,
                                const double *R_dev, int ldr,
                                const double *d_dev, const double *theta_dev,
                                double *T_dev, int ldt, void *stream);
int mbpt_mlp_destroy(mbpt_mlp *h);

#ifdef __cplusplus
}
#endif
#endif /* MBPT_HIP_H */
