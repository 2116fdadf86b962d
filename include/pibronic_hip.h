/* C-ABI of libpibronic_hip.so: the MI355X (gfx950) kernels behind Pibronic's per-block
 * PIMC estimator.
 *
 * The reference has no FFI: its boundary is the Python function API of
 * pibronic/pimc/pimc.py, whose module-level kernels take the precomputed model constants
 * (`model.const`, a TemperatureDependentClass, and `state_shift`) and mutate preallocated
 * block buffers. This header is that seam as plain C: the Python host (pibronic_amd.pimc)
 * keeps the reference's classes and computes the same constants with NumPy, then hands
 * them over once per run; each block is one or three calls.
 *
 * Conventions
 *  - All arrays are C-order float64 unless stated. Pointers named *host* are read during
 *    the call and may be freed afterwards; pointers named *dev* are device memory that the
 *    caller owns (for example torch tensors on cuda:0) and must stay valid until the work
 *    queued on `stream` has finished.
 *  - `stream` is a hipStream_t passed as void*; NULL means the default stream.
 *  - Every function returns PIB_OK (0) or a negative PIB_ERR_* code; the message is
 *    available from pib_last_error(ctx). Nothing throws or aborts across the ABI.
 *  - A context is bound to one device and is not re-entrant: calls on one context must not
 *    overlap. Use one context per host thread.
 */
#ifndef PIBRONIC_HIP_H
#define PIBRONIC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIB_ABI_VERSION 1

#define PIB_OK 0
#define PIB_ERR_ARG (-1)         /* invalid argument (reference: AssertionError)        */
#define PIB_ERR_HIP (-2)         /* HIP runtime error                                   */
#define PIB_ERR_STATE (-3)       /* required constants not set                          */
#define PIB_ERR_UNSUPPORTED (-4) /* surface count has no compiled kernel               */

/* slots of pib_set_omatrix_set */
#define PIB_SET_RHO 0       /* sampling model at tau          (data.rho.const)        */
#define PIB_SET_VIB 1       /* coupled model at tau           (data.vib.const)        */
#define PIB_SET_VIB_PLUS 2  /* coupled model at tau_plus      (data.vib.const_plus)   */
#define PIB_SET_VIB_MINUS 3 /* coupled model at tau_minus     (data.vib.const_minus)  */

/* flags of pib_eval_block */
#define PIB_EVAL_PM 1      /* also g(beta +- dbeta): block_compute_pm                     */
#define PIB_EVAL_NOSCALE 2 /* skip the S12 scaling (S_p = 1), as the explicit golden test */
#define PIB_EVAL_WIDE 4    /* force the one-wave-per-sample kernel (used for A > 8; any
                              A, A_rho <= 16); a tuning/testing switch, same results       */
#define PIB_EVAL_FP32 8    /* reduced precision (A <= 8): O-factors, V, M = exp(-tau V) (a
                              scaled and squared Taylor polynomial) and the ring products in
                              FP32 (power-of-two renormalised); FP64 log-scale and rho
                              log-sums; error vs FP64 is the caller's to report (the
                              outputs have the same meaning)                               */
#define PIB_EVAL_BF16X3 16 /* wide kernel with the ring products as BF16x3 splits on the
                              bf16 matrix cores, FP32 accumulation, renormalised every bead
                              (A, A_rho <= 16); M = exp(-tau V) and the O-factors stay FP64  */

/* rows of the per-sample output matrix written by pib_eval_block */
#define PIB_OUT_RHO 0  /* s_rho = rho / prod_p S_p             (BoxResult.scaled_rho)       */
#define PIB_OUT_G 1    /* s_g   = g   / prod_p S_p             (BoxResult.scaled_g)         */
#define PIB_OUT_GP 2   /* s_gP                                 (BoxResultPM.scaled_gofr_plus)  */
#define PIB_OUT_GM 3   /* s_gM                                 (BoxResultPM.scaled_gofr_minus) */
#define PIB_OUT_LOGS 4 /* sum_p log S_p (no reference counterpart: the dropped scale)      */
#define PIB_NUM_OUT 5

/* entries of the block-sum vector written by pib_block_sums */
#define PIB_SUM_N 0     /* samples                                  */
#define PIB_SUM_R 1     /* sum r, r = g/rho                         */
#define PIB_SUM_R2 2    /* sum r^2                                  */
#define PIB_SUM_ABSR 3  /* sum |r|                                  */
#define PIB_SUM_NEG 4   /* number of r < 0                          */
#define PIB_SUM_D1 5    /* sum d1, d1 = (g+ - g-)/(2 dbeta rho)     */
#define PIB_SUM_D1SQ 6  /* sum d1^2                                 */
#define PIB_SUM_D2 7    /* sum d2, d2 = (g+ - 2g + g-)/(dbeta^2 rho) */
#define PIB_SUM_D2SQ 8  /* sum d2^2                                 */
#define PIB_SUM_RD1 9   /* sum r d1                                 */
#define PIB_SUM_RD2 10  /* sum r d2                                 */
#define PIB_SUM_D1D2 11 /* sum d1 d2                                */
#define PIB_SUM_RP 12   /* sum g+/rho                               */
#define PIB_SUM_RM 13   /* sum g-/rho                               */
#define PIB_NUM_SUMS 14

typedef struct pib_ctx pib_ctx;

/* Library / context lifetime. No reference counterpart (the reference is in-process
 * NumPy); the context holds the per-run constants that BoxData.preprocess builds
 * (pibronic/pimc/pimc.py:647-692). */
int pib_abi_version(void);
int pib_ctx_create(int device, pib_ctx** out);
int pib_ctx_destroy(pib_ctx* ctx);
const char* pib_last_error(const pib_ctx* ctx);

/* Residual diabatic coupling of the coupled model after folding:
 * energy_host (A*A) with the diagonal zeroed, linear_host (N*A*A) with the surface
 * diagonal zeroed, quadratic_host (N*N*A*A, may be NULL for none). Replaces the model
 * state read by diagonalize_coupling_matrix (pibronic/pimc/pimc.py:1139-1160) as left by
 * ModelVibronic.finish_folding_in_terms (:208-218). */
int pib_set_coupling(pib_ctx* ctx, int A, int N, const double* energy_host,
                     const double* linear_host, const double* quadratic_host);

/* One set of O-matrix constants: coth_host, csch_host (N), prefactor_host (A_set) and the
 * model's state shift state_shift_host (A_set*N). Replaces the TemperatureDependentClass
 * fields cothAN/cschAN/omatrix_prefactor (pibronic/pimc/pimc.py:59-83) and the
 * `state_shift` argument of build_o_matrix (:1087). The reference broadcasts coth and csch
 * over surfaces; they are passed once per mode. */
int pib_set_omatrix_set(pib_ctx* ctx, int slot, int A_set, int N, const double* coth_host,
                        const double* csch_host, const double* prefactor_host,
                        const double* state_shift_host);

/* Bead count P (>= 3, pibronic/pimc/pimc.py:683) and the base tau = beta/P used for
 * M = exp(-tau V) in every ring product (build_numerator, :1183; quirk Q1). */
int pib_set_beads(pib_ctx* ctx, int P, double tau);

/* pib_eval_block for coordinates given as float offsets dq_dev (B*N*P) from the per-mode centre
 * m of pib_flow_set_basis: bead coordinate (s, j, p) = m[j] + (double)dq[(s N + j) P + p],
 * formed in FP64 inside the kernel -- the same doubles pib_flow_whiten(PIB_FLOW_UNWHITEN)
 * writes, so the outputs equal pib_eval_block's on those, bit for bit. The FP64 lanes kernel
 * (A <= 8, A_rho = A) only; flags 0 or PIB_EVAL_NOSCALE. No counterpart in the reference
 * (whose coordinates are float64 arrays); for the learned proposal's draws, whose unwhitening
 * then writes and this reads half the bytes. */
int pib_eval_block_flow_offsets(pib_ctx* ctx, const float* dq_dev, int64_t B, int flags, double* out_dev,
                                int64_t out_stride, void* stream);

/* Lanes cooperating on one sample (power of two <= 64), 0 = automatic. Tuning only. The getter
 * reports the count the FP64 lanes kernel uses for a block of B samples. */
int pib_set_lanes_per_sample(pib_ctx* ctx, int lanes);
int pib_get_lanes_per_sample(pib_ctx* ctx, int64_t B);

/* Evaluate one block of B samples from bead coordinates q_dev (B*N*P, the reference's
 * transformed qTensor[:, 0] of pibronic/pimc/pimc.py:618-624). Writes PIB_NUM_OUT rows of
 * out_stride (>= B) doubles to out_dev: the S12-scaled rho and g (and g+- with
 * PIB_EVAL_PM) of block_compute / block_compute_pm (:1341-1385, :1388-1462), computed by
 * build_o_matrix, build_scaling_factors, scale_o_matrices, build_denominator,
 * diagonalize_coupling_matrix and build_numerator (:1062-1213) fused into one launch.
 * debug_dev, if not NULL, receives per (sample, bead) records of
 * [O_rho (A_rho), O_vib (A), S_p, M (A*A)] (unscaled O), B*P records.
 * A <= 8 runs a lanes-per-sample kernel; 8 < A <= 16 (and any A with PIB_EVAL_WIDE) runs one
 * wave per sample with 16x16 padded matrices. A or A_rho > 16: PIB_ERR_UNSUPPORTED. */
int pib_eval_block(pib_ctx* ctx, const double* q_dev, int64_t B, int flags, double* out_dev,
                   int64_t out_stride, double* debug_dev, void* stream);

/* Sampling constants of the harmonic rho: surface weights weights_host (A_rho, normalised,
 * ModelSampling.compute_weight_for_each_state, pibronic/pimc/pimc.py:341-351), the
 * sampling model's coth_host and csch_host (N, its TemperatureDependentClass at tau) and its
 * state shift (A_rho*N, :372-381). Together with P they define the Gaussian of
 * compute_sampling_constants (:398-406): per mode, precision 2 coth I - csch C with C the
 * ring adjacency matrix of the circulant basis (:681-689). The library factors it once. */
int pib_set_sampler(pib_ctx* ctx, int A_rho, int N, int P, const double* weights_host,
                    const double* coth_host, const double* csch_host,
                    const double* state_shift_host);

/* Draw B samples of bead coordinates into q_dev (B*N*P) and, optionally, their surface
 * sources into src_dev (B int32). Replaces draw_sample + transform_sampled_coordinates
 * (pibronic/pimc/pimc.py:595-600, 326-334, 613-631). Counter-based: sample i of the call
 * uses the Philox counter of global sample index offset+i under key `seed`, so blocks are
 * independent streams and any split of a run gives the same samples. */
int pib_sample_block(pib_ctx* ctx, uint64_t seed, int64_t offset, int64_t B, double* q_dev,
                     int32_t* src_dev, void* stream);

/* Reduce a block's per-sample outputs (as written by pib_eval_block) into PIB_NUM_SUMS
 * doubles at sums_dev, in a fixed order (deterministic). The per-sample terms are those of
 * calculate_basic_property_terms (pibronic/stats/stats.py:38-55); with pm == 0 the
 * derivative entries are zero. */
int pib_block_sums(pib_ctx* ctx, const double* out_dev, int64_t out_stride, int64_t B, int pm,
                   double delta_beta, double* sums_dev, void* stream);

/* pib_block_sums for samples drawn from another proposal than rho (the learned flow):
 * logq_dev (B doubles) is its log density in rho's normalisation, and every r = g/q uses
 * q = exp(logq - PIB_OUT_LOGS) in place of the PIB_OUT_RHO row. */
int pib_block_sums_q(pib_ctx* ctx, const double* out_dev, int64_t out_stride, int64_t B, int pm,
                     double delta_beta, const double* logq_dev, double* sums_dev, void* stream);

/* pib_block_sums (logq_dev NULL) or pib_block_sums_q (logq_dev set) without any device
 * allocation: the partial rows live in a fixed per-GPU array of the library (1024 rows of
 * PIB_NUM_SUMS doubles, 114,688 bytes per GPU, held until the process ends) instead of memory
 * the context allocates on first use. Every B is accepted: the partial rows are capped at
 * 1024 (the reduction strides over the block). No counterpart in the reference; for callers
 * that cannot call hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation). All calls on one GPU share its array: calls that overlap on one GPU
 * (different streams or host threads) race and must be serialised by the caller. Same results,
 * bit for bit, as pib_block_sums / pib_block_sums_q. */
int pib_block_sums_noalloc(pib_ctx* ctx, const double* out_dev, int64_t out_stride, int64_t B,
                           int pm, double delta_beta, const double* logq_dev, double* sums_dev,
                           void* stream);

/* ---------------------------------------------------------------------------------------
 * Learned proposal (affine-coupling flow over bead normal-mode coordinates). No reference
 * counterpart: the reference's nearest hooks are the ML data generators
 * pibronic/pimc/pimc.py:1216-1338 (g and rho at externally chosen R). These entry points are
 * the building blocks a flow needs on the device; pibronic_amd.flow composes them.
 * Device pointers throughout; float32 unless stated.
 * ------------------------------------------------------------------------------------- */

/* whitening modes of pib_flow_whiten */
#define PIB_FLOW_WHITEN 0   /* R (double, B*N*P) -> zp (float, B*N*P), Z = (R - m) U / sigma  */
#define PIB_FLOW_UNWHITEN 1 /* zp (float) -> R (double) = m + (Z sigma) U^T                   */
#define PIB_FLOW_GRAD 2     /* dR (float) -> dzp (float) = (dR U) sigma: dL/dZ from dL/dR     */
#define PIB_FLOW_UNWHITEN_OFFSETS 3 /* zp (float) -> dR (float) = (Z sigma) U^T, so that the
                                     * coordinates are R = m + dR (P = 64 or 128): the input of
                                     * pib_eval_block_flow_offsets, at half the bytes of R */

/* C = op(A) op(B) (+ bias[N]) (relu) (* [mask > 0]) (+ C if accumulate), FP32 inputs and
 * accumulation on MFMA. M, N multiples of 64, K a multiple of 16. */
int pib_gemm_f32(pib_ctx* ctx, int trans_a, int trans_b, int64_t M, int64_t N, int64_t K,
                 const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
                 int64_t ldc, const float* bias, int relu, const float* mask, int64_t ldmask,
                 int accumulate, void* stream);

/* The same product with each FP32 operand split into bf16 hi + lo parts and
 * C = A_hi B_hi + A_hi B_lo + A_lo B_hi on the bf16 matrix cores, FP32 accumulation
 * (~2^-16 relative error per product term instead of FP32's 2^-24; K a multiple of 32). */
int pib_gemm_bf16x3(pib_ctx* ctx, int trans_a, int trans_b, int64_t M, int64_t N, int64_t K,
                 const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
                 int64_t ldc, const float* bias, int relu, const float* mask, int64_t ldmask,
                 int accumulate, void* stream);

/* BF16x3 product of operands already split into bf16 hi/lo parts (uint16_t holding bf16
 * bits): A_hi/A_lo M x K row-major (lda), B given transposed as Bt_hi/Bt_lo N x K row-major
 * (ldb). Writes C (FP32, ldc) and/or the result split again into O_hi/O_lo (ldo), after
 * bias and relu. M, N multiples of 64, K of 32, lda and ldb multiples of 8, operands 16-byte
 * aligned. Lets a chain of dense layers pass bf16 hi/lo activations without conversion. */
int pib_gemm_bf16x3_presplit(pib_ctx* ctx, int64_t M, int64_t N, int64_t K,
                             const uint16_t* A_hi_dev, const uint16_t* A_lo_dev, int64_t lda,
                             const uint16_t* Bt_hi_dev, const uint16_t* Bt_lo_dev, int64_t ldb,
                             float* C_dev, int64_t ldc, uint16_t* O_hi_dev, uint16_t* O_lo_dev,
                             int64_t ldo, const float* bias_dev, int relu, void* stream);

/* Split X (rows x cols, ldx) into bf16 hi = RNE(x), lo = RNE(x - hi). mode bits:
 * PIB_SPLIT_TRANSPOSE: the outputs are cols x rows (row length ldo);
 * PIB_SPLIT_PERM32 (with TRANSPOSE, rows a multiple of 32): within every block of 32 output
 * columns, column 32b + w is stored at 32b + 8((w % 16) / 4) + 4(w / 16) + w % 4, the operand
 * order of pib_flow_layer_fused's second and third products. */
#define PIB_SPLIT_TRANSPOSE 1
#define PIB_SPLIT_PERM32 2
/* PIB_SPLIT_FRAG32 (with TRANSPOSE only; cols a multiple of 32, rows of 16): the transposed
 * matrix (cols x rows) in the A-operand order of the 32x32x16 bf16 matrix-core product, as
 * pib_flow_layer_fused_frag reads it: element (n, k) at ((n/32 * rows/16 + k/16) * 64 +
 * 32 ((k % 16) / 8) + n % 32) * 8 + k % 8; ldo is not used. For any mode lo_dev may be NULL
 * (only hi is written). */
#define PIB_SPLIT_FRAG32 4
int pib_split_bf16(pib_ctx* ctx, int64_t rows, int64_t cols, const float* X_dev, int64_t ldx,
                   int mode, uint16_t* hi_dev, uint16_t* lo_dev, int64_t ldo, void* stream);

/* Normal-mode basis of the flow: U_host (P*P, the circulant eigenvectors, row p = bead),
 * sigma_host (N*P, harmonic standard deviations per mode and normal mode) and m_host (N, the
 * centre subtracted before rotating). P must be a multiple of 64 (MFMA tiles of the
 * whitening product; the coupling halves are the even-k and odd-k normal modes). */
int pib_flow_set_basis(pib_ctx* ctx, int N, int P, const double* U_host,
                       const double* sigma_host, const double* m_host);

/* Whitening transforms between R and the flow's permuted coordinates (modes above);
 * work_dev: 2*B*N*P floats of scratch. B*N a multiple of 64. */
int pib_flow_whiten(pib_ctx* ctx, int mode, int64_t B, const void* in_dev, float* work_dev,
                    void* out_dev, void* stream);

/* One coupling layer on z (B*D, in place on the half at b_off) from the MLP output o
 * (B*D: raw scales | shifts); logdet (double, B) += sum s (forward) or -= sum s (inverse). */
int pib_flow_coupling(pib_ctx* ctx, int inverse, int64_t B, int D, int a_off, int b_off,
                      float smax, float* z_dev, const float* o_dev, double* logdet_dev,
                      void* stream);

/* Backward of an inverse coupling layer for L = -sum_b c_b (log rho(R0_b) + logdet_b):
 * y = saved layer input, dz = dL/d(output) (B-half overwritten with dL/d(input)),
 * dout = dL/d(o). */
int pib_flow_coupling_inverse_backward(pib_ctx* ctx, int64_t B, int D, int a_off, int b_off,
                                       float smax, const float* y_dev, const float* o_dev,
                                       float* dz_dev, float* dout_dev, const double* c_dev,
                                       void* stream);

/* One whole coupling layer in one launch: the MLP of the layer (D/2 -> hidden -> hidden -> D,
 * ReLU) as BF16x3 products, then the coupling of pib_flow_coupling on z (B x D) and logdet.
 * Weights split by pib_split_bf16: W1 with PIB_SPLIT_TRANSPOSE (hidden x D/2), W2 and W3 with
 * PIB_SPLIT_TRANSPOSE | PIB_SPLIT_PERM32 (hidden x hidden, D x hidden). Same arithmetic as the
 * layer-by-layer calls (pib_gemm_bf16x3_presplit x3 + pib_flow_coupling) up to the FP32
 * summation order; the activations and the MLP output never leave the chip. Supported:
 * hidden 256, D = 512 (and D = 1024 with the lo parts NULL), B a multiple of 128.
 * W1_lo_dev, W2_lo_dev and W3_lo_dev may all be NULL:
 * the products are then single bf16 products of the hi parts (a third of the matrix-core work;
 * the coupling and its log-determinant are exact for whatever scales and shifts the network
 * produced, so densities of the map applied stay exact). No counterpart in the reference. */
int pib_flow_layer_fused(pib_ctx* ctx, int inverse, int64_t B, int D, int hidden, int a_off,
                         int b_off, float smax, float* z_dev, double* logdet_dev,
                         const uint16_t* W1_hi_dev, const uint16_t* W1_lo_dev,
                         const uint16_t* W2_hi_dev, const uint16_t* W2_lo_dev,
                         const uint16_t* W3_hi_dev, const uint16_t* W3_lo_dev, const float* b1_dev,
                         const float* b2_dev, const float* b3_dev, void* stream);

/* The layer of pib_flow_layer_fused with single bf16 products (its NULL lo parts), on weights
 * split with PIB_SPLIT_TRANSPOSE | PIB_SPLIT_FRAG32 (all three, inputs in natural order; only
 * the hi parts): each wave streams its weight tiles as contiguous 1 KB blocks. Same network,
 * coupling and exactness of the log-determinant; the FP32 summation order differs. Supported:
 * hidden 256, D = 512 or 1024, B a multiple of 128; weights and z 16-byte aligned. No counterpart in
 * the reference. */
int pib_flow_layer_fused_frag(pib_ctx* ctx, int inverse, int64_t B, int D, int hidden, int a_off,
                              int b_off, float smax, float* z_dev, double* logdet_dev,
                              const uint16_t* W1_frag_dev, const uint16_t* W2_frag_dev,
                              const uint16_t* W3_frag_dev, const float* b1_dev, const float* b2_dev,
                              const float* b3_dev, void* stream);

/* Base distribution of the flow drawn directly in its whitened, permuted coordinates: the
 * harmonic rho is the mixture sum_a w_a N(mu_a, I) there (mu_a (D, double) = the surface's
 * displacement d_a - m rotated by the circulant basis and divided by sigma; w_a the sampler's
 * surface weights, pib_set_sampler). c_a = P log F_a - |mu_a|^2 / 2 (F_a the O-matrix
 * prefactor of the sampling model), so that
 *   log rho(R0) = logsumexp_a (c_a + z . mu_a) - |z|^2 / 2
 * for R0 the unwhitened z: the same function as pib_flow_logrho (the O-matrix exponent summed
 * over the ring is (x - d)^T Lambda (x - d)). No counterpart in the reference, which draws R0
 * and evaluates rho at it; this replaces pib_sample_block + pib_flow_whiten +
 * pib_flow_logrho for flow draws. */
int pib_flow_set_base(pib_ctx* ctx, int A_rho, int D, const double* mu_host, const double* c_host);

/* B draws z (float, B x D, the flow's permuted coordinates) and log rho(R0) (double, B); the
 * surface of sample s is drawn as in pib_sample_block (Philox counter (offset + s, 0, 1)), the
 * normals from Philox counters (offset + s, m, 2), m the normal pair. */
int pib_flow_draw_base(pib_ctx* ctx, uint64_t seed, int64_t offset, int64_t B, float* z_dev,
                       double* logrho_dev, void* stream);

/* log rho(R) of the sampling model (unscaled, double, B) and, if dR_dev is not NULL,
 * dR = -c * d log rho / dR (float, B*N*P; c_dev NULL means c = 1). Uses the PIB_SET_RHO
 * constants and P. Same function as build_o_matrix + build_denominator
 * (pibronic/pimc/pimc.py:1087-1136) in log form. */
int pib_flow_logrho(pib_ctx* ctx, int64_t B, const double* R_dev, const double* c_dev,
                    float* dR_dev, double* logrho_dev, void* stream);

/* out[j] = sum_i X[i*ldx + j] for a M x N matrix. */
int pib_colsum_f32(pib_ctx* ctx, int64_t M, int64_t N, const float* X_dev, int64_t ldx,
                   float* out_dev, void* stream);

/* One Adam step (Kingma & Ba) on n parameters; step counts from 1. */
int pib_adam_f32(pib_ctx* ctx, int64_t n, float* p_dev, const float* g_dev, float* m_dev,
                 float* v_dev, float lr, float beta1, float beta2, float eps, int step,
                 void* stream);


/* ------------------------------------------------------------------------------------------
 * Batched evaluation. No counterpart in the reference, which runs one (model, temperature)
 * per process and farms them out (pibronic/server/job_boss.py:538-611, PimcSubmissionClass).
 * A batch evaluates one block of B samples for each of n problems (contexts: model, O-matrix
 * sets, beads, sampler) in ONE persistent launch; problems may differ in P (the temperatures
 * of a run) and, for the wide family, in A, A_rho and N. The contexts must outlive the batch
 * and keep their constants while it exists: the batch copies their device pointers, so after
 * any pib_set_coupling / pib_set_omatrix_set / pib_set_beads / pib_set_sampler on one of its
 * contexts, pib_batch_eval, _eval_sequential, _tune and _run return PIB_ERR_STATE without
 * launching (create a new batch). Not re-entrant (one stream at a time).
 * ---------------------------------------------------------------------------------------- */
typedef struct pib_batch pib_batch;

#define PIB_BATCH_LANES 0 /* lanes-per-sample kernel: one compiled (A, A_rho) for all   */
#define PIB_BATCH_WIDE 1  /* one wave per sample, 16x16 padded: any A, A_rho <= 16      */
#define PIB_BATCH_SEQUENTIAL 2 /* one launch per problem (tuned lanes), same graph step   */

/* flags: PIB_EVAL_PM, PIB_EVAL_NOSCALE, PIB_EVAL_FP32 (then one launch per problem); delta_beta is used by the +-dbeta sums;
 * max_blocks bounds the per-block sums kept by pib_batch_run. Buffers are the caller's
 * (pib_batch_bind, pib_batch_bind_sums); the batch owns only its schedule and scratch. */
int pib_batch_create(int n, pib_ctx* const* ctxs, int64_t B, int flags, double delta_beta,
                     int max_blocks, pib_batch** out);
int pib_batch_destroy(pib_batch* batch);
const char* pib_batch_last_error(const pib_batch* batch);

/* Problem i reads q_dev (B*N*P doubles, as pib_eval_block) and writes PIB_NUM_OUT rows of
 * out_stride (>= B) doubles at out_dev. Rebinding drops a captured graph. */
int pib_batch_bind(pib_batch* batch, int i, const double* q_dev, double* out_dev,
                   int64_t out_stride);
/* Block sums of pib_batch_run: block k of problem i goes to
 * sums_dev + (k * n + i) * PIB_NUM_SUMS (max_blocks * n * PIB_NUM_SUMS doubles). */
int pib_batch_bind_sums(pib_batch* batch, double* sums_dev);

/* Schedule: kernel family, lanes per sample of every problem (lanes family; NULL keeps the
 * current ones), tile size (waves of the lanes family, samples of the wide family) and the
 * number of persistent waves. */
int pib_batch_set_schedule(pib_batch* batch, int family, const int* lanes, int tile_waves,
                           int grid_waves);
int pib_batch_get_schedule(pib_batch* batch, int* family, int* lanes, int* tile_waves,
                           int* grid_waves);

/* Autotuner: samples every problem once (seed) into its bound q, then times schedules on the
 * device (coordinate search over family -- persistent lanes, persistent wide, or one launch
 * per problem --, beads per lane, tile size and grid size) and keeps the fastest.
 * best_ms = its time per evaluation step; sequential_ms (optional) = one launch per problem
 * with the default lanes (as pib_eval_block), for comparison. */
int pib_batch_tune(pib_batch* batch, uint64_t seed, double* best_ms, double* sequential_ms,
                   void* stream);

/* Evaluate the q currently in the batch buffers: one persistent launch. */
int pib_batch_eval(pib_batch* batch, void* stream);
/* The same problems as one launch each (reference point for the batched launch). */
int pib_batch_eval_sequential(pib_batch* batch, void* stream);

/* nblocks block steps, blocks first_block .. first_block+nblocks-1: for each, sample every
 * problem (Philox offsets block*B under `seed`, the counters of pib_sample_block), evaluate
 * all problems in one persistent launch, and reduce each problem's block into its sums row.
 * The step is captured once as a hipGraph (kernels only) and replayed. */
int pib_batch_run(pib_batch* batch, uint64_t seed, int64_t first_block, int nblocks,
                  void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PIBRONIC_HIP_H */
