// Learned-proposal (affine-coupling flow) kernels: argument blocks and launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pib {

struct GemmArgs {
  const float* A;
  const float* B;
  float* C;
  const float* bias;   // N or nullptr
  const float* mask;   // M x N (ldmask) or nullptr: C = (A B) * [mask > 0]
  int64_t M, N, K, lda, ldb, ldc, ldmask;
  int relu;            // C = max(C, 0) after the bias
  int accumulate;      // C += result
  // split-K: grid.z = splits, split z covers K range [z*kchunk, (z+1)*kchunk) and writes its
  // raw tile to partial + z*M*N; a second kernel adds the splits in order and applies the
  // epilogue (deterministic)
  int64_t kchunk;
  float* partial;
  int bf16x3;  // 1: BF16x3 split on the bf16 matrix cores (K a multiple of 32)
};

hipError_t launch_gemm_f32(bool ta, bool tb, const GemmArgs& g, hipStream_t s);

// BF16x3 product of pre-split operands (bf16 bits as unsigned short): A hi/lo M x K (lda),
// B hi/lo stored N x K (ldb); output FP32 C (ldc) and/or split hi/lo (ldo); bias, relu.
struct GemmPreArgs {
  const unsigned short* Ah;
  const unsigned short* Al;
  const unsigned short* Bh;
  const unsigned short* Bl;
  float* C;
  unsigned short* Oh;
  unsigned short* Ol;
  const float* bias;
  int64_t M, N, K, lda, ldb, ldc, ldo;
  int relu;
};
bool gemm_pre_shape_ok(const GemmPreArgs& g);
hipError_t launch_gemm_bf16x3_pre(const GemmPreArgs& g, hipStream_t s);
hipError_t launch_split_bf16(int64_t rows, int64_t cols, const float* X, int64_t ldx, int transpose,
                             unsigned short* hi, unsigned short* lo, int64_t ldo, hipStream_t s);
bool gemm_shape_ok(const GemmArgs& g);
// number of K splits worth using for this shape (1 = none)
int gemm_splits(const GemmArgs& g, int num_cu);

// Whitening between bead coordinates R (B, N, P) and the flow's permuted normal-mode
// coordinates zp (B, D): zp[b][h*D/2 + j*P/2 + kk] = Z[b][j][2 kk + h], Z = (R - m) U / sigma.
struct BasisArgs {
  int64_t B;
  int N, P;
  const float* sigma;      // N * P
  const float* inv_sigma;  // N * P
  const double* m;         // N (mixture mean of the sampling model's displacements)
};

// mode of launch_flow_pre / launch_flow_post
enum : int {
  FLOW_WHITEN = 0,   // R (double) -> X = R - m (float) ; Y -> zp = perm(Y / sigma)
  FLOW_UNWHITEN = 1, // zp -> Y = unperm(zp) * sigma ; X -> R (double) = X + m
  FLOW_GRAD = 2,     // dR (float) -> X = dR ; Y -> dzp = perm(Y * sigma)
  FLOW_UNWHITEN_OFFSETS = 3  // as FLOW_UNWHITEN, X -> dR (float) = X (R = dR + m)
};

hipError_t launch_flow_pre(int mode, const BasisArgs& b, const void* in, float* out, hipStream_t s);
hipError_t launch_flow_post(int mode, const BasisArgs& b, const float* in, void* out, hipStream_t s);
// the whole transform (pre, product with U or U^T, post) in one kernel, for P in {64, 128}
bool whiten_fused_supported(int P);
struct BaseArgs {
  int64_t B, offset;
  int D, A_rho;
  uint32_t key0, key1;
  const double* cdf;  // A_rho cumulative surface weights
  const double* mu;   // the non-zero feature quads of the displacements: nq * A_rho * 4
  const int* qidx;    // D / 4: index of a feature quad in mu, or -1 where every mu_a is 0 there
  const int* qlist;   // nq: the feature quad of each row of mu
  int nq;             // feature quads with a non-zero displacement
  const double* c;    // A_rho
  float* z;           // B * D
  double* logrho;     // B
};
hipError_t launch_flow_base(const BaseArgs& a, hipStream_t s);
hipError_t launch_whiten_fused(int mode, const BasisArgs& b, const void* in, void* out, const float* U,
                               int num_cu, hipStream_t s);

struct CouplingArgs {
  int64_t B;
  int D;             // flow dimension (N * P); halves of D/2
  int a_off, b_off;  // offsets of the conditioning and transformed halves (0 or D/2)
  float smax;        // s = smax * tanh(raw / smax)
  float* z;          // (B, D) state, B-half updated in place
  const float* o;    // (B, D) MLP output: [raw s | t]
  double* logdet;    // (B,) accumulated log|det|
  // backward
  const float* y;    // (B, D) saved layer input
  float* dz;         // (B, D) in: dL/d(layer output), out (B-half): dL/d(layer input)
  float* dout;       // (B, D) out: dL/d(o) = [dL/d raw s | dL/dt]
  const double* c;   // (B,) dL/d(-logdet) per sample
};

// column sums out[j] = sum_i X[i][j] in two deterministic passes; partial holds
// colsum_chunks(M) * N doubles
int64_t colsum_chunks(int64_t M);
hipError_t launch_colsum_f32(int64_t M, int64_t N, const float* X, int64_t ldx, float* out,
                             double* partial, hipStream_t s);

hipError_t launch_coupling(int inverse, const CouplingArgs& a, hipStream_t s);

// One coupling layer fused (MLP on pre-split weights + coupling + logdet), see
// pib_flow_fused.hip. Supported: hidden 256, D/2 in {256, 512}, B a multiple of 32.
struct FusedLayerArgs {
  int64_t B;
  int D, hidden, a_off, b_off, inverse;
  float smax;
  float* z;          // (B, D) flow state, transformed half updated in place
  double* logdet;    // (B,)
  // transposed (N x K) bf16 hi/lo; W2 and W3 with their inputs permuted within blocks of 32
  // (pib_split_bf16 with PIB_SPLIT_PERM32)
  const unsigned short *W1h, *W1l, *W2h, *W2l, *W3h, *W3l;
  const float *b1, *b2, *b3;
};
bool fused_layer_supported(int64_t B, int D, int hidden, bool lo);
hipError_t launch_flow_layer(const FusedLayerArgs& a, hipStream_t s);
// single-bf16 layer on fragment-ordered weights (W*h; the lo pointers unused)
bool frag_layer_supported(int64_t B, int D, int hidden);
hipError_t launch_flow_layer_frag(const FusedLayerArgs& a, hipStream_t s);
hipError_t launch_coupling_inverse_backward(const CouplingArgs& a, hipStream_t s);

struct LogRhoArgs {
  int64_t B;
  int N, P, A;            // A = sampling-model surfaces
  const double* coth;     // N
  const double* csch;     // N
  const double* pref;     // A
  const double* shift;    // A * N
  const double* R;        // (B, N, P)
  const double* c;        // (B,) loss coefficient or nullptr
  float* dR;              // (B, N, P): -c * grad log rho, or nullptr
  double* logrho;         // (B,) or nullptr
};

hipError_t launch_logrho(const LogRhoArgs& a, hipStream_t s);
hipError_t launch_adam_f32(int64_t n, float* p, const float* g, float* m, float* v, float lr,
                           float b1, float b2, float eps, int step, hipStream_t s);

}  // namespace pib
