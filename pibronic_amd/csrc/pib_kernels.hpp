// Kernel argument blocks and host-side launchers shared by the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pib {

// One TemperatureDependentClass (pibronic/pimc/pimc.py:59-89) plus the state shift of the
// model it belongs to. coth/csch are per mode (the reference broadcasts them over surfaces).
struct OSet {
  const double* coth;   // N
  const double* csch;   // N
  const double* pref;   // A_set, O-matrix prefactor F_a
  const double* shift;  // A_set * N, state shift d_aj
  // The exponent of O_a at bead p in the cancellation-free form
  //   sum_j th_j ((x_jp - d_aj)^2 + (x_jp' - d_aj)^2) + csch_j (x_jp - x_jp')^2
  //   = sum_j [th_j (x_jp^2 + x_jp'^2) + csch_j (x_jp - x_jp')^2 + lin_aj (x_jp + x_jp')] + k2_a
  // (p' = p + 1 around the ring), with th = coth - csch = tanh(tau w / 2) (the difference of
  // the two rounded constants: exact for tau w < 1.317, and the coefficient the unexpanded
  // form effectively uses), lin_aj = -2 th_j d_aj and k2_a = 2 sum_j th_j d_aj^2.
  const double* th;     // N
  const double* lin;    // A_set * N
  const double* k2;     // A_set
  const double* lpref;  // A_set, log F_a (lanes kernel: O-factors in the log domain)
};

enum : int { SET_RHO = 0, SET_VIB = 1, SET_VIB_PLUS = 2, SET_VIB_MINUS = 3, NUM_SETS = 4 };

// FP32 copies for the reduced-precision kernel: the O-matrix exponent in the
// cancellation-free form tc (x1^2 + x2^2) + cs (x1 - x2)^2 with tc = coth - csch =
// tanh(tau w / 2) and cs = csch, and log F_a.
struct OSetF {
  const float* tc;     // N
  const float* cs;     // N
  const float* lpref;  // A_set
  const float* k2;     // A_set, 2 sum_j th_j d_aj^2 (see OSet)
};

// Output rows of the per-sample result matrix (each row has out_stride entries).
enum : int { OUT_RHO = 0, OUT_G = 1, OUT_GP = 2, OUT_GM = 3, OUT_LOGS = 4, NUM_OUT = 5 };

struct EvalArgs {
  const double* q;      // (B, N, P) bead coordinates R, sample-major
  const float* qf;      // or (pib_eval_block_flow_offsets) float offsets: R = qcentre[j] + qf
  const double* qcentre;  // N
  int64_t B;
  int N, P;
  int G;                // lanes per sample (power of two, divides 64)
  int K;                // beads per lane = ceil(P / G)
  int noscale;          // 1: skip the S12 scaling (S_p = 1), as the reference's explicit test
  double tau;           // M-matrix step (always the base tau, quirk Q1)
  OSet set[NUM_SETS];
  const double* energy;   // A*A, diagonal zeroed
  const double* linear;   // N*A*A, diagonal zeroed
  const double* qdiag;    // N*A*A = 0.5*Q[j,j] or nullptr
  int npair;              // off-diagonal quadratic mode pairs (j != k)
  const int* pair;        // 2*npair (j, k)
  const double* qpair;    // npair*A*A = 0.5*Q[j,k]
  const double* mode_tab; // N records of ModeRec(A, AR).len() doubles (lanes kernel), then
                          // zero records up to a multiple of four
  const float* mode_tab_f; // N records of ModeRec(A, AR).len_f() floats (FP32 kernel), then
                           // zero records up to a multiple of four
  double* out;            // NUM_OUT rows of out_stride doubles
  int64_t out_stride;
  double* dbg;            // optional per-(sample, bead) dump, see eval_debug_record()
  // wide kernel (A <= 16): constants padded to 16 x 16, lower triangle mirrored
  int A, AR;
  const double2* wide_lq;     // N * 16 * 16 pairs {L[j,b,c], 0.5 Q[j,j,b,c]}
  const double* wide_e;       // 16 * 16
  const double* wide_qpair;   // npair * 16 * 16
  // FP32 kernel: constants in float (shifts stay double: x - d is formed in FP64)
  OSetF setf[NUM_SETS];
  const float* energy_f;
  const float* linear_f;
  const float* qdiag_f;       // or nullptr
  const float* qpair_f;
};

__host__ __device__ inline int eval_debug_record(int A, int AR) { return AR + A + 1 + A * A; }

// Per-mode constant record of the FP64 lanes kernel (mode j at mode_tab + j * record length),
// so that one mode's constants come in a few wide scalar loads from one address:
//   [0..4)              th_rho, csch_rho, th_vib, csch_vib
//   [4, 4+AR)           lin_rho[a]
//   [4+AR, 4+AR+A)      lin_vib[a]
//   next T = A(A+1)/2   L[j][b][c], b >= c (row-major lower triangle)
//   next T              0.5 Q[j][j][b][c], same order (zeros without quadratic terms)
//   next 4              th_plus, csch_plus, th_minus, csch_minus  (+-dbeta sets)
//   next 2A             lin_plus[a], lin_minus[a]
// padded to an even count (16-byte records).
struct ModeRec {
  int A, AR;
  __host__ __device__ constexpr ModeRec(int A_, int AR_) : A(A_), AR(AR_) {}
  __host__ __device__ constexpr int tri() const { return A * (A + 1) / 2; }
  __host__ __device__ constexpr int lin_rho() const { return 4; }
  __host__ __device__ constexpr int lin_vib() const { return 4 + AR; }
  __host__ __device__ constexpr int lin_q() const { return 4 + AR + A; }
  __host__ __device__ constexpr int quad() const { return lin_q() + tri(); }
  __host__ __device__ constexpr int pm() const { return quad() + tri(); }
  __host__ __device__ constexpr int lin_plus() const { return pm() + 4; }
  __host__ __device__ constexpr int lin_minus() const { return lin_plus() + A; }
  __host__ __device__ constexpr int len() const { return (lin_minus() + A + 1) & ~1; }
  __host__ __device__ constexpr int len_f() const { return (lin_minus() + A + 3) & ~3; }  // float copy
};

struct SampleArgs {
  int64_t B;
  int N, P, A_rho;
  int64_t offset;         // global index of this block's first sample (Philox counter)
  uint32_t key0, key1;    // Philox key (seed)
  const double* cdf;      // A_rho cumulative surface weights
  const double* chol;     // N * 3P: per mode 1/diag(L), subdiagonal of L, last row of L, where
                          // L L^T = 2 coth_j I - csch_j C (cyclic tridiagonal Cholesky)
  const double* shift;    // A_rho * N sampling-model state shift
  double* q;              // (B, N, P) output
  int32_t* src;           // (B,) surface sources, optional
  const int64_t* step;    // optional device block counter: the offset becomes offset + step * B
};

enum : int {
  SUM_N = 0, SUM_R, SUM_R2, SUM_ABSR, SUM_NEG, SUM_D1, SUM_D1SQ, SUM_D2, SUM_D2SQ,
  SUM_RD1, SUM_RD2, SUM_D1D2, SUM_RP, SUM_RM, NUM_SUMS
};

struct SumArgs {
  const double* out;
  int64_t out_stride;
  int64_t B;
  int pm;
  double delta_beta;
  const double* logq;  // nullptr: denominator rho (row OUT_RHO); else the proposal density
                       // q = exp(logq), rescaled like g: exp(logq - OUT_LOGS)
  const int64_t* step;  // optional device block counter: the sums go to sums + step * step_stride
  int64_t step_stride;  // and are skipped when step >= max_steps
  int64_t max_steps;
};

// Everything the lanes evaluation reads of one problem, passed by value in the kernel
// arguments of the batched lanes kernel (at most kBatchLaunchProbs problems per launch), so
// that every constant sits in scalar registers as in the stand-alone launch. Read from global
// memory instead, the pointers lose their address space and the constants come through
// vector loads (about 90 more VGPRs at A = 6, half the occupancy).
constexpr int kBatchLaunchProbs = 16;
struct LaneProb {
  const double* q;
  double* out;
  int64_t out_stride;
  int64_t B;
  int N, P, G, K, npair, noscale;
  double tau;
  const double* mode_tab;
  const double* energy;
  const double* qdiag;
  const int* pair;
  const double* qpair;
  const double* k2[NUM_SETS];
  const double* lpref[NUM_SETS];
};
struct LaneProbs {
  int n;
  LaneProb p[kBatchLaunchProbs];
};

// One tile of a batched launch: waves [wave0, wave0 + nwaves) of problem `prob`.
struct BatchTile {
  int prob;
  int nwaves;
  int64_t wave0;
};

// launchers (return hipError_t of the launch)
hipError_t launch_eval_fp64(int A, int AR, int pm, const EvalArgs& a, hipStream_t s);
hipError_t launch_eval_fp64_offsets(int A, const EvalArgs& a, hipStream_t s);
hipError_t launch_eval_wide(int pm, const EvalArgs& a, hipStream_t s, bool bf16x3 = false);
hipError_t launch_eval_fp32(int A, int AR, int pm, const EvalArgs& a, hipStream_t s);
hipError_t launch_eval_batched(int A, int AR, int pm, const LaneProbs& probs, const BatchTile* tiles,
                               int ntiles, int* counter, int waves, hipStream_t s);
LaneProb lane_prob(const EvalArgs& a);
hipError_t launch_eval_wide_batched(int pm, const EvalArgs* probs, const BatchTile* tiles, int ntiles,
                                    int* counter, int waves, hipStream_t s);
constexpr int kWideMaxStates = 16;
hipError_t launch_sample_fp64(const SampleArgs& a, hipStream_t s);
hipError_t launch_block_sums(const SumArgs& a, double* partials, int nparts, double* sums,
                             hipStream_t s);
// Partial rows of the block sums: one per work-group of k_sum_partials, at most this many
// whatever B is (the work-groups stride over the block).
constexpr int kSumMaxParts = 1024;
inline int sum_parts(int64_t B) {
  const int64_t n = (B + 255) / 256;
  return (int)(n < 1 ? 1 : n > kSumMaxParts ? kSumMaxParts : n);
}
// Device address, on the current device, of the file-scope partial-row array that
// pib_block_sums_noalloc uses instead of allocated memory (kSumMaxParts * NUM_SUMS doubles per
// device). Resolved once per device; the current device must be `device`.
hipError_t sum_scratch_address(int device, double** ptr);
bool eval_supported(int A, int AR);

}  // namespace pib
