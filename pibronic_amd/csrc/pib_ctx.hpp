// Internal: the context structure shared by the C-ABI translation units, and host helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/pibronic_hip.h"
#include "pib_common.hpp"
#include "pib_kernels.hpp"

namespace pib {
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};
}  // namespace pib

struct pib_ctx {
  int device = 0;
  int num_cu = 256;
  std::string err;
  // bumped by every setter of constants the evaluation reads (coupling, O-matrix sets, beads,
  // sampler): a batch records it and refuses to launch with buffers a setter may have freed
  uint64_t gen = 0;
  // coupling
  int A = 0, N = 0;
  pib::DevBuf energy, linear, qdiag, pair, qpair;
  pib::DevBuf w_lq, w_e, w_qpair;  // padded 16 x 16 constants of the wide kernel (A <= 16)
  pib::DevBuf energy_f, linear_f, qdiag_f, qpair_f;  // FP32 copies
  int npair = 0;
  bool has_qdiag = false;
  bool coupling_set = false;
  // O sets
  int set_A[pib::NUM_SETS] = {0, 0, 0, 0};
  int set_N[pib::NUM_SETS] = {0, 0, 0, 0};
  pib::DevBuf coth[pib::NUM_SETS], csch[pib::NUM_SETS], pref[pib::NUM_SETS], shift[pib::NUM_SETS];
  pib::DevBuf th[pib::NUM_SETS], lin[pib::NUM_SETS], k2[pib::NUM_SETS], lpref[pib::NUM_SETS];
  // host copies feeding the per-mode constant records of the lanes kernel
  std::vector<double> h_th[pib::NUM_SETS], h_cs[pib::NUM_SETS], h_lin[pib::NUM_SETS];
  std::vector<double> h_linear, h_qdiag;
  pib::DevBuf mode_tab, mode_tab_f, k2_f[pib::NUM_SETS];
  bool mode_tab_ok = false;
  pib::DevBuf tc_f[pib::NUM_SETS], cs_f[pib::NUM_SETS], lpref_f[pib::NUM_SETS];
  bool pref_positive[pib::NUM_SETS] = {false, false, false, false};
  // beads
  int P = 0;
  double tau = 0.0;
  int lanes = 0;
  // sampler
  int s_A = 0, s_N = 0, s_P = 0;
  pib::DevBuf s_cdf, s_chol, s_shift;
  bool sampler_set = false;
  // reduction scratch
  pib::DevBuf partials;
  int nparts = 0;
  double* sum_scratch = nullptr;  // this device's static partial rows (pib_block_sums_noalloc)
  // split-K GEMM and column-sum scratch (training calls)
  pib::DevBuf work;
  // flow basis
  int f_N = 0, f_P = 0;
  pib::DevBuf f_U, f_sigma, f_isigma, f_m;
  pib::DevBuf fb_mu, fb_c, fb_qidx, fb_qlist;  // flow base distribution (pib_flow_set_base): the
                                               // non-zero feature quads of mu and their index maps
  int fb_A = 0, fb_D = 0, fb_nq = 0;
};

namespace pib {

inline int fail(pib_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

inline int hip_fail(pib_ctx* c, hipError_t e, const char* what) {
  return fail(c, PIB_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define PIB_HIP(ctx, call)                                   \
  do {                                                       \
    hipError_t e_ = (call);                                  \
    if (e_ != hipSuccess) return hip_fail((ctx), e_, #call); \
  } while (0)

// (Re)allocate a device buffer and copy `bytes` from host. Synchronises the device first so
// that no queued kernel can still read the previous contents.
inline int upload(pib_ctx* c, DevBuf& b, const void* host, size_t bytes) {
  PIB_HIP(c, hipSetDevice(c->device));
  if (b.p && b.bytes < bytes) {
    PIB_HIP(c, hipDeviceSynchronize());
    PIB_HIP(c, hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
  }
  if (!b.p && bytes) {
    PIB_HIP(c, hipMalloc(&b.p, bytes));
    b.bytes = bytes;
  }
  if (bytes) {
    PIB_HIP(c, hipDeviceSynchronize());
    PIB_HIP(c, hipMemcpy(b.p, host, bytes, hipMemcpyHostToDevice));
  }
  return PIB_OK;
}

// Make b hold at least `bytes` of device memory (contents undefined). Synchronises the
// device before replacing a buffer that queued kernels may still use.
inline int reserve(pib_ctx* c, DevBuf& b, size_t bytes) {
  if (b.p && b.bytes >= bytes) return PIB_OK;
  PIB_HIP(c, hipSetDevice(c->device));
  if (b.p) {
    PIB_HIP(c, hipDeviceSynchronize());
    PIB_HIP(c, hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
  }
  PIB_HIP(c, hipMalloc(&b.p, bytes));
  b.bytes = bytes;
  return PIB_OK;
}

inline void release(DevBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

inline bool finite_all(const double* v, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return false;
  return true;
}

inline int auto_lanes(const pib_ctx* c, int64_t B, bool fp32) {
  // The fewest lanes per sample (a power of two, each doubling adds a level to the cross-lane
  // product tree) that give the launch enough waves and each lane at most kmax beads. The FP64
  // kernel at A >= 7 runs one wave per SIMD (its register budget): 8 waves per SIMD and up to 8
  // beads per lane; every other kernel 16 waves per SIMD and up to 4 beads. Measured on MI355X
  // (eval ms at 4/8/16/32 lanes): FP64 A=8, N=16, P=64, 65,536 samples 1.43/1.18/1.27/1.53,
  // 1e5 2.25/1.81/1.93/2.29; FP64 A=4, N=8, P=64, 65,536 0.62/0.30/0.19/0.21, 1e5
  // 0.98/0.46/0.29/0.33; FP64 A=6, N=12, P=128, 1e5 2.95/2.69/1.46/1.39; FP32 A=6, N=12, P=128,
  // 1e5 2.86/2.65/1.34/0.83; FP32 A=8, N=16, P=64, 65,536 1.30/0.68/0.57/0.63. Never more lanes
  // per sample than beads, so that each lane owns at least one bead.
  const bool one_wave = !fp32 && c->A >= 7;
  const int64_t target = (int64_t)c->num_cu * 4 * kWave * (one_wave ? 8 : 16);
  const int kmax = one_wave ? 8 : 4;
  int maxg = 1;
  while (maxg < 64 && maxg * 2 <= c->P) maxg *= 2;
  int g = 1;
  while (g < maxg && (B * g < target || (c->P + g - 1) / g > kmax)) g *= 2;
  return g;
}

}  // namespace pib

namespace pib {
// Validate the context for pib_eval_block-style evaluation of B samples with `flags` and fill
// every EvalArgs field except q, out, out_stride and dbg. Returns PIB_OK or an error code (with
// the message in c->err); *wide tells which kernel family applies.
int make_eval_args(pib_ctx* c, int flags, int64_t B, EvalArgs* a, bool* wide);
}  // namespace pib
