// Per-block statistics reduced on the device (FP64, deterministic order).
//
// The per-sample terms are those of pibronic/stats/stats.py:38-55:
//   r = g/rho, d1 = (g+ - g-)/rho/(2 dbeta), d2 = (g+ - 2g + g-)/rho/dbeta^2,
// and the sums (count, sum r, sum r^2, sum |r|, negatives, sum d1, sum d1^2, sum d2,
// sum d2^2, cross terms, sum g+/rho, sum g-/rho) are enough for Z_MC and its error
// (stats.py:84-90), the basic E and Cv (:93-106), the alpha-variant means (:58-81) and the
// Kish effective sample size. Stage 1: each work-group reduces a fixed strided slice in a
// fixed order into one partial row; stage 2: one work-group adds the partial rows in index
// order. No atomics, so the result does not depend on scheduling. With a learned proposal
// the denominator is its density q (log q given per sample) instead of rho.
#include <mutex>

#include "pib_common.hpp"
#include "pib_kernels.hpp"

namespace pib {

constexpr int kSumThreads = 256;

namespace {
// Partial rows for pib_block_sums_noalloc: callers that cannot allocate (graph capture, tight
// serving loops) get this fixed array instead of the context's allocated partials. Every GPU
// holds its own copy for the life of the process; all calls on one GPU share it, so such calls
// must not overlap on that GPU. k_sum_partials writes every row k_sum_final reads, so nothing
// needs clearing between calls.
static __device__ __align__(256) double scratch[kSumMaxParts * NUM_SUMS];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};
}  // namespace

hipError_t sum_scratch_address(int device, double** ptr) {
  *ptr = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    // the array's name in host code is only a handle: the device address comes from the
    // runtime, per device (the caller has made `device` current)
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<double*>(p);
  }
  *ptr = g_scratch[device];
  return hipSuccess;
}

__device__ __forceinline__ void block_tree(double* sh, double (&v)[NUM_SUMS]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int i = 0; i < NUM_SUMS; ++i) sh[i * kSumThreads + t] = v[i];
  __syncthreads();
  for (int w = kSumThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int i = 0; i < NUM_SUMS; ++i) sh[i * kSumThreads + t] += sh[i * kSumThreads + t + w];
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kSumThreads) k_sum_partials(const SumArgs a, double* partials) {
  __shared__ double sh[NUM_SUMS * kSumThreads];
  double v[NUM_SUMS];
#pragma unroll
  for (int i = 0; i < NUM_SUMS; ++i) v[i] = 0.0;
  const int64_t st = a.out_stride;
  const double idb = 1.0 / (2.0 * a.delta_beta);
  const double idb2 = 1.0 / (a.delta_beta * a.delta_beta);
  for (int64_t i = (int64_t)blockIdx.x * kSumThreads + threadIdx.x; i < a.B;
       i += (int64_t)gridDim.x * kSumThreads) {
    const double rho = a.logq ? exp(a.logq[i] - a.out[OUT_LOGS * st + i]) : a.out[OUT_RHO * st + i];
    const double g = a.out[OUT_G * st + i];
    const double r = g / rho;
    v[SUM_N] += 1.0;
    v[SUM_R] += r;
    v[SUM_R2] = fma(r, r, v[SUM_R2]);
    v[SUM_ABSR] += fabs(r);
    v[SUM_NEG] += (r < 0.0) ? 1.0 : 0.0;
    if (a.pm) {
      const double gp = a.out[OUT_GP * st + i], gm = a.out[OUT_GM * st + i];
      const double d1 = ((gp - gm) / rho) * idb;
      const double d2 = ((gp - 2.0 * g + gm) / rho) * idb2;
      v[SUM_D1] += d1;
      v[SUM_D1SQ] = fma(d1, d1, v[SUM_D1SQ]);
      v[SUM_D2] += d2;
      v[SUM_D2SQ] = fma(d2, d2, v[SUM_D2SQ]);
      v[SUM_RD1] = fma(r, d1, v[SUM_RD1]);
      v[SUM_RD2] = fma(r, d2, v[SUM_RD2]);
      v[SUM_D1D2] = fma(d1, d2, v[SUM_D1D2]);
      v[SUM_RP] += gp / rho;
      v[SUM_RM] += gm / rho;
    }
  }
  block_tree(sh, v);
  if (threadIdx.x < NUM_SUMS)
    partials[(int64_t)blockIdx.x * NUM_SUMS + threadIdx.x] = sh[threadIdx.x * kSumThreads];
}

__global__ void __launch_bounds__(kSumThreads) k_sum_final(const double* partials, int nparts,
                                                           double* sums, const int64_t* step,
                                                           int64_t step_stride, int64_t max_steps) {
  if (step) {
    const int64_t k = *step;
    if (k < 0 || k >= max_steps) return;  // uniform over the work-group
    sums += k * step_stride;
  }
  __shared__ double sh[NUM_SUMS * kSumThreads];
  double v[NUM_SUMS];
#pragma unroll
  for (int i = 0; i < NUM_SUMS; ++i) v[i] = 0.0;
  for (int b = threadIdx.x; b < nparts; b += kSumThreads) {
#pragma unroll
    for (int i = 0; i < NUM_SUMS; ++i) v[i] += partials[(int64_t)b * NUM_SUMS + i];
  }
  block_tree(sh, v);
  if (threadIdx.x < NUM_SUMS) sums[threadIdx.x] = sh[threadIdx.x * kSumThreads];
}

hipError_t launch_block_sums(const SumArgs& a, double* partials, int nparts, double* sums,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_sum_partials, dim3(nparts), dim3(kSumThreads), 0, s, a, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kSumThreads), 0, s, partials, nparts, sums, a.step,
                     a.step_stride, a.max_steps);
  return hipGetLastError();
}

}  // namespace pib
