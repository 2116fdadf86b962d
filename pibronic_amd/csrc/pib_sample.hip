// Harmonic-rho bead sampler (draw_sample + transform_sampled_coordinates).
//
// Replaces pibronic/pimc/pimc.py:326-334 (normal draws with the source surface's standard
// deviation), :386-389 (categorical surface source) and :613-631 (R = U_circ . cc + d_src).
// Random numbers come from Philox4x32-10 keyed by the run seed with the counter
// (global sample, mode, pair index); Box-Muller turns two 53-bit uniforms into two normals.
// The surface source of each sample is drawn from its own counter (w = 1), per sample
// rather than once per run (the reference's quirk Q5 draws all sources up front; the
// distribution is the same).
//
// Layout: one work-group handles a strip of rows (sample, mode); for each row the P normal
// draws are written to LDS, then lane p forms R_p = sum_k U[p][k] cc_k with U read
// transposed so that a wave's loads of one k are contiguous.
#include "pib_common.hpp"
#include "pib_kernels.hpp"

namespace pib {

constexpr int kSampleThreads = 256;
constexpr int kMaxLdsRow = 4096;  // doubles of LDS for normal draws per work-group (32 KB)

__device__ __forceinline__ int draw_source(const SampleArgs& a, int64_t gsample) {
  if (a.A_rho == 1) return 0;
  const u32x4 r = philox4x32_10(u32x4{(uint32_t)gsample, (uint32_t)(gsample >> 32), 0u, 1u},
                                a.key0, a.key1);
  const double u = u53(r.x, r.y);
  int s = 0;
  while (s < a.A_rho - 1 && u >= a.cdf[s]) ++s;
  return s;
}

__global__ void __launch_bounds__(kSampleThreads) k_sample_fp64(const SampleArgs a) {
  __shared__ double cc[kMaxLdsRow];
  const int P = a.P, N = a.N;
  const int64_t rows = a.B * (int64_t)N;
  // rows handled together: as many as fit in the block's threads (one thread per bead)
  const int rpi = max(1, min(kSampleThreads / P, kMaxLdsRow / P));
  const int tid = threadIdx.x;
  for (int64_t r0 = (int64_t)blockIdx.x * rpi; r0 < rows; r0 += (int64_t)gridDim.x * rpi) {
    // 1. normal draws cc[row][k] = sigma[src][j][k] * z
    for (int e = tid; e < rpi * P; e += blockDim.x) {
      const int lr = e / P, k = e - lr * P;
      const int64_t row = r0 + lr;
      if (row < rows) {
        const int64_t s = row / N;
        const int j = (int)(row - s * N);
        const int64_t gs = a.offset + s;
        const int src = draw_source(a, gs);
        const uint64_t grow = (uint64_t)gs * (uint64_t)N + (uint64_t)j;
        const u32x4 r = philox4x32_10(
            u32x4{(uint32_t)grow, (uint32_t)(grow >> 32), (uint32_t)(k >> 1), 0u}, a.key0, a.key1);
        const double u1 = 1.0 - u53(r.x, r.y);  // (0, 1]
        const double u2 = u53(r.z, r.w);
        const double rad = sqrt(-2.0 * log(u1));
        double sn, cs;
        sincospi(2.0 * u2, &sn, &cs);
        const double z = (k & 1) ? rad * sn : rad * cs;
        cc[lr * P + k] = a.sigma[((int64_t)src * N + j) * P + k] * z;
      }
    }
    __syncthreads();
    // 2. R_p = sum_k U[p][k] cc_k + d_src,j
    for (int e = tid; e < rpi * P; e += blockDim.x) {
      const int lr = e / P, p = e - lr * P;
      const int64_t row = r0 + lr;
      if (row < rows) {
        const int64_t s = row / N;
        const int j = (int)(row - s * N);
        const int src = draw_source(a, a.offset + s);
        const double* c = cc + lr * P;
        double acc = 0.0;
        for (int k = 0; k < P; ++k) acc = fma(a.Ut[(int64_t)k * P + p], c[k], acc);
        a.q[row * P + p] = acc + a.shift[src * N + j];
        if (a.src && j == 0 && p == 0) a.src[s] = src;
      }
    }
    __syncthreads();
  }
}

hipError_t launch_sample_fp64(const SampleArgs& a, hipStream_t s) {
  if (a.P > kMaxLdsRow) return hipErrorInvalidValue;
  const int rpi = max(1, min(kSampleThreads / a.P, kMaxLdsRow / a.P));
  const int64_t rows = a.B * (int64_t)a.N;
  int64_t groups = (rows + rpi - 1) / rpi;
  if (groups > 8192) groups = 8192;
  if (groups < 1) groups = 1;
  hipLaunchKernelGGL(k_sample_fp64, dim3((unsigned)groups), dim3(kSampleThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace pib
