// Harmonic-rho bead sampler (draw_sample + transform_sampled_coordinates).
//
// Replaces pibronic/pimc/pimc.py:326-334 (normal draws with the source surface's standard
// deviations), :386-389 (categorical surface source) and :613-631 (R = U_circ . cc + d_src).
//
// Distribution: for mode j the bead coordinates of the reference are Gaussian with mean
// d_src,j and covariance Lambda_j^-1, Lambda_j = 2 coth_j I - csch_j C (C the ring adjacency
// matrix; pimc.py:401-404 lists its eigenvalues), independent of the surface. The reference
// draws in C's eigenbasis (a dense P x P transform). Here Lambda_j = L_j L_j^T is factored
// once on the host: a cyclic tridiagonal matrix has a Cholesky factor with only the diagonal,
// the subdiagonal and the last row non-zero, so x = L_j^-T z, z ~ N(0, I), is an exact draw
// from N(0, Lambda_j^-1) by an O(P) back substitution (the standard sampler for Gaussian
// Markov random fields, Rue 2001). Same distribution, O(P) instead of O(P^2) per row.
//
// Random numbers: Philox4x32-10 keyed by the run seed. The normal pair m of row
// (global sample s, mode j) uses the counter (s*N + j, m, 0): one call, two 53-bit uniforms,
// two normals by Box-Muller. The surface source of sample s uses the counter (s, 0, 1).
// Sources are drawn per sample rather than once per run (quirk Q5 of the reference draws them
// all up front; same distribution).
//
// Layout: one lane per row (sample, mode), 64 consecutive rows per wave. Each lane walks its
// beads downwards in chunks of 16; a chunk is staged in LDS and the wave then writes the 64
// rows' chunk with consecutive lanes on consecutive addresses (full cache lines).
#include "pib_common.hpp"
#include "pib_kernels.hpp"

namespace pib {

constexpr int kSampleWaves = 4;   // waves per work-group, each independent
constexpr int kChunk = 16;        // beads staged per lane before a coalesced write-out
constexpr int kLd = kChunk + 1;   // padded LDS row

__device__ __forceinline__ int draw_source(const SampleArgs& a, int64_t gsample) {
  if (a.A_rho == 1) return 0;
  const u32x4 r = philox4x32_10(u32x4{(uint32_t)gsample, (uint32_t)(gsample >> 32), 0u, 1u},
                                a.key0, a.key1);
  const double u = u53(r.x, r.y);
  int s = 0;
  while (s < a.A_rho - 1 && u >= a.cdf[s]) ++s;
  return s;
}

__device__ __forceinline__ void normal_pair(uint64_t grow, int m, uint32_t k0, uint32_t k1,
                                            double& ze, double& zo) {
  const u32x4 r = philox4x32_10(u32x4{(uint32_t)grow, (uint32_t)(grow >> 32), (uint32_t)m, 0u}, k0, k1);
  const double u1 = 1.0 - u53(r.x, r.y);  // (0, 1]
  const double u2 = u53(r.z, r.w);
  const double rad = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincospi(2.0 * u2, &sn, &cs);
  ze = rad * cs;
  zo = rad * sn;
}

__global__ void __launch_bounds__(kWave * kSampleWaves) k_sample_fp64(const SampleArgs a) {
  __shared__ double stage_all[kSampleWaves][kWave * kLd];
  const int P = a.P, N = a.N;
  const int64_t rows = a.B * (int64_t)N;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  double* stage = stage_all[w];
  const int64_t step = (int64_t)gridDim.x * kSampleWaves * kWave;
  const int64_t offset = a.step ? a.offset + *a.step * a.B : a.offset;
  for (int64_t r0 = ((int64_t)blockIdx.x * kSampleWaves + w) * kWave; r0 < rows; r0 += step) {
    // wave-uniform loop: every lane of the wave runs the same chunks
    const int64_t row = r0 + lane;
    const bool live = row < rows;
    const int64_t s = live ? row / N : 0;
    const int j = live ? (int)(row - s * N) : 0;
    const int64_t gs = offset + s;
    const int src = live ? draw_source(a, gs) : 0;
    const uint64_t grow = (uint64_t)gs * (uint64_t)N + (uint64_t)j;
    const double* __restrict__ dinv = a.chol + (int64_t)j * 3 * P;
    const double* __restrict__ sub = dinv + P;
    const double* __restrict__ last = sub + P;
    const double d = a.shift[src * N + j];
    if (live && a.src && j == 0) a.src[s] = src;
    const int64_t nrow = (rows - r0) < kWave ? (rows - r0) : kWave;
    double ze = 0.0, zo = 0.0, xl = 0.0, xn = 0.0;
    // beads in descending order, in chunks [kb, ke)
    for (int ke = P; ke > 0; ke -= kChunk) {
      const int kb = ke - kChunk > 0 ? ke - kChunk : 0;
      if (live) {
        for (int k = ke - 1; k >= kb; --k) {
          if (k == P - 1 || (k & 1)) normal_pair(grow, k >> 1, a.key0, a.key1, ze, zo);
          const double z = (k & 1) ? zo : ze;
          double x;
          if (k == P - 1) {
            x = z * dinv[k];
            xl = x;
          } else if (k == P - 2) {
            x = fma(-last[k], xl, z) * dinv[k];
          } else {
            x = fma(-last[k], xl, fma(-sub[k], xn, z)) * dinv[k];
          }
          xn = x;
          stage[lane * kLd + (k - kb)] = x + d;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // write the chunk: row rr's beads [kb, ke) are contiguous in q
      const int w_ = ke - kb;
      for (int e = lane; e < nrow * w_; e += kWave) {
        const int rr = e / w_, kk = e - rr * w_;
        a.q[(r0 + rr) * P + kb + kk] = stage[rr * kLd + kk];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
}

hipError_t launch_sample_fp64(const SampleArgs& a, hipStream_t s) {
  const int64_t rows = a.B * (int64_t)a.N;
  const int64_t per = (int64_t)kSampleWaves * kWave;
  int64_t groups = (rows + per - 1) / per;
  if (groups > 65536) groups = 65536;
  if (groups < 1) groups = 1;
  hipLaunchKernelGGL(k_sample_fp64, dim3((unsigned)groups), dim3(kWave * kSampleWaves), 0, s, a);
  return hipGetLastError();
}

}  // namespace pib
