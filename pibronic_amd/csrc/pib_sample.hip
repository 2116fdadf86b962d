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
// Random numbers: Philox4x32-10 keyed by the run seed. The beads 4m..4m+3 of row (global
// sample s, mode j) take the four normals of the counter (s*N + j, m, 0): two Box-Muller
// pairs, each from a 32-bit radius uniform and a 32-bit angle uniform (FP32 transform, FP64
// recurrence); a radius uniform below 2^-20 is refined to 64 bits by the counter
// (s*N + j, m, 2), so that the tails reach ~9.4 sigma. The surface source of sample s uses
// the counter (s, 0, 1).
// Sources are drawn per sample rather than once per run (quirk Q5 of the reference draws them
// all up front; same distribution).
//
// Layout: a wave takes one mode of 64 consecutive samples, one lane per row (sample, mode), so
// the mode's factor entries are wave-uniform scalar loads. Each lane walks its beads
// downwards in chunks of 8; a chunk is staged in LDS and the wave then writes the 64 rows'
// chunk with consecutive lanes on consecutive addresses (64-byte row segments).
#include "pib_common.hpp"
#include "pib_kernels.hpp"

namespace pib {

constexpr int kSampleWaves = 4;   // waves per work-group, each independent
constexpr int kChunk = 8;         // beads staged per lane before a coalesced write-out (64 B
                                  // per row: LDS for ~6 waves per SIMD)
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

// the four normals of beads 4m..4m+3 of row grow: two Box-Muller pairs from one Philox call
// in FP32 with the hardware log2, square root and sin/cos (v_sin_f32 / v_cos_f32 take their
// argument in revolutions), widened to FP64. The radius uniform below 2^-20 is refined to 64
// bits (rounded to FP32 before its log2), so the tails reach ~9.4 sigma instead of the
// 6.7 sigma of a 32-bit uniform. The normals carry ~1e-7 relative error against an exact transform of the same
// uniforms: a shift of the sampled distribution far below the statistical error of any block
// (the estimator evaluates rho exactly at the coordinates drawn).
__device__ __forceinline__ void normal_quad(uint64_t grow, int m, uint32_t k0, uint32_t k1, double (&z)[4]) {
  const u32x4 r = philox4x32_10(u32x4{(uint32_t)grow, (uint32_t)(grow >> 32), (uint32_t)m, 0u}, k0, k1);
  float l2a = __log2f(((float)r.x + 0.5f) * 0x1.0p-32f);
  float l2b = __log2f(((float)r.z + 0.5f) * 0x1.0p-32f);
  if (r.x < (1u << 12) || r.z < (1u << 12)) {  // rare: P = 2^-19 per call
    const u32x4 t = philox4x32_10(u32x4{(uint32_t)grow, (uint32_t)(grow >> 32), (uint32_t)m, 2u}, k0, k1);
    if (r.x < (1u << 12)) l2a = __log2f((float)(((double)r.x * 0x1.0p32 + (double)t.x + 0.5) * 0x1.0p-64));
    if (r.z < (1u << 12)) l2b = __log2f((float)(((double)r.z * 0x1.0p32 + (double)t.z + 0.5) * 0x1.0p-64));
  }
  const float ra = __fsqrt_rn(-1.3862943611198906f * l2a), rb = __fsqrt_rn(-1.3862943611198906f * l2b);
  const float ua = ((float)r.y + 0.5f) * 0x1.0p-32f, ub = ((float)r.w + 0.5f) * 0x1.0p-32f;
  z[0] = (double)(ra * __builtin_amdgcn_cosf(ua));
  z[1] = (double)(ra * __builtin_amdgcn_sinf(ua));
  z[2] = (double)(rb * __builtin_amdgcn_cosf(ub));
  z[3] = (double)(rb * __builtin_amdgcn_sinf(ub));
}

__global__ void __launch_bounds__(kWave * kSampleWaves) k_sample_fp64(const SampleArgs a) {
  __shared__ double stage_all[kSampleWaves][kWave * kLd];
  const int P = a.P, N = a.N;
  const int64_t groups = (a.B + kWave - 1) / kWave;  // groups of 64 consecutive samples
  const int64_t units = groups * N;                  // (sample group, mode) pairs
  const int lane = threadIdx.x & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);  // wave-uniform
  double* stage = stage_all[w];
  const int64_t offset = a.step ? a.offset + *a.step * a.B : a.offset;
  for (int64_t u = (int64_t)blockIdx.x * kSampleWaves + w; u < units; u += (int64_t)gridDim.x * kSampleWaves) {
    // wave-uniform: the sample group and the mode; every lane of the wave runs the same chunks
    const int64_t sg = u / N;
    const int j = (int)(u - sg * N);
    const int64_t s0 = sg * kWave;
    const int64_t s = s0 + lane;
    const bool live = s < a.B;
    const int64_t gs = offset + (live ? s : s0);
    const int src = live ? draw_source(a, gs) : 0;
    const uint64_t grow = (uint64_t)gs * (uint64_t)N + (uint64_t)j;
    // the mode's factor entries are the same for the whole wave: read through the constant
    // address space (the kernel never writes them), so they come by scalar loads
    typedef const double __attribute__((address_space(4))) cdouble;
    cdouble* dinv = (cdouble*)(a.chol + (int64_t)j * 3 * P);
    cdouble* sub = dinv + P;
    cdouble* last = sub + P;
    const double d = a.shift[src * N + j];
    if (live && a.src && j == 0) a.src[s] = src;
    const int nrow = (a.B - s0) < kWave ? (int)(a.B - s0) : kWave;
    double zq[4], xl = 0.0, xn = 0.0;
    // beads in descending order, in chunks [kb, ke) aligned to multiples of kChunk, and within
    // a chunk quad by quad. One formula for every bead: the host stores sub[k] = 0 for
    // k >= P-2 and last[P-1] = 0, and x_{P-1} (xl) and x_{k+1} (xn) start at 0, so bead P-1
    // gives z dinv and bead P-2 (z - last xl) dinv exactly.
    // Lanes past the block's end (the last group only) run the same uniform loop on the
    // group's first sample; their rows are never written out.
    for (int kb = ((P - 1) / kChunk) * kChunk; kb >= 0; kb -= kChunk) {
      const int ke = kb + kChunk < P ? kb + kChunk : P;
#pragma unroll 2
      for (int m = (ke - 1) >> 2; m >= (kb >> 2); --m) {
        double cd[4], cs[4], cl[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int k = 4 * m + e < P ? 4 * m + e : P - 1;  // the top quad may be partial
          cd[e] = dinv[k];
          cs[e] = sub[k];
          cl[e] = last[k];
        }
        normal_quad(grow, m, a.key0, a.key1, zq);
#pragma unroll
        for (int e = 3; e >= 0; --e) {
          const int k = 4 * m + e;
          if (k < ke) {  // uniform over the wave
            const double x = fma(-cl[e], xl, fma(-cs[e], xn, zq[e])) * cd[e];
            if (k == P - 1) xl = x;
            xn = x;
            stage[lane * kLd + (k - kb)] = x + d;
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // write the chunk: row rr (sample s0 + rr, mode j) has its beads [kb, ke) contiguous in q
      const int w_ = ke - kb;
      double* __restrict__ qj = a.q + (s0 * N + j) * (int64_t)P + kb;
      const int64_t rstride = (int64_t)N * P;
      if (w_ == kChunk && nrow == kWave) {  // full chunk: no divisions
#pragma unroll 4
        for (int it = 0; it < kChunk; ++it) {
          const int e = lane + kWave * it, rr = e / kChunk, kk = e % kChunk;
          qj[rr * rstride + kk] = stage[rr * kLd + kk];
        }
      } else {
        for (int e = lane; e < nrow * w_; e += kWave) {
          const int rr = e / w_, kk = e - rr * w_;
          qj[rr * rstride + kk] = stage[rr * kLd + kk];
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
}

hipError_t launch_sample_fp64(const SampleArgs& a, hipStream_t s) {
  const int64_t units = (a.B + kWave - 1) / kWave * a.N;  // one wave each
  int64_t groups = (units + kSampleWaves - 1) / kSampleWaves;
  if (groups > 65536) groups = 65536;
  if (groups < 1) groups = 1;
  hipLaunchKernelGGL(k_sample_fp64, dim3((unsigned)groups), dim3(kWave * kSampleWaves), 0, s, a);
  return hipGetLastError();
}

}  // namespace pib
