// FP32 GEMM on MFMA (v_mfma_f32_16x16x4f32, FP32 inputs, FP32 accumulation) for the flow's
// dense layers and the normal-mode whitening transforms.
//
//   C[M x N] = op(A)[M x K] * op(B)[K x N]  (+ bias[N]) (relu) (* [mask > 0])
//   op(A)(i,k) = TA ? A[k*lda + i] : A[i*lda + k],  op(B)(k,j) = TB ? B[j*ldb + k] : B[k*ldb + j]
//
// Tiling: a 256-thread work-group computes a 64 x 64 tile of C; each of its four waves a
// 32 x 32 sub-tile as 2 x 2 MFMA blocks of 16 x 16. K advances in slabs of 16 staged in LDS
// (k-major, so a lane's A and B fragments are single LDS reads). MFMA 16x16x4 lane layout
// (one block, 64 lanes): lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]; the result
// D[4 * (l / 16) + r][l % 16] is element r of the lane's 4-vector.
// Shapes must be multiples of the tile (M, N of 64, K of 16); the launcher checks.
#include "pib_common.hpp"
#include "pib_flow.hpp"

#include <algorithm>

namespace pib {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTM = 64, kTN = 64, kTK = 16, kPad = 4;

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) k_gemm_f32(GemmArgs g) {
  __shared__ float As[kTK][kTM + kPad];
  __shared__ float Bs[kTK][kTN + kPad];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int64_t m0 = (int64_t)blockIdx.y * kTM, n0 = (int64_t)blockIdx.x * kTN;
  const int64_t kbeg = g.partial ? (int64_t)blockIdx.z * g.kchunk : 0;
  const int64_t kend = g.partial ? kbeg + g.kchunk : g.K;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int64_t k0 = kbeg; k0 < kend; k0 += kTK) {
    // stage A (kTM x kTK) and B (kTK x kTN): 1024 elements each, 4 per thread
#pragma unroll
    for (int e = tid; e < kTM * kTK; e += 256) {
      int i, k;
      if (TA) { k = e / kTM; i = e % kTM; } else { i = e / kTK; k = e % kTK; }
      const int64_t gi = m0 + i, gk = k0 + k;
      As[k][i] = TA ? g.A[gk * g.lda + gi] : g.A[gi * g.lda + gk];
    }
#pragma unroll
    for (int e = tid; e < kTK * kTN; e += 256) {
      int j, k;
      if (TB) { j = e / kTK; k = e % kTK; } else { k = e / kTN; j = e % kTN; }
      const int64_t gj = n0 + j, gk = k0 + k;
      Bs[k][j] = TB ? g.B[gj * g.ldb + gk] : g.B[gk * g.ldb + gj];
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTK; kk += 4) {
      float a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[kk + (lane >> 4)][wr * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[kk + (lane >> 4)][wc * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  if (g.partial) {  // split-K: raw tile, epilogue in k_gemm_reduce
    float* P = g.partial + (int64_t)blockIdx.z * g.M * g.N;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int64_t col = n0 + wc * 32 + j * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = m0 + wr * 32 + i * 16 + 4 * (lane >> 4) + r;
          P[row * g.N + col] = acc[i][j][r];
        }
      }
    return;
  }
  // epilogue
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int64_t col = n0 + wc * 32 + j * 16 + (lane & 15);
      const float bv = g.bias ? g.bias[col] : 0.f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wr * 32 + i * 16 + 4 * (lane >> 4) + r;
        float v = acc[i][j][r] + bv;
        if (g.relu) v = v > 0.f ? v : 0.f;
        if (g.mask) v = g.mask[row * g.ldmask + col] > 0.f ? v : 0.f;
        if (g.accumulate) v += g.C[row * g.ldc + col];
        g.C[row * g.ldc + col] = v;
      }
    }
}

// sum of the K splits in split order, then the same epilogue as k_gemm_f32
__global__ void k_gemm_reduce(GemmArgs g, int splits) {
  const int64_t MN = g.M * g.N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < MN;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = e / g.N, col = e - row * g.N;
    float v = g.partial[e];
    for (int z = 1; z < splits; ++z) v += g.partial[z * MN + e];
    if (g.bias) v += g.bias[col];
    if (g.relu) v = v > 0.f ? v : 0.f;
    if (g.mask) v = g.mask[row * g.ldmask + col] > 0.f ? v : 0.f;
    if (g.accumulate) v += g.C[row * g.ldc + col];
    g.C[row * g.ldc + col] = v;
  }
}

int gemm_splits(const GemmArgs& g, int num_cu) {
  const int64_t tiles = (g.M / kTM) * (g.N / kTN);
  int s = 1;
  // split while the grid is under one work-group per CU and each split keeps >= 256 of K
  while (tiles * s * 2 <= num_cu && g.K % (kTK * s * 2) == 0 && g.K / (s * 2) >= 256) s *= 2;
  return s;
}

bool gemm_shape_ok(const GemmArgs& g) {
  return g.M > 0 && g.N > 0 && g.K > 0 && g.M % kTM == 0 && g.N % kTN == 0 && g.K % kTK == 0;
}

hipError_t launch_gemm_f32(bool ta, bool tb, const GemmArgs& g, hipStream_t s) {
  if (!gemm_shape_ok(g)) return hipErrorInvalidValue;
  const int splits = g.partial ? (int)(g.K / g.kchunk) : 1;
  if (g.partial && (g.kchunk <= 0 || g.kchunk % kTK || g.K % g.kchunk)) return hipErrorInvalidValue;
  const dim3 grid((unsigned)(g.N / kTN), (unsigned)(g.M / kTM), (unsigned)splits);
  if (!ta && !tb) hipLaunchKernelGGL((k_gemm_f32<false, false>), grid, dim3(256), 0, s, g);
  else if (!ta && tb) hipLaunchKernelGGL((k_gemm_f32<false, true>), grid, dim3(256), 0, s, g);
  else if (ta && !tb) hipLaunchKernelGGL((k_gemm_f32<true, false>), grid, dim3(256), 0, s, g);
  else hipLaunchKernelGGL((k_gemm_f32<true, true>), grid, dim3(256), 0, s, g);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !g.partial) return e;
  const int64_t MN = g.M * g.N;
  const unsigned blocks = (unsigned)std::min<int64_t>((MN + 255) / 256, 4096);
  hipLaunchKernelGGL(k_gemm_reduce, dim3(blocks), dim3(256), 0, s, g, splits);
  return hipGetLastError();
}

}  // namespace pib
