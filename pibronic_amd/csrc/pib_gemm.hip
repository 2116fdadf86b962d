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
constexpr int kTK3 = 32, kLd3 = kTK3 + 8;  // BF16x3: K slab, LDS row of 40 bf16 = 80 B (16 B aligned)

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
  const int64_t kq = g.bf16x3 ? kTK3 : kTK;
  int s = 1;
  // split while the grid is under one work-group per CU and each split keeps >= 256 of K
  while (tiles * s * 2 <= num_cu && g.K % (kq * s * 2) == 0 && g.K / (s * 2) >= 256) s *= 2;
  return s;
}

// ---------------------------------------------------------------------------------------
// BF16x3 variant: each FP32 operand is split into bf16 hi + lo (hi = RNE(x), lo = RNE(x - hi))
// while it is staged into LDS, and C += A_hi B_hi + A_hi B_lo + A_lo B_hi on the bf16 matrix
// cores (v_mfma_f32_16x16x32_bf16) with FP32 accumulation; the dropped lo*lo term leaves
// ~2^-16 relative error per product instead of bf16's 2^-8. Same tiling as k_gemm_f32 with
// K slabs of 32. MFMA lane layout: lane l supplies A[l % 16][8 (l / 16) + e] and
// B[8 (l / 16) + e][l % 16], e = 0..7; the result layout is k_gemm_f32's.
// ---------------------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned uvec4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (operands are finite)
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) k_gemm_bf16x3(GemmArgs g) {
  __shared__ __align__(16) unsigned short Ah[kTM][kLd3], Al[kTM][kLd3];
  __shared__ __align__(16) unsigned short Bh[kTN][kLd3], Bl[kTN][kLd3];
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

  for (int64_t k0 = kbeg; k0 < kend; k0 += kTK3) {
#pragma unroll
    for (int e = tid; e < kTM * kTK3; e += 256) {
      int i, k;
      if (TA) { k = e / kTM; i = e % kTM; } else { i = e / kTK3; k = e % kTK3; }
      const int64_t gi = m0 + i, gk = k0 + k;
      const float x = TA ? g.A[gk * g.lda + gi] : g.A[gi * g.lda + gk];
      const unsigned short h = f2bf(x);
      Ah[i][k] = h;
      Al[i][k] = f2bf(x - bf2f(h));
    }
#pragma unroll
    for (int e = tid; e < kTK3 * kTN; e += 256) {
      int j, k;
      if (TB) { j = e / kTK3; k = e % kTK3; } else { k = e / kTN; j = e % kTN; }
      const int64_t gj = n0 + j, gk = k0 + k;
      const float x = TB ? g.B[gj * g.ldb + gk] : g.B[gk * g.ldb + gj];
      const unsigned short h = f2bf(x);
      Bh[j][k] = h;
      Bl[j][k] = f2bf(x - bf2f(h));
    }
    __syncthreads();
    const int kq = 8 * (lane >> 4);
    bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int r = wr * 32 + i * 16 + (lane & 15);
      ah[i] = __builtin_bit_cast(bf16x8, *(const uvec4*)&Ah[r][kq]);
      al[i] = __builtin_bit_cast(bf16x8, *(const uvec4*)&Al[r][kq]);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = wc * 32 + j * 16 + (lane & 15);
      bh[j] = __builtin_bit_cast(bf16x8, *(const uvec4*)&Bh[c][kq]);
      bl[j] = __builtin_bit_cast(bf16x8, *(const uvec4*)&Bl[c][kq]);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
      }
    __syncthreads();
  }
  if (g.partial) {
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

// Pipelined BF16x3 variant for 16-byte aligned operands: float4 global loads, and the next K
// slab is loaded into registers while the matrix cores work on the current one.
template <bool TA, bool TB>
__global__ void __launch_bounds__(256) k_gemm_bf16x3_v(GemmArgs g) {
  __shared__ __align__(16) unsigned short Ah[kTM][kLd3], Al[kTM][kLd3];
  __shared__ __align__(16) unsigned short Bh[kTN][kLd3], Bl[kTN][kLd3];
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

  // slab element groups (4 floats each) of this thread: q = tid + 256 r, r = 0, 1
  float4 ra[2], rb[2];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int q = tid + 256 * r;
      if (TA) {  // A stored K x M: 4 consecutive m
        const int k = q / 16, i = (q % 16) * 4;
        ra[r] = *(const float4*)(g.A + (k0 + k) * g.lda + m0 + i);
      } else {   // A stored M x K: 4 consecutive k
        const int i = q / 8, k = (q % 8) * 4;
        ra[r] = *(const float4*)(g.A + (m0 + i) * g.lda + k0 + k);
      }
      if (TB) {  // B stored N x K: 4 consecutive k
        const int j = q / 8, k = (q % 8) * 4;
        rb[r] = *(const float4*)(g.B + (n0 + j) * g.ldb + k0 + k);
      } else {   // B stored K x N: 4 consecutive n
        const int k = q / 16, j = (q % 16) * 4;
        rb[r] = *(const float4*)(g.B + (k0 + k) * g.ldb + n0 + j);
      }
    }
  };
  auto split = [](float x, unsigned short& h, unsigned short& l) {
    h = f2bf(x);
    l = f2bf(x - bf2f(h));
  };
  auto store = [&]() {
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int q = tid + 256 * r;
      const float av[4] = {ra[r].x, ra[r].y, ra[r].z, ra[r].w};
      const float bv[4] = {rb[r].x, rb[r].y, rb[r].z, rb[r].w};
      unsigned short h[4], l[4];
      if (TA) {
        const int k = q / 16, i = (q % 16) * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e) { split(av[e], h[e], l[e]); Ah[i + e][k] = h[e]; Al[i + e][k] = l[e]; }
      } else {
        const int i = q / 8, k = (q % 8) * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e) split(av[e], h[e], l[e]);
        *(uint2*)&Ah[i][k] = make_uint2(h[0] | ((unsigned)h[1] << 16), h[2] | ((unsigned)h[3] << 16));
        *(uint2*)&Al[i][k] = make_uint2(l[0] | ((unsigned)l[1] << 16), l[2] | ((unsigned)l[3] << 16));
      }
      if (TB) {
        const int j = q / 8, k = (q % 8) * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e) split(bv[e], h[e], l[e]);
        *(uint2*)&Bh[j][k] = make_uint2(h[0] | ((unsigned)h[1] << 16), h[2] | ((unsigned)h[3] << 16));
        *(uint2*)&Bl[j][k] = make_uint2(l[0] | ((unsigned)l[1] << 16), l[2] | ((unsigned)l[3] << 16));
      } else {
        const int k = q / 16, j = (q % 16) * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e) { split(bv[e], h[e], l[e]); Bh[j + e][k] = h[e]; Bl[j + e][k] = l[e]; }
      }
    }
  };

  load(kbeg);
  for (int64_t k0 = kbeg; k0 < kend; k0 += kTK3) {
    store();
    __syncthreads();
    if (k0 + kTK3 < kend) load(k0 + kTK3);  // in flight during the matrix-core work below
    const int kq = 8 * (lane >> 4);
    bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int r = wr * 32 + i * 16 + (lane & 15);
      ah[i] = __builtin_bit_cast(bf16x8, *(const uvec4*)&Ah[r][kq]);
      al[i] = __builtin_bit_cast(bf16x8, *(const uvec4*)&Al[r][kq]);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = wc * 32 + j * 16 + (lane & 15);
      bh[j] = __builtin_bit_cast(bf16x8, *(const uvec4*)&Bh[c][kq]);
      bl[j] = __builtin_bit_cast(bf16x8, *(const uvec4*)&Bl[c][kq]);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
      }
    __syncthreads();
  }
  if (g.partial) {
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

static bool aligned16(const void* p, int64_t ld) {
  return ((uintptr_t)p % 16) == 0 && ld % 4 == 0;
}

bool gemm_shape_ok(const GemmArgs& g) {
  const int64_t kq = g.bf16x3 ? kTK3 : kTK;
  return g.M > 0 && g.N > 0 && g.K > 0 && g.M % kTM == 0 && g.N % kTN == 0 && g.K % kq == 0;
}

hipError_t launch_gemm_f32(bool ta, bool tb, const GemmArgs& g, hipStream_t s) {
  if (!gemm_shape_ok(g)) return hipErrorInvalidValue;
  const int splits = g.partial ? (int)(g.K / g.kchunk) : 1;
  const int64_t kq = g.bf16x3 ? kTK3 : kTK;
  if (g.partial && (g.kchunk <= 0 || g.kchunk % kq || g.K % g.kchunk)) return hipErrorInvalidValue;
  const dim3 grid((unsigned)(g.N / kTN), (unsigned)(g.M / kTM), (unsigned)splits);
  if (g.bf16x3 && aligned16(g.A, g.lda) && aligned16(g.B, g.ldb)) {
    if (!ta && !tb) hipLaunchKernelGGL((k_gemm_bf16x3_v<false, false>), grid, dim3(256), 0, s, g);
    else if (!ta && tb) hipLaunchKernelGGL((k_gemm_bf16x3_v<false, true>), grid, dim3(256), 0, s, g);
    else if (ta && !tb) hipLaunchKernelGGL((k_gemm_bf16x3_v<true, false>), grid, dim3(256), 0, s, g);
    else hipLaunchKernelGGL((k_gemm_bf16x3_v<true, true>), grid, dim3(256), 0, s, g);
  } else if (g.bf16x3) {
    if (!ta && !tb) hipLaunchKernelGGL((k_gemm_bf16x3<false, false>), grid, dim3(256), 0, s, g);
    else if (!ta && tb) hipLaunchKernelGGL((k_gemm_bf16x3<false, true>), grid, dim3(256), 0, s, g);
    else if (ta && !tb) hipLaunchKernelGGL((k_gemm_bf16x3<true, false>), grid, dim3(256), 0, s, g);
    else hipLaunchKernelGGL((k_gemm_bf16x3<true, true>), grid, dim3(256), 0, s, g);
  } else if (!ta && !tb) hipLaunchKernelGGL((k_gemm_f32<false, false>), grid, dim3(256), 0, s, g);
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

// ---------------------------------------------------------------------------------------
// BF16x3 with operands already split: A as hi/lo bf16 (M x K, row-major, lda), B as hi/lo bf16
// stored transposed (N x K, row-major, ldb). Staging is a plain 16-byte copy. The result goes
// to FP32 C and/or, split again, to hi/lo bf16 (M x N, ldo) for a following product.
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_gemm_bf16x3_pre(GemmPreArgs g) {
  __shared__ __align__(16) unsigned short Ah[kTM][kLd3], Al[kTM][kLd3];
  __shared__ __align__(16) unsigned short Bh[kTN][kLd3], Bl[kTN][kLd3];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int64_t m0 = (int64_t)blockIdx.y * kTM, n0 = (int64_t)blockIdx.x * kTN;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  // one 16-byte group (8 bf16 along K) of each of the four slabs per thread
  const int row = tid / 4, kk = (tid % 4) * 8;
  uvec4 rah, ral, rbh, rbl;
  auto load = [&](int64_t k0) {
    rah = *(const uvec4*)(g.Ah + (m0 + row) * g.lda + k0 + kk);
    ral = *(const uvec4*)(g.Al + (m0 + row) * g.lda + k0 + kk);
    rbh = *(const uvec4*)(g.Bh + (n0 + row) * g.ldb + k0 + kk);
    rbl = *(const uvec4*)(g.Bl + (n0 + row) * g.ldb + k0 + kk);
  };
  load(0);
  for (int64_t k0 = 0; k0 < g.K; k0 += kTK3) {
    *(uvec4*)&Ah[row][kk] = rah;
    *(uvec4*)&Al[row][kk] = ral;
    *(uvec4*)&Bh[row][kk] = rbh;
    *(uvec4*)&Bl[row][kk] = rbl;
    __syncthreads();
    if (k0 + kTK3 < g.K) load(k0 + kTK3);
    const int kq = 8 * (lane >> 4);
    bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int r = wr * 32 + i * 16 + (lane & 15);
      ah[i] = __builtin_bit_cast(bf16x8, *(const uvec4*)&Ah[r][kq]);
      al[i] = __builtin_bit_cast(bf16x8, *(const uvec4*)&Al[r][kq]);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = wc * 32 + j * 16 + (lane & 15);
      bh[j] = __builtin_bit_cast(bf16x8, *(const uvec4*)&Bh[c][kq]);
      bl[j] = __builtin_bit_cast(bf16x8, *(const uvec4*)&Bl[c][kq]);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
      }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int64_t col = n0 + wc * 32 + j * 16 + (lane & 15);
      const float bv = g.bias ? g.bias[col] : 0.f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t rw = m0 + wr * 32 + i * 16 + 4 * (lane >> 4) + r;
        float v = acc[i][j][r] + bv;
        if (g.relu) v = v > 0.f ? v : 0.f;
        if (g.C) g.C[rw * g.ldc + col] = v;
        if (g.Oh) {
          const unsigned short h = f2bf(v);
          g.Oh[rw * g.ldo + col] = h;
          g.Ol[rw * g.ldo + col] = f2bf(v - bf2f(h));
        }
      }
    }
}

// X (rows x cols, ldx) -> hi/lo bf16; transpose: out[c][r] (cols x rows, row length rows)
// Position of input index k in a row permuted for register-resident products: within each
// block of 32, index 32b + w goes to 32b + 8((w % 16) / 4) + 4(w / 16) + w % 4 (the order in
// which two 16x16 accumulator tiles hand their rows to the next product's operand).
__device__ __forceinline__ int64_t perm32(int64_t k) {
  const int64_t b = k & ~(int64_t)31;
  const int w = (int)(k & 31);
  return b + 8 * ((w & 15) >> 2) + 4 * (w >> 4) + (w & 3);
}

// Offset of element (n, k) of the transposed matrix (N x K, N a multiple of 32, K of 16) in
// the A-operand fragment order of v_mfma_f32_32x32x16_bf16: tile m = n / 32 of 32 rows, block
// b = k / 16 of 16 inputs; lane 32((k % 16) / 8) + n % 32 holds elements k % 8 = 0..7, 16
// contiguous bytes, and the 64 lanes of a (tile, block) pair 1 KB.
__device__ __forceinline__ int64_t frag32(int64_t n, int64_t k, int64_t K) {
  const int64_t tile = n >> 5, blk = k >> 4;
  const int lane = 32 * (int)((k & 15) >> 3) + (int)(n & 31);
  return ((tile * (K >> 4) + blk) * 64 + lane) * 8 + (k & 7);
}

__global__ void k_split_bf16(int64_t rows, int64_t cols, const float* X, int64_t ldx, int mode,
                             unsigned short* hi, unsigned short* lo, int64_t ldo) {
  const bool transpose = (mode & 1) != 0, perm = (mode & 2) != 0, frag = (mode & 4) != 0;
  for (int64_t r = blockIdx.y; r < rows; r += gridDim.y) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < cols; c += (int64_t)gridDim.x * blockDim.x) {
      const float x = X[r * ldx + c];
      const unsigned short h = f2bf(x);
      const int64_t o = frag ? frag32(c, r, rows) : transpose ? c * ldo + (perm ? perm32(r) : r) : r * ldo + c;
      hi[o] = h;
      if (lo) lo[o] = f2bf(x - bf2f(h));
    }
  }
}

bool gemm_pre_shape_ok(const GemmPreArgs& g) {
  return g.M > 0 && g.N > 0 && g.K > 0 && g.M % kTM == 0 && g.N % kTN == 0 && g.K % kTK3 == 0 &&
         g.lda % 8 == 0 && g.ldb % 8 == 0 && ((uintptr_t)g.Ah % 16) == 0 && ((uintptr_t)g.Al % 16) == 0 &&
         ((uintptr_t)g.Bh % 16) == 0 && ((uintptr_t)g.Bl % 16) == 0 && (g.C || g.Oh) &&
         (!g.Oh == !g.Ol);
}

hipError_t launch_gemm_bf16x3_pre(const GemmPreArgs& g, hipStream_t s) {
  if (!gemm_pre_shape_ok(g)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_gemm_bf16x3_pre, dim3((unsigned)(g.N / kTN), (unsigned)(g.M / kTM)), dim3(256), 0, s, g);
  return hipGetLastError();
}

hipError_t launch_split_bf16(int64_t rows, int64_t cols, const float* X, int64_t ldx, int transpose,
                             unsigned short* hi, unsigned short* lo, int64_t ldo, hipStream_t s) {
  if ((transpose & ~7) || ((transpose & 2) && (!(transpose & 1) || rows % 32)) ||
      ((transpose & 4) && (transpose != 5 || cols % 32 || rows % 16)))
    return hipErrorInvalidValue;
  const dim3 grid((unsigned)std::max<int64_t>(1, std::min<int64_t>((cols + 255) / 256, 64)),
                  (unsigned)std::max<int64_t>(1, std::min<int64_t>(rows, 65535)));
  hipLaunchKernelGGL(k_split_bf16, grid, dim3(256), 0, s, rows, cols, X, ldx, transpose, hi, lo, ldo);
  return hipGetLastError();
}

}  // namespace pib
