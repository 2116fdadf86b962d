// One affine-coupling layer in one kernel (flow sampling and density directions).
//
// Per work-group: 32 samples (rows of the flow state zp). The conditioning half is split into
// bf16 hi/lo in LDS; the layer's MLP (H -> h -> h -> D, ReLU, ReLU) runs as BF16x3 products on
// the bf16 matrix cores with FP32 accumulation, hidden activations staying in LDS as hi/lo
// pairs; the weights (split once per parameter update, stored transposed N x K) stream from
// L2 straight into MFMA operand registers. The last product's epilogue applies the coupling
// in registers -- s = smax tanh(raw / smax), y = z e^s + t (or z = (y - t) e^-s) -- updates
// the transformed half in place and adds sum s to the row's log-determinant. Identical
// arithmetic to the layer-by-layer path (same splits, same product order), so the result is
// the same bit for bit; what changes is that o and the activations never go to HBM.
//
// Four waves split the output columns; MFMA tile 16x16x32 (lane l: A[l%16][8(l/16)+e],
// B[8(l/16)+e][l%16], D[4(l/16)+r][l%16]).
#include <hip/hip_runtime.h>

#include "pib_common.hpp"
#include "pib_flow.hpp"

namespace pib {

namespace {

typedef float f32x4f __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8f __attribute__((ext_vector_type(8)));
typedef unsigned uvec4f __attribute__((ext_vector_type(4)));

constexpr int kRows = 32;   // samples per work-group
constexpr int kPadF = 8;    // LDS row padding (bf16 elements)

__device__ __forceinline__ unsigned short f2bf_f(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f_f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ bf16x8f ld8(const unsigned short* p) {
  return __builtin_bit_cast(bf16x8f, *(const uvec4f*)p);
}

// acc[rb][cb] += X[rows 16 rb..][0..K) * W^T for the wave's column blocks; X hi/lo in LDS with
// row stride ldx, W^T hi/lo in global memory (N x K, row stride K), columns col0 + 16 cb.
template <int NB>
__device__ __forceinline__ void mlp_product(f32x4f (&acc)[2][NB], const unsigned short* Xh,
                                            const unsigned short* Xl, int ldx, const unsigned short* Wh,
                                            const unsigned short* Wl, int K, int col0, int lane) {
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) acc[rb][cb] = f32x4f{0.f, 0.f, 0.f, 0.f};
  const int kq = 8 * (lane >> 4), li = lane & 15;
  // weight fragments double-buffered in registers: the next K step's loads are in flight
  // while the matrix cores work on the current one
  bf16x8f bh[NB], bl[NB];
#pragma unroll
  for (int cb = 0; cb < NB; ++cb) {
    const int64_t wrow = (int64_t)(col0 + 16 * cb + li) * K + kq;
    bh[cb] = ld8(Wh + wrow);
    bl[cb] = ld8(Wl + wrow);
  }
  for (int k0 = 0; k0 < K; k0 += 32) {
    bf16x8f nh[NB], nl[NB];
    const int kn = (k0 + 32 < K) ? k0 + 32 : k0;  // last step reloads its own slab (in bounds)
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) {
      const int64_t wrow = (int64_t)(col0 + 16 * cb + li) * K + kn + kq;
      nh[cb] = ld8(Wh + wrow);
      nl[cb] = ld8(Wl + wrow);
    }
    bf16x8f ah[2], al[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
      ah[rb] = ld8(Xh + (16 * rb + li) * ldx + k0 + kq);
      al[rb] = ld8(Xl + (16 * rb + li) * ldx + k0 + kq);
    }
#pragma unroll
    for (int cb = 0; cb < NB; ++cb)
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rb], bh[cb], acc[rb][cb], 0, 0, 0);
        acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], bl[cb], acc[rb][cb], 0, 0, 0);
        acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], bh[cb], acc[rb][cb], 0, 0, 0);
      }
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) {
      bh[cb] = nh[cb];
      bl[cb] = nl[cb];
    }
  }
}

// bias + ReLU, then split into the LDS activation buffer (row stride ldy)
template <int NB>
__device__ __forceinline__ void store_hidden(const f32x4f (&acc)[2][NB], const float* bias,
                                             unsigned short* Yh, unsigned short* Yl, int ldy, int col0,
                                             int lane) {
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) {
      const int col = col0 + 16 * cb + (lane & 15);
      const float bv = bias[col];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * rb + 4 * (lane >> 4) + r;
        float v = acc[rb][cb][r] + bv;
        v = v > 0.f ? v : 0.f;
        const unsigned short h = f2bf_f(v);
        Yh[row * ldy + col] = h;
        Yl[row * ldy + col] = f2bf_f(v - bf2f_f(h));
      }
    }
}

// NBH: column blocks per wave of the hidden products (h = 64 NBH);
// NBD: column blocks per wave of each half of the last product (H = 64 NBD)
template <int NBH, int NBD>
__global__ void __launch_bounds__(256) k_flow_layer(FusedLayerArgs a) {
  extern __shared__ __align__(16) unsigned short lds[];
  const int H = 64 * NBD, h = 64 * NBH, D = 2 * H;
  const int ldx = (H > h ? H : h) + kPadF, ldh = h + kPadF;
  unsigned short* Xh = lds;                 // conditioning half, later the second hidden layer
  unsigned short* Xl = Xh + kRows * ldx;
  unsigned short* Hh = Xl + kRows * ldx;    // first hidden layer
  unsigned short* Hl = Hh + kRows * ldh;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t r0 = (int64_t)blockIdx.x * kRows;
  float* zrow = a.z + r0 * D;

  // stage and split the conditioning half (kRows x H, FP32 -> bf16 hi/lo)
  for (int e = tid; e < kRows * H; e += 256) {
    const int r = e / H, c = e - r * H;
    const float x = zrow[(int64_t)r * D + a.a_off + c];
    const unsigned short hb = f2bf_f(x);
    Xh[r * ldx + c] = hb;
    Xl[r * ldx + c] = f2bf_f(x - bf2f_f(hb));
  }
  __syncthreads();
  {
    f32x4f acc[2][NBH];
    mlp_product<NBH>(acc, Xh, Xl, ldx, a.W1h, a.W1l, H, wave * 16 * NBH, lane);
    store_hidden<NBH>(acc, a.b1, Hh, Hl, ldh, wave * 16 * NBH, lane);
  }
  __syncthreads();
  {
    f32x4f acc[2][NBH];
    mlp_product<NBH>(acc, Hh, Hl, ldh, a.W2h, a.W2l, h, wave * 16 * NBH, lane);
    __syncthreads();  // every wave has finished reading the conditioning half from X
    store_hidden<NBH>(acc, a.b2, Xh, Xl, ldx, wave * 16 * NBH, lane);
  }
  __syncthreads();
  // last product: raw scales (columns 0..H) and shifts (H..D) for the same elements
  f32x4f as[2][NBD], at[2][NBD];
  const int c0 = wave * 16 * NBD;
  mlp_product<NBD>(as, Xh, Xl, ldx, a.W3h, a.W3l, h, c0, lane);
  mlp_product<NBD>(at, Xh, Xl, ldx, a.W3h, a.W3l, h, H + c0, lane);
  double part[2][4];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int r = 0; r < 4; ++r) part[rb][r] = 0.0;
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < NBD; ++cb) {
      const int i = c0 + 16 * cb + (lane & 15);
      const float bs = a.b3[i], bt = a.b3[H + i];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * rb + 4 * (lane >> 4) + r;
        const float s = a.smax * tanhf((as[rb][cb][r] + bs) / a.smax);
        const float t = at[rb][cb][r] + bt;
        float* zp = zrow + (int64_t)row * D + a.b_off + i;
        *zp = a.inverse ? (*zp - t) * expf(-s) : fmaf(*zp, expf(s), t);
        part[rb][r] += (double)s;
      }
    }
  // row sums of s: over the 16 lanes that share a row, then over the four waves
  __syncthreads();  // LDS is reused below
  double* red = (double*)lds;  // [4 waves][kRows]
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double v = part[rb][r];
#pragma unroll
      for (int m = 1; m < 16; m <<= 1) v += __shfl_xor(v, m, 64);
      if ((lane & 15) == 0) red[wave * kRows + 16 * rb + 4 * (lane >> 4) + r] = v;
    }
  __syncthreads();
  if (tid < kRows) {
    const double tot = ((red[tid] + red[kRows + tid]) + red[2 * kRows + tid]) + red[3 * kRows + tid];
    a.logdet[r0 + tid] += a.inverse ? -tot : tot;
  }
}

template <int NBH, int NBD>
hipError_t launch_nb(const FusedLayerArgs& a, hipStream_t s) {
  const int H = 64 * NBD, h = 64 * NBH;
  const int ldx = (H > h ? H : h) + kPadF, ldh = h + kPadF;
  const size_t lds = (size_t)2 * kRows * (ldx + ldh) * sizeof(unsigned short);
  static bool attr_set = false;  // more than the default 64 KB of dynamic LDS
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void*)k_flow_layer<NBH, NBD>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL((k_flow_layer<NBH, NBD>), dim3((unsigned)(a.B / kRows)), dim3(256), lds, s, a);
  return hipGetLastError();
}

}  // namespace

bool fused_layer_supported(int64_t B, int D, int hidden) {
  const int H = D / 2;
  return B % kRows == 0 && hidden == 256 && (H == 256 || H == 512);
}

hipError_t launch_flow_layer(const FusedLayerArgs& a, hipStream_t s) {
  if (!fused_layer_supported(a.B, a.D, a.hidden)) return hipErrorInvalidValue;
  if (a.D / 2 == 256) return launch_nb<4, 4>(a, s);
  return launch_nb<4, 8>(a, s);
}

}  // namespace pib
